/*
 * afid_gpu.h -- C ABI of the MI355X implementation of the AFiD-MuRPhFi RK3
 * substep (TimeMarcher and its callees).
 *
 * The reference has no plugin or FFI layer: its boundary is the Fortran call
 * tree under TimeMarcher (src/flow_solver/TimeMarcher.F90:13-165), made of
 * argument-less subroutines acting on module arrays.  Each entry point below
 * replaces one of those subroutines (cited beside it); the Fortran host binds
 * them with iso_c_binding (afid-murphfi_amd/fortran/afid_gpu_mod.F90).
 *
 * Conventions
 *  - every function returns 0 on success, nonzero on error (AFID_E_*);
 *    afid_last_error() gives a message;
 *  - calls are made from one host thread; work is enqueued on the context's
 *    HIP stream and runs asynchronously; functions that return scalars and
 *    afid_download synchronise;
 *  - host arrays are Fortran column-major (x fastest) with explicit lower
 *    bounds and extents per dimension, exactly as the reference allocates them
 *    (AuxiliaryRoutines.F90:79-99): e.g. vx(1:nx, -1:nym+2, -1:nzm+2) is
 *    lbound = {1,-1,-1}, extent = {nx, nym+4, nzm+4}.  Compact arrays use
 *    lbound = {1,1,1}.  Only the interior is read on upload; download writes
 *    the interior and fills any periodic y/z halo the host array has.
 */
#ifndef AFID_GPU_H
#define AFID_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AFID_OK 0
#define AFID_E_ARG 1      /* invalid argument or unsupported configuration */
#define AFID_E_HIP 2      /* HIP runtime error */
#define AFID_E_MEM 3      /* device allocation failed */
#define AFID_E_STATE 4    /* call out of order (e.g. no substep set) */

/* Arithmetic form of the context. */
#define AFID_FP64 0       /* every field and every operation in FP64 */
#define AFID_MIXED 1      /* FP64 state (u, T, p) and x-tridiagonals; FP32 storage and
                             arithmetic of the explicit terms, the refined scalar
                             and velocities and the Hermite transfers; FP32 Poisson
                             transforms inside FP64-residual iterative refinement */

/* The bou.in value lines (ReadInputFile.F90:20-76) plus the arithmetic form. */
typedef struct afid_config {
    int32_t nxm, nym, nzm, nsst;
    int32_t multires, nxmr, nymr, nzmr;
    int32_t salinity, phasefield;
    double alx3, ylen, zlen;
    int32_t istr3;  double str3; int32_t istr3r;   /* istr3 8 = x_k = L/2 (1-cos(pi k/nxm)) */
    double rayt, prat, rays, pras; int32_t ffscales;
    int32_t inslws, inslwn, tfixs, tfixn, sfixs, sfixn;
    int32_t active_t, active_s, gaxis;
    double xplusu, xminusu, dpdy, dpdz;
    int32_t precision;                             /* AFID_FP64 or AFID_MIXED */
    double div_tol;        /* AFID_MIXED: refinement stops once max|div u| <= div_tol (0: 1e-10) */
    int32_t max_sweeps;    /* AFID_MIXED: Poisson sweeps per substep, 1..4 (0: 3) */
    /* phase field: pf_A, pf_S, pf_Tm of the bou.in PF line (pf_D follows,
       ReadInputFile.F90:125); interface width pf_eps (0: 1/nxmr, :126);
       radius pf_r0 of the solid sphere of the seeded initial state (0: 0.1,
       phasefield.F90:109-110) */
    double pf_a, pf_s, pf_tm, pf_eps, pf_r0;
    /* moisture (src/moisture.F90, switched on when humid.in exists,
       main.F90:38-39): the humid.in values of ReadMoistParameters
       (moisture.F90:64-117): alpha_q, beta_q, gamma_q (< 0: the value
       that makes the moist-static jump 1, :104-110), tau_q (times pect when
       q_ifdiff), qfixN (2: unsaturated top), qfixS, and Sm, the
       humidity-to-thermal diffusivity ratio (pecq = pect / Sm) */
    int32_t moist;
    double q_alpha, q_beta, q_gamma, q_tau;
    int32_t q_ifdiff, q_fixn, q_fixs;
    double q_sm;
} afid_config;

typedef struct afid_ctx afid_ctx;

/* Field identifiers (module arrays of param.F90:100-126, salinity.F90:14-35). */
enum afid_field {
    AFID_VX = 0, AFID_VY, AFID_VZ, AFID_TEMP, AFID_PR,
    AFID_RUX, AFID_RUY, AFID_RUZ, AFID_RUTEMP,
    AFID_QCAP, AFID_DPH, AFID_DQ, AFID_HRO, AFID_RHS,
    AFID_SAL, AFID_RUSAL, AFID_HSAL, AFID_RHSR, AFID_VXR, AFID_VYR, AFID_VZR, AFID_SALC,
    AFID_TEMPBP, AFID_TEMPTP, AFID_SALBP, AFID_SALTP,
    AFID_PHI, AFID_RUPHI, AFID_HPHI, AFID_PHIC, AFID_TEMPR,   /* phasefield.F90:12-16 */
    /* moisture.F90:10-13, base grid: humid, hkhumid (RK history), rkhumid
       (explicit term of the substep), qsat, and the wall planes humbp/humtp */
    AFID_HUMID, AFID_RUHUMID, AFID_HHUMID, AFID_QSAT, AFID_HUMBP, AFID_HUMTP,
    AFID_NFIELDS
};

/* Set-up: ReadInputFile derived constants (ReadInputFile.F90:101-122),
 * InitTimeMarchScheme (InitTimeMarchScheme.F90:11-52), CreateGrid
 * (CreateGrid.F90:11-179), InitPressureSolver (pressure.F90:61-136),
 * CreateMgrdGrid (CreateMgrdGrid.F90:11-166), CreateMgrdStencil
 * (CreateMgrdStencil.F90:1-81) and the allocation of InitVariables /
 * InitMgrdVariables / InitSalVariables.  All device memory is allocated here. */
int afid_gpu_init(const afid_config *cfg, afid_ctx **out);
/* QuitRoutine.F90:46-59 deallocation. */
int afid_gpu_finalize(afid_ctx *ctx);
const char *afid_last_error(const afid_ctx *ctx);
/* Device bytes held by the context. */
int64_t afid_device_bytes(const afid_ctx *ctx);

/* Host <-> device copies of a module array (HdfReadContinua.F90:11-107 /
 * HdfWriteRealHalo3D, HdfRoutines.F90:428-520, read and write these arrays). */
int afid_upload(afid_ctx *ctx, int field, const double *host, const int64_t lbound[3], const int64_t extent[3]);
int afid_download(afid_ctx *ctx, int field, double *host, const int64_t lbound[3], const int64_t extent[3]);

/* TimeMarcher.F90:28,38-40: beta = dt/ren/2 and (al, ga, ro) of substep ns (1-based). */
int afid_set_substep(afid_ctx *ctx, double dt, int ns);

/* Explicit terms (stage 1 of a substep). */
int afid_explicit_terms_vx(afid_ctx *ctx);      /* ExplicitTermsVX.F90:11-131   -> qcap */
int afid_explicit_terms_vy(afid_ctx *ctx);      /* ExplicitTermsVY.F90:11-126   -> dph  */
int afid_explicit_terms_vz(afid_ctx *ctx);      /* ExplicitTermsVZ.F90:11-121   -> dq   */
int afid_explicit_terms_temp(afid_ctx *ctx);    /* ExplicitTermsTemp.F90:11-85  -> hro  */
int afid_explicit_salinity(afid_ctx *ctx);      /* salinity.F90:268-366         -> hsal */
int afid_add_sal_buoyancy(afid_ctx *ctx);       /* salinity.F90:495-506, TimeMarcher.F90:54-62 */

/* Implicit update (RHS assembly + wall-normal tridiagonal solve + update). */
int afid_implicit_and_update_vx(afid_ctx *ctx);   /* ImplicitAndUpdateVX.F90:14-89 + SolveImpEqnUpdate_X.F90:12-69 */
int afid_implicit_and_update_vy(afid_ctx *ctx);   /* ImplicitAndUpdateVY.F90:14-95 + SolveImpEqnUpdate_YZ.F90:12-60 */
int afid_implicit_and_update_vz(afid_ctx *ctx);   /* ImplicitAndUpdateVZ.F90:14-80 + SolveImpEqnUpdate_YZ.F90:12-60 */
int afid_implicit_and_update_temp(afid_ctx *ctx); /* ImplicitAndUpdateTemp.F90:13-77 + SolveImpEqnUpdate_Temp.F90:11-63 */
int afid_implicit_salinity(afid_ctx *ctx);        /* salinity.F90:369-454 */

/* Pressure projection. */
int afid_calc_local_divergence(afid_ctx *ctx);     /* pressure.F90:141-162 */
int afid_solve_pressure_correction(afid_ctx *ctx); /* pressure.F90:165-236 */
int afid_correct_velocity(afid_ctx *ctx);          /* CorrectVelocity.F90:24 -> RemoveDivergence, pressure.F90:240-267 */
int afid_correct_pressure(afid_ctx *ctx);          /* pressure.F90:270-301 */

/* Phase field (phasefield.F90).  The velocity solves of
 * afid_implicit_and_update_vx/vy/vz switch to the masked forms
 * SolveImpEqnUpdate_X_pf / _YZ_pf (:821-930) when the phase field is on. */
int afid_explicit_phase(afid_ctx *ctx);           /* phasefield.F90:345-383 -> hphi */
int afid_add_volume_penalty(afid_ctx *ctx);       /* phasefield.F90:525-566 */
int afid_implicit_phase(afid_ctx *ctx);           /* phasefield.F90:386-466 (rhsr keeps d phi) */
int afid_add_latent_heat(afid_ctx *ctx);          /* phasefield.F90:651-725 */
int afid_interp_temp_multigrid(afid_ctx *ctx);    /* phasefield.F90:498-520 */
int afid_interp_phi_multigrid(afid_ctx *ctx);     /* phasefield.F90:470-494 */

/* Immersed boundary (src/ibm/ibm_param.F90): the mask of one wall-normal
 * solve, field = AFID_VX (ibmaskx, distx), AFID_VY (ibmasky, disty), AFID_VZ
 * (ibmaskz, distz), AFID_TEMP (ibmaskt, distbt) on (nx, nym, nzm), or
 * AFID_SAL (ibmaskr) on (nxr, nymr, nzmr), x fastest: 2 liquid, 0 solid,
 * 1 / -1 upper / lower boundary point.  dist: the boundary distances in the
 * reference's loop order (z, y, x over rows 1..nxm), ndist of them (none for
 * S, whose solve takes 1).  The implicit updates of that field then solve
 * SolveImpEqnUpdate_X_ibm / _YZ_ibm / _Temp_ibm / _Sal_ibm (src/ibm/): solid
 * rows force the increment to -q (to Tb - T, Tb = 1, for T), boundary rows
 * tie the point to its neighbour.  T keeps the plain solve when the phase
 * field is on (ImplicitAndUpdateTemp.F90:69-70).  mask = NULL switches the
 * solve back.  FP64 call tree only: afid_time_marcher returns AFID_E_STATE
 * while a mask is set. */
int afid_set_ibm(afid_ctx *ctx, int field, const int32_t *mask, const double *dist, int64_t ndist);

/* Moisture (src/moisture.F90), for a context with moist = 1.  The
 * per-routine calls follow the call tree (TimeMarcher.F90:79-81, :97, :154);
 * afid_time_marcher runs them fused.  FP64 arithmetic in both forms. */
int afid_set_humidity_bcs(afid_ctx *ctx);     /* SetHumidityBCs, moisture.F90:129-148 (from tempbp / temptp) */
int afid_explicit_humidity(afid_ctx *ctx);    /* ExplicitHumidity, :220-292 (into hhumid = rkhumid) */
int afid_add_condensation(afid_ctx *ctx);     /* AddCondensation, :296-318 (into hro) */
int afid_implicit_humidity(afid_ctx *ctx);    /* ImplicitHumidity + SolveImpEqnUpdate_Humidity, :336-418 */
int afid_update_saturation(afid_ctx *ctx);    /* UpdateSaturation, :321-333 */

/* Multi-resolution transfers. */
int afid_interp_vel_mgrd(afid_ctx *ctx);       /* InterpVelMgrd.F90:1-243 */
int afid_interp_sal_multigrid(afid_ctx *ctx);  /* salinity.F90:458-491 (fast path, HermiteInterpolations.F90:256-312) */

/* One full RK3 step, the whole call tree of TimeMarcher.F90:13-165.  The
 * only stepping call of an AFID_MIXED context (its per-routine calls return
 * AFID_E_STATE): the mixed form replaces SolvePressureCorrection's single
 * FP64 solve (pressure.F90:165-197) by FP32 sweeps inside FP64-residual
 * iterative refinement. */
int afid_time_marcher(afid_ctx *ctx, double dt);

/* No counterpart in the reference: nsteps calls of afid_time_marcher(ctx, dt)
 * replayed from a HIP graph of two captured steps (the small grids of a
 * 64^3 run are bound by kernel launches).  Bit-identical to the eager
 * steps.  The graph is captured on the first call, and again when dt changes
 * or the fields were reallocated; an odd last step runs eagerly. */
int afid_time_marcher_graph(afid_ctx *ctx, double dt, int nsteps);

/* AFID_MIXED: max|div u| (CheckDivergence units) measured before each
 * refinement sweep 2..max_sweeps of the last substep; 0 where the sweep was
 * not needed.  count = max_sweeps - 1 (0 for an AFID_FP64 context).  No
 * counterpart in the reference, which has no mixed form. */
int afid_refinement_residuals(afid_ctx *ctx, double *out, int max, int *count);

/* Reductions read back to the host. */
int afid_calc_max_cfl(afid_ctx *ctx, double *cflm, double *cflmr);      /* CalcMaxCFL.F90:11-72 */
int afid_check_divergence(afid_ctx *ctx, double *qmax, double *qmaxr);  /* CheckDivergence.F90:11-87 */
/* GlobalQuantities.F90:11-102: out = {Nu_vol, mean kinetic energy, rms velocity, mean T}. */
int afid_global_quantities(afid_ctx *ctx, double out[4]);
/* CalcMeanProfiles.F90:11-229: the 15 wall-parallel mean profiles of the base
 * grid, out[q * nxm + k] (k = 0..nxm-1) with q in the order the reference
 * writes them to means.h5 (:230-283): Tbar, vybar, vzbar, vxT, vyT, vzT, Trms,
 * vxrms, vyrms, vzrms, vxvy, vxvz, vyvz, chiT, epsilon.  out holds 15*nxm
 * doubles; the host writes the file (HdfSerialWriteReal1D).  FP64 in both
 * forms.  Needs nxm >= 2.  Waits for the result. */
int afid_calc_mean_profiles(afid_ctx *ctx, double *out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight diagnostic loops, hosts that
 * forbid hipMalloc): afid_calc_mean_profiles with the partial sums in a
 * static device array of AFID_NOALLOC_SCRATCH_DOUBLES doubles (16 MiB, held
 * on every GPU the library runs on until the process ends) instead of the
 * buffers afid_calc_mean_profiles allocates on its first call.  out_dev is
 * DEVICE memory of 15*nxm doubles, written in the same order; the call only
 * enqueues work on the context's stream (no allocation, no copy, no wait:
 * afid_synchronize before reading out_dev).  Grids whose partials exceed the
 * array (15 * nxm * ceil(ny/4) doubles) get AFID_E_ARG.  All contexts on one
 * GPU share the array: calls that overlap in time on one GPU race. */
#define AFID_NOALLOC_SCRATCH_DOUBLES (2 * 1024 * 1024)
int afid_calc_mean_profiles_noalloc(afid_ctx *ctx, double *out_dev);
/* CalcSalStats + CalcDissipationSal (salinity.F90:510-690, non-IBM): the 6
 * refined-grid profiles Sbar, Srms, vxS, vyS, vzS, chiS, out[q * nxmr + k],
 * the order of means.h5 (:583-597); out holds 6*nxmr doubles.  Sums in FP64
 * in both forms (the mixed form reads its FP32 refined fields).  Needs the
 * salinity field and nxmr >= 2.  Waits for the result. */
int afid_calc_sal_stats(afid_ctx *ctx, double *out);
/* CalcPhiStats (phasefield.F90:750-798): phibar and phirms of the refined
 * phase field, out[q * nxmr + k], q = 0 phibar, 1 phirms (2*nxmr doubles).
 * FP64 sums in both forms.  Needs the phase field.  Waits for the result. */
int afid_calc_phi_stats(afid_ctx *ctx, double *out);
/* CalcMoistStats (moisture.F90:422-505): qbar, qsbar, qrms, qrel, vxq, vyq,
 * vzq, out[q * nxm + k] (7*nxm doubles, the order of means.h5).  FP64 sums.
 * Needs moisture.  Waits for the result. */
int afid_calc_moist_stats(afid_ctx *ctx, double *out);

/* Restart onto a new grid (ReadFlowInterp.F90:90-141): the grid of the old run,
 * as its continua_master.h5 records it (nx, ny, nz, istr3, str3, ylen, zlen
 * and, with multires, nxr, nyr, nzr, istr3r; sizes count the faces). */
typedef struct afid_old_grid {
    int32_t nxo, nyo, nzo, istr3o;
    double str3o, ylen, zlen;
    int32_t nxro, nyro, nzro, istr3ro;
} afid_old_grid;
/* CreateOldGrid + CreateNewInputStencil + InterpInputVel
 * (src/multires/IC_interpolation/): vx, vy, vz and temp of the old run, host
 * arrays of the checkpoint's dims (nxo, nyo-1, nzo-1), x fastest, moved to
 * this context's base grid by the 4-point Hermite transfer: T directly (with
 * the ghosts of its wall values), the velocity through its interpolated
 * gradients (vx integrated from the wall, vy and vz from first-guess planes
 * at the first y / z face), so a divergence-free old field stays
 * divergence-free off the periodic seams.  pr and the RK history restart
 * from zero, as in the reference.  The domain lengths must match
 * (ReadFlowInterp.F90:76-82).  Waits for the result; the host then runs the
 * start-up transfers (InterpVelMgrd, InterpSalMultigrid, ...). */
int afid_interp_input_vel(afid_ctx *ctx, const afid_old_grid *og, const double *vx, const double *vy,
                          const double *vz, const double *temp);
/* CreateNewSalStencil + InterpInputSal: the old refined S, dims (nxro,
 * nyro-1, nzro-1), with its Dirichlet (sign of Ra_S; 1 / 0 with the phase
 * field) or zero-gradient wall ghosts, onto the refined grid. */
int afid_interp_input_sal(afid_ctx *ctx, const afid_old_grid *og, const double *sal);
/* InterpInputPhi: the old refined phi with zero-gradient ghosts onto the refined grid. */
int afid_interp_input_phi(afid_ctx *ctx, const afid_old_grid *og, const double *phi);
/* InterpInputHum (moisture.F90:539-582): the old humidity, dims (nxo, nyo-1,
 * nzo-1), onto the base grid, with the ghosts of its wall values (run
 * afid_set_humidity_bcs first); the history restarts from zero. */
int afid_interp_input_humid(afid_ctx *ctx, const afid_old_grid *og, const double *humid);

/* Seeded initial state (CreateInitialConditions.F90:150-201, RBC branch, with
 * u = 0): T = tempbp + (temptp - tempbp) x + noise, the wall-clamped noise of
 * amplitude eps being U[0,1) of splitmix64(seed ^ linear index) mapped to
 * [-eps, eps]; with salinity, S likewise on the refined grid from salbp/saltp
 * with seed+1, then salc.  With the phase field, a solid sphere of radius
 * pf_r0 at the centre (set_ice_sphere, phasefield.F90:311-341): phi =
 * (1 - tanh((r - pf_r0)/(2 pf_eps)))/2 on the refined grid, T = 0 in the
 * base cells with r <= pf_r0, then tempr and phic (main.F90:210-212).  The
 * wall planes must be uploaded first.  The reference seeds from the system
 * clock (CreateInitialConditions.F90:24); a counter-based generator makes
 * CPU and GPU states identical. */
int afid_create_initial_conditions(afid_ctx *ctx, uint64_t seed, double eps);

/* Wait for all work enqueued on the context. */
int afid_synchronize(afid_ctx *ctx);

/* Kernel timing of the last afid_time_marcher call, for the roofline: per
 * kernel group, accumulated device milliseconds measured with HIP events on
 * the context's stream.  Enabled by afid_set_timing(ctx, 1). */
int afid_set_timing(afid_ctx *ctx, int on);
int afid_get_timing(afid_ctx *ctx, const char **names, double *ms, int max, int *count);

/* 1-D set-up arrays, for the set-up tests (copied into out, at most n). */
int afid_grid_array(afid_ctx *ctx, const char *name, double *out, int64_t n, int64_t *count);

#ifdef __cplusplus
}
#endif
#endif
