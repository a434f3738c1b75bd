/*
 * afid_oracle.h -- FP64 CPU restatement of the AFiD-MuRPhFi RK3 substep.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py (its cpu_baseline
 * leg) may load it.  Nothing in the product links it or falls back to it.
 *
 * Each routine restates the arithmetic of the reference Fortran routine cited
 * beside it (paths relative to the reference repository root).  Arrays are
 * stored x-fastest, without halos: periodic y/z halos of the reference
 * (src/flow_solver/2decomp/halo_common.F90:66-71, a memcpy at one rank) become
 * index wraps.  Base arrays hold nx = nxm+1 x-entries (index 1..nx),
 * refined arrays nxr = nxmr+1.
 */
#ifndef AFID_ORACLE_H
#define AFID_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors the bou.in value lines read by src/flow_solver/ReadInputFile.F90:20-76. */
typedef struct orc_config {
    int32_t nxm, nym, nzm, nsst;
    int32_t multires, nxmr, nymr, nzmr;
    int32_t salinity, phasefield;
    double alx3, ylen, zlen;
    int32_t istr3;  double str3; int32_t istr3r;
    double rayt, prat, rays, pras; int32_t ffscales;
    int32_t inslws, inslwn, tfixs, tfixn, sfixs, sfixn;
    int32_t active_t, active_s, gaxis;
    double xplusu, xminusu, dpdy, dpdz;
    /* phase field: pf_A, pf_S, pf_Tm of the bou.in PF line; interface width
       pf_eps (0 selects the reference's 1/nxmr, ReadInputFile.F90:126) */
    double pf_a, pf_s, pf_tm, pf_eps;
    /* moisture (src/moisture.F90, humid.in: ReadMoistParameters :64-117):
       moist on/off, alpha_q, beta_q, gamma_q (< 0: derived, :104-110),
       tau_q (times pect when q_ifdiff), qfixN, qfixS, and Sm, the
       humidity-to-thermal diffusivity ratio (pecq = pect / Sm) */
    int32_t moist;
    double q_alpha, q_beta, q_gamma, q_tau;
    int32_t q_ifdiff, q_fixn, q_fixs;
    double q_sm;
} orc_config;

typedef struct orc_state orc_state;

/* field ids (shared numbering with the product's C ABI) */
enum {
    ORC_VX = 0, ORC_VY, ORC_VZ, ORC_TEMP, ORC_PR,
    ORC_RUX, ORC_RUY, ORC_RUZ, ORC_RUTEMP,
    ORC_QCAP, ORC_DPH, ORC_DQ, ORC_HRO, ORC_RHS,
    ORC_SAL, ORC_RUSAL, ORC_HSAL, ORC_RHSR, ORC_VXR, ORC_VYR, ORC_VZR, ORC_SALC,
    ORC_TEMPBP, ORC_TEMPTP, ORC_SALBP, ORC_SALTP,
    ORC_PHI, ORC_RUPHI, ORC_HPHI, ORC_PHIC, ORC_TEMPR,
    /* moisture.F90:10-13 (base grid): humid, hkhumid (RK history), rkhumid
       (this substep's explicit term), qsat, and the wall planes humbp/humtp */
    ORC_HUMID, ORC_RUHUMID, ORC_HHUMID, ORC_QSAT, ORC_HUMBP, ORC_HUMTP,
    ORC_NFIELDS
};

/* 1-D grid arrays exposed for the grid tests */
enum {
    ORC_G_XC = 0, ORC_G_XM, ORC_G_G3RC, ORC_G_G3RM, ORC_G_D3XC, ORC_G_D3XM,
    ORC_G_UDX3C, ORC_G_UDX3M,
    ORC_G_AP3CK, ORC_G_AC3CK, ORC_G_AM3CK,
    ORC_G_AP3SK, ORC_G_AC3SK, ORC_G_AM3SK,
    ORC_G_AP3SSK, ORC_G_AC3SSK, ORC_G_AM3SSK,
    ORC_G_AK1, ORC_G_AK2, ORC_G_AMPHK, ORC_G_ACPHK, ORC_G_APPHK,
    ORC_G_XCR, ORC_G_XMR, ORC_G_UDX3MR, ORC_G_G3RMR,
    ORC_G_AP3SSKR, ORC_G_AC3SSKR, ORC_G_AM3SSKR,
    ORC_G_CXRS, ORC_G_CYRS, ORC_G_CZRS, ORC_G_CXSALC, ORC_G_CYSALC, ORC_G_CZSALC,
    ORC_G_AP3SPKR, ORC_G_AC3SPKR, ORC_G_AM3SPKR, ORC_G_CXPHIC, ORC_G_CYPHIC, ORC_G_CZPHIC,
    ORC_G_NGRID
};

orc_state *orc_create(const orc_config *cfg);
void orc_destroy(orc_state *s);
double *orc_field(orc_state *s, int id, int64_t dims[3]);  /* dims = (nz, ny, nx) */
double *orc_grid(orc_state *s, int id, int64_t *n);
int orc_int_grid(orc_state *s, int id, int64_t *n, int32_t **out); /* 0 irangs 1 jrangs 2 krangs 3 irangb */
void orc_scalars(orc_state *s, double out[8]);  /* ren, pect, pecs, byct, bycs, dx, dy, dz */

void orc_set_substep(orc_state *s, double dt, int ns);
void orc_explicit_terms_vx(orc_state *s);
void orc_explicit_terms_vy(orc_state *s);
void orc_explicit_terms_vz(orc_state *s);
void orc_explicit_terms_temp(orc_state *s);
void orc_implicit_and_update_vx(orc_state *s);
void orc_implicit_and_update_vy(orc_state *s);
void orc_implicit_and_update_vz(orc_state *s);
void orc_implicit_and_update_temp(orc_state *s);
void orc_calc_local_divergence(orc_state *s);
void orc_solve_pressure_correction(orc_state *s);
void orc_correct_velocity(orc_state *s);
void orc_correct_pressure(orc_state *s);
void orc_explicit_salinity(orc_state *s);
void orc_add_sal_buoyancy(orc_state *s);
void orc_implicit_salinity(orc_state *s);
void orc_interp_vel_mgrd(orc_state *s);
void orc_interp_sal_multigrid(orc_state *s);
/* phase field (phasefield.F90) */
void orc_explicit_phase(orc_state *s);
void orc_add_volume_penalty(orc_state *s);
void orc_implicit_phase(orc_state *s);
void orc_add_latent_heat(orc_state *s);
void orc_interp_temp_multigrid(orc_state *s);
void orc_interp_phi_multigrid(orc_state *s);
void orc_pf_scalars(orc_state *s, double out[4]);   /* pf_D, pf_eps, pf_eta, pf_B */
/* moisture (moisture.F90) */
void orc_set_humidity_bcs(orc_state *s);        /* SetHumidityBCs :129-148 */
void orc_explicit_humidity(orc_state *s);       /* ExplicitHumidity :220-292 */
void orc_add_condensation(orc_state *s);        /* AddCondensation :296-318 */
void orc_implicit_humidity(orc_state *s);       /* ImplicitHumidity + SolveImpEqnUpdate_Humidity :336-418 */
void orc_update_saturation(orc_state *s);       /* UpdateSaturation :321-333 */
void orc_moist_scalars(orc_state *s, double out[5]);   /* alpha_q, beta_q, gamma_q, tau_q, pecq */
void orc_time_marcher(orc_state *s, double dt);

void orc_calc_max_cfl(orc_state *s, double *cflm, double *cflmr);
void orc_check_divergence(orc_state *s, double *qmax, double *qmaxr);
void orc_global_quantities(orc_state *s, double out[4]); /* Nu_vol, KE, vrms, Tmean */
/* CalcMeanProfiles: 15 profiles of nxm values (Tbar, vybar, vzbar, vxT, vyT, vzT, Trms,
 * vxrms, vyrms, vzrms, vxvy, vxvz, vyvz, chiT, epsilon) */
void orc_mean_profiles(orc_state *s, double *out);
/* CalcSalStats: 6 refined profiles of nxmr values (Sbar, Srms, vxS, vyS, vzS, chiS) */
void orc_sal_stats(orc_state *s, double *out);
/* CalcPhiStats: phibar, phirms of nxmr values each */
void orc_phi_stats(orc_state *s, double *out);
/* CalcMoistStats: qbar, qsbar, qrms, qrel, vxq, vyq, vzq of nxm values each */
void orc_moist_stats(orc_state *s, double *out);

/* Restart onto a new grid (ReadFlowInterp.F90:90-141): the old run's grid
 * as continua_master.h5 describes it (nx, ny, nz, istr3, str3 and, with
 * multires, nxr, nyr, nzr, istr3r), and the interpolation of its fields,
 * (n?o-1, ..., nxo) x fastest as the checkpoints hold them. */
typedef struct orc_old_grid {
    int32_t nxo, nyo, nzo, istr3o;
    double str3o;
    int32_t nxro, nyro, nzro, istr3ro;
} orc_old_grid;
int orc_interp_input_vel(orc_state *s, const orc_old_grid *og, const double *vxo, const double *vyo,
                         const double *vzo, const double *tempo);   /* InterpInputVel.F90 */
int orc_interp_input_sal(orc_state *s, const orc_old_grid *og, const double *salo);   /* InterpInputSal.F90 */
int orc_interp_input_phi(orc_state *s, const orc_old_grid *og, const double *phio);   /* InterpInputPhi.F90 */
int orc_interp_input_humid(orc_state *s, const orc_old_grid *og, const double *qo);   /* InterpInputHum, moisture.F90:539-582 */

/* Immersed boundary (src/ibm/ibm_param.F90): the mask of the solve of field
 * (ORC_VX, ORC_VY, ORC_VZ, ORC_TEMP on (nzm, nym, nx); ORC_SAL on (nzmr,
 * nymr, nxr)), values 2 liquid, 0 solid, 1 / -1 upper / lower boundary
 * point, and the boundary distances in the loop order (ic, jc, kc) (none for
 * S).  The implicit updates then solve SolveImpEqnUpdate_*_ibm.  Returns 0,
 * or nonzero for a bad field or too short a list. */
int orc_set_ibm(orc_state *s, int field, const int32_t *mask, const double *dist, int64_t ndist);

/* free-standing numerics used by the pins */
void orc_fft(double *re, double *im, int n, int sign);      /* in place, unnormalised */
void orc_dft_naive(const double *re, const double *im, double *ore, double *oim, int n, int sign);
void orc_tridiag_solve(int n, const double *a, const double *b, const double *c, double *d);
void orc_set_threads(int n);
/* Planted mistake n (0: none); the pin tests show each one is caught. */
void orc_set_plant(int n);

#ifdef __cplusplus
}
#endif
#endif
