/*
 * afid_ensemble.h -- C ABI of the batched ensemble: many independent 2-D
 * runs (nym = 1, the x-z plane) of uneven size, Rayleigh and Prandtl number,
 * advanced together on one GPU.
 *
 * The reference runs one simulation per process: main.F90:281-422 picks dt
 * from the CFL number (CalcMaxCFL, main.F90:285-302), calls TimeMarcher
 * (TimeMarcher.F90:13-165), advances the time and checks for the exit codes
 * 165/166/169/333/555 (main.F90:355-385).  An ensemble holds that driver
 * state per run on the device and advances every run by one RK3 substep per
 * kernel launch; afid_ens_step replays a captured graph of such launches.
 *
 * A run is described by the afid_config of its bou.in (nym must be 1; no
 * multires, salinity or phase field; every run of an ensemble has the same
 * precision: AFID_FP64, or AFID_MIXED with FP32 explicit terms and Poisson
 * sweeps, FP64 state and solves, and further sweeps while the FP64 max|div u|
 * exceeds div_tol, which runs only the phased launch shape) and an
 * afid_run_control with the
 * time-stepping line of its bou.in (IDTV, DT, RESID, CFLMAX, DTMIN, DTMAX)
 * and its NTST / TMAX.  v is identically zero in such a run (every y
 * difference vanishes when nym = 1), so the ensemble does not advance it.
 *
 * Conventions as in afid_gpu.h: 0 on success, AFID_E_* otherwise with
 * afid_ens_last_error(); one host thread; host arrays column-major with
 * lower bounds and extents (e.g. temp(1:nx, 1:1, 1:nzm)).
 */
#ifndef AFID_ENSEMBLE_H
#define AFID_ENSEMBLE_H
#include <stdint.h>
#include "afid_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Time-stepping line and run limits of one run's bou.in (ReadInputFile.F90:
 * 20-76; main.F90:287-302, 355-385). */
typedef struct afid_run_control {
    int32_t variable_dt;  /* IDTV: 1 = dt = CFLMAX/CFL capped at DTMAX from the third step on */
    double dt;            /* DT: the first steps' (or the fixed) time step */
    double resid;         /* RESID: exit 169 when max|div u| exceeds it */
    double limit_cfl;     /* CFLMAX */
    double dtmin, dtmax;  /* DTMIN (exit 166 below it), DTMAX */
    int32_t ntst;         /* NTST: the last step index (steps 0..NTST), exit 555 after it */
    double tmax;          /* TMAX: exit 333 once the time exceeds it */
} afid_run_control;

/* Launch shape of a substep, chosen by afid_ens_autotune or set.
 * phased = 0: one persistent launch per substep; each workgroup owns whole
 *             runs (threads, workgroups, lines apply);
 * phased = 1: one launch per phase (factors, columns, forward transform,
 *             modes, inverse transform, projection, finish), each over the
 *             work items of every run (lines applies; 256 threads). */
typedef struct afid_ens_tuning {
    int32_t threads;      /* workgroup size of the persistent launch: 256, 512 or 1024 */
    int32_t workgroups;   /* persistent grid: runs are dealt largest-first to this many workgroups */
    int32_t lines;        /* transform lines per LDS batch (upper bound) */
    int32_t phased;       /* 0 persistent launch per substep, 1 one launch per phase */
    double ms_per_step;   /* measured time of one RK3 step of the whole ensemble (autotune) */
} afid_ens_tuning;

/* Per-run state read back by afid_ens_run_state. */
typedef struct afid_run_state {
    double time, dt;      /* time reached and dt of the next step */
    double cfl;           /* CFL number of the current state (CalcMaxCFL, without dt) */
    double max_div;       /* max|div u| of the current state (CheckDivergence) */
    int32_t ntime;        /* index of the next step */
    int32_t status;       /* 0 running, else the exit code the reference would stop with */
    int32_t steps;        /* RK3 steps done */
} afid_run_state;

typedef struct afid_ens afid_ens;

/* Create an ensemble of nruns runs.  Each run gets its grid, metrics and
 * Poisson factors as afid_gpu_init builds them; dt starts at ctl[r].dt. */
int afid_ens_init(const afid_config *cfgs, const afid_run_control *ctl, int32_t nruns, afid_ens **out);
int afid_ens_finalize(afid_ens *e);
const char *afid_ens_last_error(const afid_ens *e);
int64_t afid_ens_device_bytes(const afid_ens *e);

/* Fields of run r (AFID_VX, AFID_VZ, AFID_TEMP, AFID_PR, AFID_RUX, AFID_RUZ,
 * AFID_RUTEMP, AFID_DPH, AFID_TEMPBP, AFID_TEMPTP; AFID_VY downloads zeros
 * and accepts only zeros). */
int afid_ens_upload(afid_ens *e, int32_t run, int32_t field, const double *host, const int64_t lb[3], const int64_t ext[3]);
int afid_ens_download(afid_ens *e, int32_t run, int32_t field, double *host, const int64_t lb[3], const int64_t ext[3]);

/* CreateInitialConditions (RBC branch, u = 0, T = 1 - x plus eps noise
 * seeded with seed0 + r) for every run, and the driver state of step 0. */
int afid_ens_create_initial_conditions(afid_ens *e, uint64_t seed0, double eps);

/* Restart the driver state of every run at step 0 with its DT (after
 * uploads); time 0, status 0. */
int afid_ens_reset(afid_ens *e);

/* Advance every running run by nsteps RK3 steps (a run stops at its exit
 * code).  graph != 0 replays a captured hipGraph of the substep launches;
 * graph == 0 launches them one by one.  Asynchronous. */
int afid_ens_step(afid_ens *e, int32_t nsteps, int32_t graph);

/* afid_ens_step bracketed by HIP events on the ensemble's stream; waits for
 * the steps and returns their elapsed time in *ms. */
int afid_ens_step_timed(afid_ens *e, int32_t nsteps, int32_t graph, double *ms);

int afid_ens_run_state(afid_ens *e, int32_t run, afid_run_state *out);

/* GlobalQuantities of run r (afid_global_quantities layout: Nu_vol, KE, vrms, Tmean). */
int afid_ens_global_quantities(afid_ens *e, int32_t run, double out[4]);

/* Time candidate launch shapes on the current ensemble (steps RK3 steps
 * each; the state is restored afterwards) and keep the fastest.  The
 * candidates and their times go to log/nlog (may be null). */
int afid_ens_autotune(afid_ens *e, int32_t steps, afid_ens_tuning *best, afid_ens_tuning *log, int32_t maxlog,
                      int32_t *nlog);
int afid_ens_set_tuning(afid_ens *e, const afid_ens_tuning *t);
int afid_ens_get_tuning(afid_ens *e, afid_ens_tuning *t);

int afid_ens_synchronize(afid_ens *e);

#ifdef __cplusplus
}
#endif
#endif
