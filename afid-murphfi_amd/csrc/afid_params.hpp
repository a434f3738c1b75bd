// afid_params.hpp -- the kernel parameter block.
//
// One plain struct, passed by value to every kernel: grid sizes, the RK3
// substep scalars and device pointers.  Device layout of a base field is
// [nz][ny][ldx] doubles with x fastest (x index k = kc-1 for the reference's
// kc), y/z periodic without halo cells; a refined field is [nzr][nyr][ldxr].
// The wall-normal stride ldx is nx rounded up to 8 (64-byte rows).
#pragma once
#include <cstddef>
#include <cstdint>

namespace afid {

enum MatKind { MAT_X = 0, MAT_YZ = 1, MAT_T = 2, MAT_S = 3, MAT_P = 4, MAT_Q = 5, MAT_NKIND = 6 };

// LU factors of one constant wall-normal tridiagonal system, in the form the
// wave solver consumes (row r, 0-based, padded with zeros to 64*chunk rows):
//   forward   y_r = s_r d_r - l_r y_{r-1}
//   backward  x_r = g_r y_r - e_r x_{r+1}
// a, c: the normalised off-diagonals of each row (row r: a_r x_{r-1} + x_r +
// c_r x_{r+1} = s_r d_r), for solves whose rows a mask may replace.
// Folded form of the same factors (the pivot g moved into the forward sweep,
// the row scale s into the right-hand side), two tables instead of four:
//   forward   y'_r = sn_r d_r - ln_r y'_{r-1}   (sn = g s, ln_r = g_r l_r / g_{r-1})
//   backward  x_r  = y'_r - e_r x_{r+1}
struct TriFactors {
    const double *l, *e, *g, *s, *a, *c;
    const double *ln, *sn;
    int n, chunk;
};

struct KP {
    // base grid (nx = nxm+1 faces; ny = nym, nz = nzm periodic cells)
    int nxm, nx, ny, nz, ldx;
    long long plane;            // ny*ldx
    int nymh;                   // ny/2+1 spectral modes in y
    int ldt;                    // x stride of the gradient scratch (x index -1..nx+1)
    // refined grid
    int nxmr, nxr, nyr, nzr, ldxr;
    long long planer;
    // substep scalars (TimeMarcher.F90:28,38-40)
    double dt, al, ga, ro, beta;
    // physics (ReadInputFile.F90:101-122)
    double ren, pect, pecs, byct, bycs, Gy;
    double dx, dy, dz, dxq, dyq, dzq;
    double dxr, dyr, dzr, dxqr, dyqr, dzqr;
    int inslws, inslwn, tfixs, tfixn, sfixs, sfixn, active_t, active_s, gaxis;
    double xplusu, xminusu;
    // base fields
    double *vx, *vy, *vz, *temp, *pr;
    double *rux, *ruy, *ruz, *rut;
    double *qcap, *dph, *dq, *hro;
    double *rhsx, *rhsy, *rhsz, *rhst;
    double *salc, *tpdv;
    const double *tempbp, *temptp;      // [nz][ny]
    // refined fields
    double *sal, *rusal, *hsal, *rhsr, *vxr, *vyr, *vzr;
    const double *salbp, *saltp;        // [nzr][nyr]
    // mixed form: FP32 storage of the refined fields and of the base
    // explicit-term history (null in the FP64 form)
    float *salf, *rusalf, *vxrf, *vyrf, *vzrf;
    float *ruxf, *ruyf, *ruzf, *rutf, *salcf;
    // base metrics (0-based by kc-1)
    const double *udx3m, *udx3c, *rudx3m;
    const double *ap3ck, *ac3ck, *am3ck, *ap3sk, *ac3sk, *am3sk, *ap3ssk, *ac3ssk, *am3ssk;
    const double *amphk, *acphk, *apphk, *g3rc;
    // refined metrics
    const double *udx3mr, *ap3sskr, *ac3sskr, *am3sskr, *lxa;
    // Hermite tables coarse -> refined (per refined index): source + 4 weights
    const int *xs_src, *ys_src, *zs_src;
    const double *xs_w, *ys_w, *zs_w;
    // FP32 copies of the Hermite weights and of lxa (mixed form)
    const float *xs_wf, *ys_wf, *zs_wf, *lxaf;
    // phase field (phasefield.F90): refined phi, its RK arrays and the
    // refined temperature; coarse phic and the latent-heat source qlat
    double *phi, *ruphi, *hphi, *tempr, *phic, *qlat;
    float *phif, *ruphif, *temprf, *rhsrf;      // FP32 storage of the refined phase-field arrays (mixed form)
    float *dphf;                                // FP32 pressure correction of the mixed form's Poisson sweeps
    double pf_d, pf_b, pf_a, pf_tm, pf_s, heta;  // pf_D, pf_D/eps^2, A, T_m, S_t, 0.5/eta
    int pf;
    const double *ap3spkr, *ac3spkr, *am3spkr;   // zero-gradient x operator of phi (refined)
    // moisture (moisture.F90): humidity, its RK history (hkhumid), this
    // substep's explicit term (rkhumid), saturation, wall values; alpha_q,
    // beta_q, gamma_q/tau_q, 1/tau_q, kappa_q = 1/pecq, and the wall flags
    double *humid, *ruhumid, *hhumid, *qsat;
    const double *humbp, *humtp, *xm;
    double q_alpha, q_beta, q_gtau, q_itau, kappa_q, pecq;
    int q_fixn, q_fixs, moist;
    // mixed form, gAxis = 1: per-row FP64 reference buoyancy subtracted from
    // the FP32 vx explicit term and added back in the RHS (null otherwise)
    const double *bref;
    // reductions
    unsigned long long *redmax;         // [4] max slots (bit patterns of non-negative doubles)
    double *redsum;                     // per-block partial sums
};

}  // namespace afid
