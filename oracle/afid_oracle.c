/*
 * afid_oracle.c -- FP64 CPU restatement of the AFiD-MuRPhFi RK3 substep.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path (see the
 * header).  Every routine cites the reference Fortran it restates; paths are
 * relative to the reference repository root.  Indices in the comments and in
 * the arithmetic follow the reference's 1-based Fortran convention:
 * k = x (wall-normal, contiguous), j = y, i = z.
 *
 * Departures from the reference, all deliberate:
 *  - periodic halos are index wraps (one rank: halo_common.F90:66-71 copies);
 *  - LAPACK dgttrf/dgttrs and zgttrf/zgttrs are replaced by Thomas
 *    elimination without pivoting (the matrices are diagonally dominant);
 *  - FFTW is replaced by a textbook mixed-radix FFT (orc_fft);
 *  - the singular (ky,kz)=(0,0) Poisson mode, which the reference regularises
 *    with a 1e-15 diagonal shift (pressure.F90:217-219) and so fixes the
 *    pressure's additive constant by round-off, is solved with its first
 *    unknown pinned to zero; every other mode keeps the 1e-15 shift.
 */
#include "afid_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXALLOC 256

struct orc_state {
    orc_config c;
    int nx, ny, nz, nxm, nym, nzm;
    int nxr, nyr, nzr, nxmr, nymr, nzmr;
    double ren, pect, pecs, byct, bycs, rrho, pi;
    double dx, dy, dz, dxq, dyq, dzq;
    double dxr, dyr, dzr, dxqr, dyqr, dzqr;
    double gam[3], rom[3], alm[3];
    double dt, al, ga, ro, beta;
    /* 1-D arrays, Fortran index bounds noted */
    double *xc, *xm, *yc, *ym, *zc, *zm;          /* xc(1:nx) xm(0:nx) yc(1:ny) ym(0:ny) zc(1:nz) zm(0:nz) */
    double *g3rc, *g3rm, *d3xc, *d3xm, *udx3c, *udx3m;   /* (1:nx) */
    int *kmv, *kpv, *kmc, *kpc;                    /* (1:nx) */
    double *ap3ck, *ac3ck, *am3ck, *ap3sk, *ac3sk, *am3sk, *ap3ssk, *ac3ssk, *am3ssk;
    double *ak1, *ak2, *amphk, *acphk, *apphk;
    double *xcr, *xmr, *ycr, *ymr, *zcr, *zmr, *g3rcr, *g3rmr, *d3xcr, *d3xmr, *udx3cr, *udx3mr;
    double *ap3sskr, *ac3sskr, *am3sskr;
    int *irangs, *jrangs, *krangs, *irangb, *jrangb, *krangb;
    double *cxrs, *cyrs, *czrs, *cxsalc, *cysalc, *czsalc;   /* (4, n) column major */
    /* phase field (phasefield.F90:17-29, ReadInputFile.F90:125-126) */
    double pf_a, pf_s, pf_tm, pf_d, pf_eps;
    double *ap3spkr, *ac3spkr, *am3spkr;                     /* (1:nxr), Neumann (CreateMgrdGrid.F90:163) */
    double *cxphic, *cyphic, *czphic;                        /* refined -> coarse centres, (4, n) */
    /* moisture (moisture.F90:15-23, ReadMoistParameters :64-117) */
    double q_alpha, q_beta, q_gamma, q_tau, pecq, kappa_q;
    int q_fixn, q_fixs;
    int *irangr, *jrangr, *krangr;                           /* refined -> coarse centre indices (0:n_ref+1) */
    /* 3-D fields (x fastest, no halo) */
    double *f[ORC_NFIELDS];
    double *tpdv, *tpdvr;   /* x index -1..nx+1 / -1..nxr+1 */
    /* immersed boundary (src/ibm/ibm_param.F90): per solve (vx, vy, vz, T,
     * S) the mask (2 liquid, 0 solid, 1 / -1 upper / lower boundary point;
     * x fastest) and the boundary distances in the reference's loop order */
    int32_t *ibm_mask[5];
    double *ibm_dist[5];
    int64_t ibm_ndist[5];
    void *allocs[MAXALLOC];
    int nalloc;
};

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }

/* Planted mistakes, for the tests that show each pin catches a plausible
 * error (0 = none, the faithful restatement): 1 flips the sign of the
 * pressure gradient in the vx RHS; 2 drops the alpha*beta term of
 * CorrectPressure; 3 drops the 1/2 of the vx buoyancy average; 4 leaves vy
 * out of the kinetic energy; 5 takes vx at the lower face instead of the
 * cell-centre average in the Nusselt sum; 6 drops alpha from the x diffusion
 * of the vz RHS; 7 flips the sign of the condensation heating; 8 takes the
 * humidity x-diffusion with pect instead of pecq. */
static int g_plant = 0;
void orc_set_plant(int n) { g_plant = n; }
static void apply_threads(void) {
#ifdef _OPENMP
    if (g_threads > 0) omp_set_num_threads(g_threads);
#endif
}

static void *track(orc_state *s, size_t bytes) {
    void *p = calloc(1, bytes ? bytes : 8);
    if (!p) { fprintf(stderr, "oracle: out of memory (%zu bytes)\n", bytes); abort(); }
    if (s->nalloc >= MAXALLOC) { fprintf(stderr, "oracle: too many allocations\n"); abort(); }
    s->allocs[s->nalloc++] = p;
    return p;
}
/* real array valid on Fortran bounds lo..hi */
static double *r1(orc_state *s, int lo, int hi) { return (double *)track(s, sizeof(double) * (size_t)(hi - lo + 1)) - lo; }
static int *i1(orc_state *s, int lo, int hi) { return (int *)track(s, sizeof(int) * (size_t)(hi - lo + 1)) - lo; }

/* 1-based periodic index -> 0-based storage index */
static inline int W(int j, int n) { int r = (j - 1) % n; return r < 0 ? r + n : r; }

/* ------------------------------------------------------------------ */
/* Grid module: src/grid.F90                                           */
/* ------------------------------------------------------------------ */
/* uniform_grid, src/grid.F90:12-28 ; c(1:Nm+1), m(1:Nm) */
static void uniform_grid(double *c, double *m, int Nm, double L) {
    for (int i = 1; i <= Nm + 1; i++) c[i] = L * (double)(i - 1) / (double)Nm;
    for (int i = 1; i <= Nm; i++) m[i] = 0.5 * (c[i] + c[i + 1]);
}
/* tanh_grid, src/grid.F90:30-53 */
static void tanh_grid(double *c, double *m, int Nm, double L, double str) {
    double tstr = tanh(str);
    c[1] = 0.0;
    for (int i = 2; i <= Nm + 1; i++) {
        double z2 = (double)(2 * (i - 1) - Nm) / (double)Nm;
        c[i] = 0.5 * (1 + tanh(str * z2) / tstr) * L;
    }
    for (int i = 1; i <= Nm; i++) m[i] = 0.5 * (c[i] + c[i + 1]);
}
/* cheb_grid, src/grid.F90:55-94 (clipped Chebyshev, nclip = int(str)) */
static void cheb_grid(double *c, double *m, int Nm, double L, double str, double pi) {
    int nclip = (int)str, Next = Nm + 1 + 2 * nclip;
    double *etazm = malloc(sizeof(double) * (Next + 1)), *etaz = malloc(sizeof(double) * (Nm + 2));
    for (int i = 1; i <= Next; i++) etazm[i] = cos(pi * ((double)i - 0.5) / (double)Next);
    for (int i = 1; i <= Nm + 1; i++) etaz[i] = etazm[i + nclip];
    double delet = etaz[1] - etaz[Nm + 1];
    for (int i = 1; i <= Nm + 1; i++) etaz[i] = 2.0 * etaz[i] / delet;
    c[1] = 0.0;
    for (int i = 2; i <= Nm; i++) c[i] = 0.5 * L * (1.0 - etaz[i]);
    c[Nm + 1] = L;
    for (int i = 1; i <= Nm; i++) m[i] = 0.5 * (c[i] + c[i + 1]);
    free(etazm); free(etaz);
}
/* asym_cheb_grid, src/grid.F90:96-131 */
static void asym_cheb_grid(double *c, double *m, int Nm, double L, double str, double pi) {
    int nclip = (int)str, Next = Nm + 1 + nclip;
    double *etazm = malloc(sizeof(double) * (Next + 1)), *etaz = malloc(sizeof(double) * (Nm + 2));
    for (int i = 1; i <= Next; i++) etazm[i] = cos(pi * (double)i / (double)Next / 2.0);
    for (int i = 1; i <= Nm + 1; i++) etaz[i] = etazm[i + nclip];
    double delet = etaz[1];
    for (int i = 1; i <= Nm + 1; i++) etaz[i] = etaz[i] / delet;
    c[1] = 0.0;
    for (int i = 2; i <= Nm; i++) c[i] = L * (1.0 - etaz[i]);
    c[Nm + 1] = L;
    for (int i = 1; i <= Nm; i++) m[i] = 0.5 * (c[i] + c[i + 1]);
    free(etazm); free(etaz);
}
/* centre_focus_grid, src/grid.F90:133-152 */
static void centre_focus_grid(double *c, double *m, int Nm, double L, double str) {
    c[1] = 0.0;
    for (int i = 2; i <= Nm; i++) {
        double q = 2.0 * (i - 1) / Nm - 1;
        c[i] = L / 2.0 * (str * pow(q, 7) + q) / (str + 1) + L / 2.0;
    }
    c[Nm + 1] = L;
    for (int i = 1; i <= Nm; i++) m[i] = 0.5 * (c[i] + c[i + 1]);
}
/* natural_BL_grid, src/grid.F90:154-180 */
static void natural_BL_grid(double *c, double *m, int Nm, double L) {
    double jb = 16.0, ceta = 0.8, alpha = 1.5, dyw = 0.1;
    for (int j = 1; j <= Nm + 1; j++) {
        double jm = j - 1;
        c[j] = 1 / (1.0 + pow(jm / jb, 2)) * (dyw * jm + pow(0.75 * alpha * ceta * jm, 4.0 / 3.0) * pow(jm / jb, 2));
    }
    double top = c[Nm + 1];
    for (int j = 1; j <= Nm + 1; j++) c[j] = c[j] / top * L;
    for (int j = 1; j <= Nm; j++) m[j] = 0.5 * (c[j] + c[j + 1]);
}
/* sym_natural_BL_grid, src/grid.F90:182-215 */
static void sym_natural_BL_grid(double *c, double *m, int Nm, double L, double Sc) {
    double jb = 16.0, ceta = 0.8, alpha = 1.5, dyw = 0.1;
    int nh = Nm / 2;
    for (int j = 1; j <= nh + 1; j++) {
        double jm = j - 1;
        c[j] = 1 / (1.0 + pow(jm / jb, 2)) * (dyw * jm + pow(0.75 * alpha * ceta * jm / sqrt(Sc), 4.0 / 3.0) * pow(jm / jb, 2));
    }
    double Lh = 2.0 * c[nh + 1];
    for (int j = 1; j <= nh; j++) c[Nm + 2 - j] = Lh - c[j];
    double top = c[Nm + 1];
    for (int j = 1; j <= Nm + 1; j++) c[j] = c[j] / top * L;
    for (int j = 1; j <= Nm; j++) m[j] = 0.5 * (c[j] + c[j + 1]);
}
/* scallop_grid, src/grid.F90:217-262 (c(1) assumed 0 on entry) */
static void scallop_grid(double *c, double *m, int Nm, double L, double Retau, double dw) {
    double kb = 0.2 * Retau / dw;
    double alpha = 8.0 / (Nm + 1 - kb) * sqrt(Retau) * (1.0 - pow(5.0, -0.5));
    int ks = 20;
    c[1] = 0.0;
    for (int k = 1; k <= Nm; k++) {
        double dxlo = dw;
        double dxup = alpha / 4.0 * (alpha / 8.0 * (k - kb) + sqrt(Retau / 5.0));
        double sig = 0.5 * (1.0 + tanh((k - kb) / ks));
        c[k + 1] = c[k] + (1.0 - sig) * dxlo + sig * dxup;
    }
    double top = c[Nm + 1];
    for (int k = 1; k <= Nm + 1; k++) c[k] = c[k] / top * L;
    for (int k = 1; k <= Nm; k++) m[k] = 0.5 * (c[k] + c[k + 1]);
}
/* Full-cosine clustering x_k = L/2 (1 - cos(pi k / Nm)), k = 0..Nm.  Not a
 * reference option: the workload definition asks for this grid; istr3 = 8. */
static void cosine_grid(double *c, double *m, int Nm, double L, double pi) {
    c[1] = 0.0;
    for (int i = 2; i <= Nm; i++) c[i] = 0.5 * L * (1.0 - cos(pi * (double)(i - 1) / (double)Nm));
    c[Nm + 1] = L;
    for (int i = 1; i <= Nm; i++) m[i] = 0.5 * (c[i] + c[i + 1]);
}
/* second_derivative_coeff, src/grid.F90:264-307; x(1:n) */
static void second_derivative_coeff(double *ap3, double *ac3, double *am3, const double *x, int n,
                                    double xlen, int fix_up, int fix_low) {
    int k = 1;
    double a33 = 2.0 / (x[k] + x[k + 1]);
    ap3[k] = a33 / (x[k + 1] - x[k]);
    am3[k] = 0.0;
    ac3[k] = -ap3[k] - fix_low * a33 / x[k];
    if (x[k] == 0.0) { ap3[k] = 0.0; ac3[k] = 1.0; }
    for (k = 2; k <= n - 1; k++) {
        a33 = 2.0 / (x[k + 1] - x[k - 1]);
        ap3[k] = a33 / (x[k + 1] - x[k]);
        am3[k] = a33 / (x[k] - x[k - 1]);
        ac3[k] = -ap3[k] - am3[k];
    }
    k = n;
    a33 = 2.0 / (2.0 * xlen - x[k] - x[k - 1]);
    ap3[k] = 0.0;
    am3[k] = a33 / (x[k] - x[k - 1]);
    ac3[k] = -am3[k] - fix_up * a33 / (xlen - x[k]);
    if (x[k] == xlen) { am3[k] = 0.0; ac3[k] = 1.0; }
}

static int make_xgrid(orc_state *s, int istr3, double *c, double *m, int Nm, int refined) {
    double L = s->c.alx3, str = s->c.str3;
    switch (istr3) {
    case 0: uniform_grid(c, m, Nm, L); break;
    case 1: centre_focus_grid(c, m, Nm, L, str); break;
    case 2: if (refined) return -1; natural_BL_grid(c, m, Nm, L); break;
    case 3: sym_natural_BL_grid(c, m, Nm, L, refined ? s->c.pras : 1.0); break;
    case 4: tanh_grid(c, m, Nm, L, str); break;
    case 5: scallop_grid(c, m, Nm, L, s->c.dpdy, 0.5); break;
    case 6: cheb_grid(c, m, Nm, L, str, s->pi); break;
    case 7: asym_cheb_grid(c, m, Nm, L, str, s->pi); break;
    case 8: cosine_grid(c, m, Nm, L, s->pi); break;
    default: return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Hermite stencils: src/multires/HermiteInterpolations.F90:28-178       */
/* ------------------------------------------------------------------ */
/* interpolation_indices, HermiteInterpolations.F90:28-58 ; idx(0:n_old+1) */
static void interpolation_indices(int *idx, const double *xo, int n_old, const double *xnew, int n_new, double L) {
    double *xn = malloc(sizeof(double) * (n_new + 2));
    for (int i = 1; i <= n_new; i++) xn[i] = xnew[i];
    xn[0] = -xn[1];
    if (xn[0] == xn[1]) xn[0] = -xn[2];
    xn[n_new + 1] = 2.0 * L - xn[n_new];
    if (xn[n_new + 1] == xn[n_new]) xn[n_new + 1] = 2.0 * L - xn[n_new - 1];
    idx[0] = 1;
    for (int io = 1; io <= n_old; io++) {
        idx[io] = 0;
        for (int in = 0; in <= n_new; in++)
            if (xn[in] < xo[io] && xn[in + 1] >= xo[io]) idx[io] = in + 1;
    }
    idx[n_old + 1] = n_new + 1;
    free(xn);
}
/* extend_grid, HermiteInterpolations.F90:62-103 ; returns array valid on 1-next..n+next */
static double *extend_grid(const double *x, int n, double L, int next) {
    double *base = malloc(sizeof(double) * (n + 2 * next)), *e = base + next - 1;
    for (int i = 1; i <= n; i++) e[i] = x[i];
    if (fabs(x[1]) < 1e-10) { for (int i = 1; i <= next; i++) e[1 - i] = -x[1 + i]; }
    else                    { for (int i = 1; i <= next; i++) e[1 - i] = -x[i]; }
    if (fabs(x[n] - L) < 1e-10) { for (int i = 1; i <= next; i++) e[n + i] = 2 * L - x[n - i]; }
    else                        { for (int i = 1; i <= next; i++) e[n + i] = 2 * L - x[n + 1 - i]; }
    return e;
}
/* construct_stencil, HermiteInterpolations.F90:107-178 ; cx(4, n_new) column-major: cx[(in-1)*4 + r] */
static void construct_stencil(double *cx, const double *xold, int n_old, const double *xnew, int n_new,
                              double L, const int *idx, char axis) {
    int next = axis == 'x' ? 1 : 2;
    memset(cx, 0, sizeof(double) * 4 * (size_t)n_new);
    double *xo = extend_grid(xold, n_old, L, next), *xn = extend_grid(xnew, n_new, L, next);
    for (int io = 0; io <= n_old; io++) {
        int lo = idx[io] > 1 ? idx[io] : 1, hi = idx[io + 1] - 1 < n_new ? idx[io + 1] - 1 : n_new;
        if (axis == 'x' && (io == 0 || io == n_old)) {
            double dlc = xo[io + 1] - xo[io];
            for (int in = lo; in <= hi; in++) {
                double t = (xn[in] - xo[io]) / dlc;
                double *w = cx + (size_t)(in - 1) * 4;
                w[0] = 0.0; w[1] = 1 - t; w[2] = t; w[3] = 0.0;
            }
        } else {
            double dlm = xo[io] - xo[io - 1], dlc = xo[io + 1] - xo[io], dlp = xo[io + 2] - xo[io + 1];
            for (int in = lo; in <= hi; in++) {
                double t = (xn[in] - xo[io]) / dlc;
                double *w = cx + (size_t)(in - 1) * 4;
                if (t < 1e-12) { w[0] = 0.0; w[1] = 1.0; w[2] = 0.0; w[3] = 0.0; }
                else {
                    double h00 = (1.0 + 2.0 * t) * (1.0 - t) * (1.0 - t);
                    double h10 = t * (1.0 - t) * (1.0 - t);
                    double h01 = (1.0 + 2.0 * (1.0 - t)) * t * t;
                    double h11 = -(1.0 - t) * t * t;
                    w[0] = -h10 * dlc * dlc / dlm / (dlc + dlm);
                    w[1] = h00 - h11 * dlp / (dlp + dlc) + h10 * (dlc - dlm) / dlm;
                    w[2] = h01 + h10 * dlm / (dlm + dlc) + h11 * (dlp - dlc) / dlp;
                    w[3] = h11 * dlc * dlc / dlp / (dlp + dlc);
                }
            }
        }
    }
    free(xo + (1 - next)); free(xn + (1 - next));
}

static void get_twiddles(int n, const double **twr, const double **twi);

/* ------------------------------------------------------------------ */
/* creation                                                            */
/* ------------------------------------------------------------------ */
static size_t nbase(orc_state *s) { return (size_t)s->nx * s->nym * s->nzm; }
static size_t nref(orc_state *s) { return (size_t)s->nxr * s->nymr * s->nzmr; }

orc_state *orc_create(const orc_config *cfg) {
    orc_state *s = calloc(1, sizeof(orc_state));
    s->c = *cfg;
    s->pi = 2.0 * asin(1.0);
    s->nxm = cfg->nxm; s->nym = cfg->nym; s->nzm = cfg->nzm;
    s->nx = s->nxm + 1; s->ny = s->nym + 1; s->nz = s->nzm + 1;
    int mr = cfg->multires || cfg->salinity || cfg->phasefield;
    s->nxmr = mr ? cfg->nxmr : s->nxm; s->nymr = mr ? cfg->nymr : s->nym; s->nzmr = mr ? cfg->nzmr : s->nzm;
    s->nxr = s->nxmr + 1; s->nyr = s->nymr + 1; s->nzr = s->nzmr + 1;
    if (cfg->phasefield && cfg->salinity) { fprintf(stderr, "oracle: phase field with salinity not restated\n"); free(s); return NULL; }

    /* derived constants, ReadInputFile.F90:101-122 */
    s->rrho = fabs(cfg->rayt) * cfg->pras / (fabs(cfg->rays) * cfg->prat);
    if (cfg->salinity && cfg->ffscales == 1) { s->ren = sqrt(fabs(cfg->rays) / cfg->pras); s->byct = s->rrho; s->bycs = 1.0; }
    else { s->ren = sqrt(fabs(cfg->rayt) / cfg->prat); s->byct = 1.0; s->bycs = 1.0 / s->rrho; }
    s->pect = s->ren * cfg->prat;
    s->pecs = s->ren * cfg->pras;
    /* ReadInputFile.F90:125-126: pf_D follows from the Stefan number and A */
    s->pf_a = cfg->pf_a; s->pf_s = cfg->pf_s; s->pf_tm = cfg->pf_tm;
    if (cfg->phasefield) {
        s->pf_d = 1.2 / s->pect / cfg->pf_s / cfg->pf_a;
        s->pf_eps = cfg->pf_eps > 0.0 ? cfg->pf_eps : 1.0 / s->nxmr;
    }
    /* moisture.F90:99-110: pecq = pect / Sm, kappa_q = 1 / pecq, tau_q in
     * diffusive units when ifdiff, and gamma_q < 0 chosen so that the
     * moist-static jump is 1 (beta_q with the unsaturated top, qfixN = 2) */
    if (cfg->moist) {
        s->q_alpha = cfg->q_alpha; s->q_beta = cfg->q_beta;
        s->pecq = s->pect / cfg->q_sm;
        s->kappa_q = 1.0 / s->pecq;
        s->q_tau = cfg->q_ifdiff ? s->pect * cfg->q_tau : cfg->q_tau;
        s->q_gamma = cfg->q_gamma;
        if (s->q_gamma < 0.0) s->q_gamma = cfg->q_fixn == 2 ? s->q_beta : s->q_beta / (1.0 - exp(-s->q_alpha));
        /* SetHumidityBCs (:139-147) turns qfixN = 2 into a zero top value with qfixN = 1 */
        s->q_fixn = cfg->q_fixn == 2 ? 1 : cfg->q_fixn;
        s->q_fixs = cfg->q_fixs;
    }

    /* InitTimeMarchScheme.F90:17-49 */
    if (cfg->nsst > 1) {
        s->gam[0] = 8.0 / 15.0; s->gam[1] = 5.0 / 12.0; s->gam[2] = 3.0 / 4.0;
        s->rom[0] = 0.0; s->rom[1] = -17.0 / 60.0; s->rom[2] = -5.0 / 12.0;
    } else {
        s->gam[0] = 1.5; s->gam[1] = 0; s->gam[2] = 0; s->rom[0] = -0.5; s->rom[1] = 0; s->rom[2] = 0;
    }
    for (int n = 0; n < 3; n++) s->alm[n] = s->gam[n] + s->rom[n];

    int nx = s->nx, ny = s->ny, nz = s->nz, nxm = s->nxm, nym = s->nym, nzm = s->nzm;
    s->xc = r1(s, 1, nx); s->xm = r1(s, 0, nx); s->yc = r1(s, 1, ny); s->ym = r1(s, 0, ny);
    s->zc = r1(s, 1, nz); s->zm = r1(s, 0, nz);
    s->g3rc = r1(s, 1, nx); s->g3rm = r1(s, 1, nx); s->d3xc = r1(s, 1, nx); s->d3xm = r1(s, 1, nx);
    s->udx3c = r1(s, 1, nx); s->udx3m = r1(s, 1, nx);
    s->kmv = i1(s, 1, nx); s->kpv = i1(s, 1, nx); s->kmc = i1(s, 1, nx); s->kpc = i1(s, 1, nx);
    s->ap3ck = r1(s, 1, nx); s->ac3ck = r1(s, 1, nx); s->am3ck = r1(s, 1, nx);
    s->ap3sk = r1(s, 1, nx); s->ac3sk = r1(s, 1, nx); s->am3sk = r1(s, 1, nx);
    s->ap3ssk = r1(s, 1, nx); s->ac3ssk = r1(s, 1, nx); s->am3ssk = r1(s, 1, nx);
    s->ak1 = r1(s, 1, nz); s->ak2 = r1(s, 1, ny);
    s->amphk = r1(s, 1, nx); s->acphk = r1(s, 1, nx); s->apphk = r1(s, 1, nx);

    /* CreateGrid.F90:19-178 */
    for (int kc = 1; kc <= nxm; kc++) {
        s->kmv[kc] = kc - 1; s->kpv[kc] = kc + 1;
        if (kc == 1) s->kmv[kc] = kc;
        if (kc == nxm) s->kpv[kc] = kc;
    }
    for (int kc = 1; kc <= nxm; kc++) { s->kpc[kc] = s->kpv[kc] - kc; s->kmc[kc] = kc - s->kmv[kc]; }
    uniform_grid(s->zc, s->zm, nzm, cfg->zlen);
    s->zm[0] = 2.0 * s->zm[1] - (nzm > 1 ? s->zm[2] : s->zm[1] + cfg->zlen);
    s->zm[nz] = 2.0 * s->zm[nzm] - (nzm > 1 ? s->zm[nzm - 1] : s->zm[nzm] - cfg->zlen);
    uniform_grid(s->yc, s->ym, nym, cfg->ylen);
    s->ym[0] = 2.0 * s->ym[1] - (nym > 1 ? s->ym[2] : s->ym[1] + cfg->ylen);
    s->ym[ny] = 2.0 * s->ym[nym] - (nym > 1 ? s->ym[nym - 1] : s->ym[nym] - cfg->ylen);
    if (make_xgrid(s, cfg->istr3, s->xc, s->xm, nxm, 0)) { fprintf(stderr, "oracle: istr3 %d unsupported\n", cfg->istr3); orc_destroy(s); return NULL; }
    s->dx = (double)nxm / cfg->alx3; s->dy = (double)nym / cfg->ylen; s->dz = (double)nzm / cfg->zlen;
    s->dxq = s->dx * s->dx; s->dyq = s->dy * s->dy; s->dzq = s->dz * s->dz;
    for (int kc = 1; kc <= nxm; kc++) s->xm[kc] = (s->xc[kc] + s->xc[kc + 1]) * 0.5;
    s->xm[nx] = 2 * s->xc[nx] - s->xm[nxm];
    for (int kc = 2; kc <= nxm; kc++) {
        s->g3rc[kc] = (s->xc[kc + 1] - s->xc[kc - 1]) * s->dx * 0.5;
        s->g3rm[kc] = (s->xm[kc + 1] - s->xm[kc - 1]) * s->dx * 0.5;
    }
    s->g3rm[1] = (s->xm[2] - (2 * s->xc[1] - s->xm[1])) * s->dx * 0.5;
    s->g3rc[1] = (s->xc[2] - s->xc[1]) * s->dx;
    s->g3rc[nx] = (s->xc[nx] - s->xc[nxm]) * s->dx;
    for (int kc = 1; kc <= nxm; kc++) {
        s->d3xc[kc] = (s->xc[kc + 1] - s->xc[kc]) * s->dx;
        s->d3xm[kc] = (s->xm[kc + 1] - s->xm[kc]) * s->dx;
    }
    for (int kc = 1; kc <= nxm; kc++) { s->udx3m[kc] = s->dx / s->d3xc[kc]; s->udx3c[kc] = s->dx / s->g3rc[kc]; }
    s->udx3c[nx] = s->dx / s->g3rc[nx];
    second_derivative_coeff(s->ap3ck, s->ac3ck, s->am3ck, s->xc, nx, cfg->alx3, 1, 1);
    second_derivative_coeff(s->ap3sk, s->ac3sk, s->am3sk, s->xm, nxm, cfg->alx3, cfg->inslwn, cfg->inslws);
    second_derivative_coeff(s->ap3ssk, s->ac3ssk, s->am3ssk, s->xm, nxm, cfg->alx3, cfg->tfixn, cfg->tfixs);

    /* InitPressureSolver, pressure.F90:68-114 */
    {
        int nymh = nym / 2 + 1, nymp = nymh + 1, nzmh = nzm / 2 + 1, nzmp = nzmh + 1;
        double *ao = malloc(sizeof(double) * (nz + 1)), *ap = malloc(sizeof(double) * (ny + 1));
        for (int i = 1; i <= nzmh; i++) ao[i] = (i - 1) * 2.0 * s->pi;
        for (int i = nzmp; i <= nzm; i++) ao[i] = -(nzm - i + 1) * 2.0 * s->pi;
        for (int i = 1; i <= nzm; i++) s->ak1[i] = 2.0 * (1.0 - cos(ao[i] / nzm)) * pow((double)nzm / cfg->zlen, 2);
        for (int j = 1; j <= nymh; j++) ap[j] = (j - 1) * 2.0 * s->pi;
        for (int j = nymp; j <= nym; j++) ap[j] = -(nym - j + 1) * 2.0 * s->pi;
        for (int j = 1; j <= nym; j++) s->ak2[j] = 2.0 * (1.0 - cos(ap[j] / nym)) * pow((double)nym / cfg->ylen, 2);
        free(ao); free(ap);
        for (int kc = 1; kc <= nxm; kc++) {
            int km = s->kmv[kc];
            double a33icc = s->kmc[kc] * s->dxq / s->d3xm[km];
            double a33icp = s->kpc[kc] * s->dxq / s->d3xm[kc];
            double ugmmm = 1.0 / s->d3xc[kc];
            s->amphk[kc] = a33icc * ugmmm;
            s->apphk[kc] = a33icp * ugmmm;
            s->acphk[kc] = -(s->amphk[kc] + s->apphk[kc]);
        }
    }

    { const double *a, *b; get_twiddles(nym, &a, &b); get_twiddles(nzm, &a, &b); }

    /* fields */
    for (int id = ORC_VX; id <= ORC_RHS; id++) s->f[id] = track(s, sizeof(double) * nbase(s));
    s->f[ORC_SALC] = track(s, sizeof(double) * nbase(s));
    s->f[ORC_TEMPBP] = track(s, sizeof(double) * (size_t)nym * nzm);
    s->f[ORC_TEMPTP] = track(s, sizeof(double) * (size_t)nym * nzm);
    s->tpdv = track(s, sizeof(double) * (size_t)(nx + 3) * nym * nzm);
    if (cfg->moist) {
        for (int id = ORC_HUMID; id <= ORC_QSAT; id++) s->f[id] = track(s, sizeof(double) * nbase(s));
        s->f[ORC_HUMBP] = track(s, sizeof(double) * (size_t)nym * nzm);
        s->f[ORC_HUMTP] = track(s, sizeof(double) * (size_t)nym * nzm);
    }

    if (cfg->salinity || cfg->phasefield) {
        int nxr = s->nxr, nyr = s->nyr, nzr = s->nzr, nxmr = s->nxmr, nymr = s->nymr, nzmr = s->nzmr;
        if (cfg->salinity) {
            for (int id = ORC_SAL; id <= ORC_VZR; id++) s->f[id] = track(s, sizeof(double) * nref(s));
            s->f[ORC_SALBP] = track(s, sizeof(double) * (size_t)nymr * nzmr);
            s->f[ORC_SALTP] = track(s, sizeof(double) * (size_t)nymr * nzmr);
        } else {
            /* phase field: refined phi, its RK arrays, tempr and rhsr; coarse phic */
            for (int id = ORC_PHI; id <= ORC_HPHI; id++) s->f[id] = track(s, sizeof(double) * nref(s));
            s->f[ORC_TEMPR] = track(s, sizeof(double) * nref(s));
            s->f[ORC_RHSR] = track(s, sizeof(double) * nref(s));
            s->f[ORC_PHIC] = track(s, sizeof(double) * nbase(s));
        }
        s->tpdvr = track(s, sizeof(double) * (size_t)(nxr + 3) * nymr * nzmr);
        s->xcr = r1(s, 1, nxr); s->xmr = r1(s, 0, nxr); s->ycr = r1(s, 1, nyr); s->ymr = r1(s, 0, nyr);
        s->zcr = r1(s, 1, nzr); s->zmr = r1(s, 0, nzr);
        s->g3rcr = r1(s, 1, nxr); s->g3rmr = r1(s, 1, nxr); s->d3xcr = r1(s, 1, nxr); s->d3xmr = r1(s, 1, nxr);
        s->udx3cr = r1(s, 1, nxr); s->udx3mr = r1(s, 1, nxr);
        s->ap3sskr = r1(s, 1, nxr); s->ac3sskr = r1(s, 1, nxr); s->am3sskr = r1(s, 1, nxr);
        /* CreateMgrdGrid.F90:38-161 */
        uniform_grid(s->zcr, s->zmr, nzmr, cfg->zlen);
        s->zmr[0] = 2.0 * s->zmr[1] - s->zmr[2];
        s->zmr[nzr] = 2.0 * s->zmr[nzmr] - s->zmr[nzmr - 1];
        uniform_grid(s->ycr, s->ymr, nymr, cfg->ylen);
        s->ymr[0] = 2.0 * s->ymr[1] - s->ymr[2];
        s->ymr[nyr] = 2.0 * s->ymr[nymr] - s->ymr[nymr - 1];
        /* CreateMgrdGrid.F90:77 tests istr3 (not istr3r) for the scallop option */
        int rc = (cfg->istr3 == 5) ? make_xgrid(s, 5, s->xcr, s->xmr, nxmr, 1)
                                   : make_xgrid(s, cfg->istr3r, s->xcr, s->xmr, nxmr, 1);
        if (rc) { fprintf(stderr, "oracle: istr3r %d unsupported\n", cfg->istr3r); orc_destroy(s); return NULL; }
        s->dxr = (double)nxmr / cfg->alx3; s->dyr = (double)nymr / cfg->ylen; s->dzr = (double)nzmr / cfg->zlen;
        s->dxqr = s->dxr * s->dxr; s->dyqr = s->dyr * s->dyr; s->dzqr = s->dzr * s->dzr;
        for (int kc = 1; kc <= nxmr; kc++) s->xmr[kc] = (s->xcr[kc] + s->xcr[kc + 1]) * 0.5;
        s->xmr[0] = 2.0 * s->xcr[1] - s->xmr[1];
        s->xmr[nxr] = 2.0 * s->xcr[nxr] - s->xmr[nxmr];
        for (int kc = 2; kc <= nxmr; kc++) {
            s->g3rcr[kc] = (s->xcr[kc + 1] - s->xcr[kc - 1]) * s->dxr * 0.5;
            s->g3rmr[kc] = (s->xmr[kc + 1] - s->xmr[kc - 1]) * s->dxr * 0.5;
        }
        s->g3rmr[1] = (s->xmr[2] - (2 * s->xcr[1] - s->xmr[1])) * s->dxr * 0.5;
        s->g3rcr[1] = (s->xcr[2] - s->xcr[1]) * s->dxr;
        s->g3rcr[nxr] = (s->xcr[nxr] - s->xcr[nxmr]) * s->dxr;
        for (int kc = 1; kc <= nxmr; kc++) {
            s->d3xcr[kc] = (s->xcr[kc + 1] - s->xcr[kc]) * s->dxr;
            s->d3xmr[kc] = (s->xmr[kc + 1] - s->xmr[kc]) * s->dxr;
        }
        for (int kc = 1; kc <= nxmr; kc++) { s->udx3mr[kc] = s->dxr / s->d3xcr[kc]; s->udx3cr[kc] = s->dxr / s->g3rcr[kc]; }
        s->udx3cr[nxr] = s->dxr / s->g3rcr[nxr];
        second_derivative_coeff(s->ap3sskr, s->ac3sskr, s->am3sskr, s->xmr, nxmr, cfg->alx3, cfg->sfixn, cfg->sfixs);
        s->ap3spkr = r1(s, 1, nxr); s->ac3spkr = r1(s, 1, nxr); s->am3spkr = r1(s, 1, nxr);
        second_derivative_coeff(s->ap3spkr, s->ac3spkr, s->am3spkr, s->xmr, nxmr, cfg->alx3, 0, 0);

        /* CreateMgrdStencil.F90:17-78 (the stencils the substep uses) */
        s->irangs = i1(s, 0, nx); s->jrangs = i1(s, 0, ny); s->krangs = i1(s, 0, nz); s->irangb = i1(s, 0, nx + 1);
        s->jrangb = i1(s, 0, ny); s->krangb = i1(s, 0, nz);
        s->cxrs = track(s, sizeof(double) * 4 * nxmr); s->cyrs = track(s, sizeof(double) * 4 * nymr);
        s->czrs = track(s, sizeof(double) * 4 * nzmr);
        s->cxsalc = track(s, sizeof(double) * 4 * nx); s->cysalc = track(s, sizeof(double) * 4 * nym);
        s->czsalc = track(s, sizeof(double) * 4 * nzm);
        interpolation_indices(s->irangs, s->xm, nxm, s->xmr, nxmr, cfg->alx3);
        construct_stencil(s->cxrs, s->xm, nxm, s->xmr, nxmr, cfg->alx3, s->irangs, 'x');
        interpolation_indices(s->jrangs, s->ym, nym, s->ymr, nymr, cfg->ylen);
        construct_stencil(s->cyrs, s->ym, nym, s->ymr, nymr, cfg->ylen, s->jrangs, 'y');
        interpolation_indices(s->krangs, s->zm, nzm, s->zmr, nzmr, cfg->zlen);
        construct_stencil(s->czrs, s->zm, nzm, s->zmr, nzmr, cfg->zlen, s->krangs, 'z');
        {
            int *irangr = i1(s, 0, nxr), *jrangr = i1(s, 0, nyr), *krangr = i1(s, 0, nzr);
            s->irangr = irangr; s->jrangr = jrangr; s->krangr = krangr;
            double *cxphic = track(s, sizeof(double) * 4 * nxm);
            s->cxphic = cxphic;
            s->cyphic = track(s, sizeof(double) * 4 * nym); s->czphic = track(s, sizeof(double) * 4 * nzm);
            interpolation_indices(irangr, s->xmr, nxmr, s->xm, nxm, cfg->alx3);
            construct_stencil(cxphic, s->xmr, nxmr, s->xm, nxm, cfg->alx3, irangr, 'x');
            interpolation_indices(jrangr, s->ymr, nymr, s->ym, nym, cfg->ylen);
            construct_stencil(s->cyphic, s->ymr, nymr, s->ym, nym, cfg->ylen, jrangr, 'y');
            interpolation_indices(krangr, s->zmr, nzmr, s->zm, nzm, cfg->zlen);
            construct_stencil(s->czphic, s->zmr, nzmr, s->zm, nzm, cfg->zlen, krangr, 'z');
            /* S to the coarse grid: faces along gAxis (CreateMgrdStencil.F90:25-32, :47-54, :69-76) */
            for (int k = 0; k <= nx; k++) s->irangb[k] = s->irangs[k];
            for (int j = 0; j <= ny; j++) s->jrangb[j] = s->jrangs[j];
            for (int k = 0; k <= nz; k++) s->krangb[k] = s->krangs[k];
            memcpy(s->cxsalc, cxphic, sizeof(double) * 4 * nxm);
            memcpy(s->cysalc, s->cyphic, sizeof(double) * 4 * nym);
            memcpy(s->czsalc, s->czphic, sizeof(double) * 4 * nzm);
            if (cfg->gaxis == 1) {
                int *irange = malloc(sizeof(int) * (nxr + 1));
                interpolation_indices(irange, s->xmr, nxmr, s->xc, nx, cfg->alx3);
                construct_stencil(s->cxsalc, s->xmr, nxmr, s->xc, nx, cfg->alx3, irange, 'x');
                interpolation_indices(s->irangb, s->xc, nx, s->xmr, nxmr, cfg->alx3);
                free(irange);
            } else if (cfg->gaxis == 2) {
                int *jrange = malloc(sizeof(int) * (nyr + 1));
                interpolation_indices(jrange, s->ymr, nymr, s->yc, nym, cfg->ylen);
                construct_stencil(s->cysalc, s->ymr, nymr, s->yc, nym, cfg->ylen, jrange, 'y');
                interpolation_indices(s->jrangb, s->yc, nym, s->ymr, nymr, cfg->ylen);
                free(jrange);
            } else if (cfg->gaxis == 3) {
                int *krange = malloc(sizeof(int) * (nzr + 1));
                interpolation_indices(krange, s->zmr, nzmr, s->zc, nzm, cfg->zlen);
                construct_stencil(s->czsalc, s->zmr, nzmr, s->zc, nzm, cfg->zlen, krange, 'z');
                interpolation_indices(s->krangb, s->zc, nzm, s->zmr, nzmr, cfg->zlen);
                free(krange);
            }
        }
    }
    return s;
}

void orc_destroy(orc_state *s) {
    if (!s) return;
    for (int i = 0; i < s->nalloc; i++) free(s->allocs[i]);
    free(s);
}

double *orc_field(orc_state *s, int id, int64_t dims[3]) {
    if (id < 0 || id >= ORC_NFIELDS || !s->f[id]) return NULL;
    int refined = (id >= ORC_SAL && id <= ORC_VZR) || id == ORC_SALBP || id == ORC_SALTP ||
                  id == ORC_PHI || id == ORC_RUPHI || id == ORC_HPHI || id == ORC_TEMPR;
    int plane = id == ORC_TEMPBP || id == ORC_TEMPTP || id == ORC_SALBP || id == ORC_SALTP ||
                id == ORC_HUMBP || id == ORC_HUMTP;
    dims[0] = refined ? s->nzmr : s->nzm;
    dims[1] = refined ? s->nymr : s->nym;
    dims[2] = plane ? 1 : (refined ? s->nxr : s->nx);
    return s->f[id];
}

void orc_scalars(orc_state *s, double out[8]) {
    out[0] = s->ren; out[1] = s->pect; out[2] = s->pecs; out[3] = s->byct; out[4] = s->bycs;
    out[5] = s->dx; out[6] = s->dy; out[7] = s->dz;
}

double *orc_grid(orc_state *s, int id, int64_t *n) {
    int nx = s->nx, nxm = s->nxm;
    switch (id) {
    case ORC_G_XC: *n = nx; return s->xc + 1;
    case ORC_G_XM: *n = nx; return s->xm + 1;
    case ORC_G_G3RC: *n = nx; return s->g3rc + 1;
    case ORC_G_G3RM: *n = nx; return s->g3rm + 1;
    case ORC_G_D3XC: *n = nxm; return s->d3xc + 1;
    case ORC_G_D3XM: *n = nxm; return s->d3xm + 1;
    case ORC_G_UDX3C: *n = nx; return s->udx3c + 1;
    case ORC_G_UDX3M: *n = nxm; return s->udx3m + 1;
    case ORC_G_AP3CK: *n = nx; return s->ap3ck + 1;
    case ORC_G_AC3CK: *n = nx; return s->ac3ck + 1;
    case ORC_G_AM3CK: *n = nx; return s->am3ck + 1;
    case ORC_G_AP3SK: *n = nxm; return s->ap3sk + 1;
    case ORC_G_AC3SK: *n = nxm; return s->ac3sk + 1;
    case ORC_G_AM3SK: *n = nxm; return s->am3sk + 1;
    case ORC_G_AP3SSK: *n = nxm; return s->ap3ssk + 1;
    case ORC_G_AC3SSK: *n = nxm; return s->ac3ssk + 1;
    case ORC_G_AM3SSK: *n = nxm; return s->am3ssk + 1;
    case ORC_G_AK1: *n = s->nzm; return s->ak1 + 1;
    case ORC_G_AK2: *n = s->nym; return s->ak2 + 1;
    case ORC_G_AMPHK: *n = nxm; return s->amphk + 1;
    case ORC_G_ACPHK: *n = nxm; return s->acphk + 1;
    case ORC_G_APPHK: *n = nxm; return s->apphk + 1;
    default: break;
    }
    if (!s->c.salinity && !s->c.phasefield) { *n = 0; return NULL; }
    switch (id) {
    case ORC_G_AP3SPKR: *n = s->nxmr; return s->ap3spkr + 1;
    case ORC_G_AC3SPKR: *n = s->nxmr; return s->ac3spkr + 1;
    case ORC_G_AM3SPKR: *n = s->nxmr; return s->am3spkr + 1;
    case ORC_G_CXPHIC: *n = 4 * s->nxm; return s->cxphic;
    case ORC_G_CYPHIC: *n = 4 * s->nym; return s->cyphic;
    case ORC_G_CZPHIC: *n = 4 * s->nzm; return s->czphic;
    case ORC_G_XCR: *n = s->nxr; return s->xcr + 1;
    case ORC_G_XMR: *n = s->nxr; return s->xmr + 1;
    case ORC_G_UDX3MR: *n = s->nxmr; return s->udx3mr + 1;
    case ORC_G_G3RMR: *n = s->nxr; return s->g3rmr + 1;
    case ORC_G_AP3SSKR: *n = s->nxmr; return s->ap3sskr + 1;
    case ORC_G_AC3SSKR: *n = s->nxmr; return s->ac3sskr + 1;
    case ORC_G_AM3SSKR: *n = s->nxmr; return s->am3sskr + 1;
    case ORC_G_CXRS: *n = 4 * s->nxmr; return s->cxrs;
    case ORC_G_CYRS: *n = 4 * s->nymr; return s->cyrs;
    case ORC_G_CZRS: *n = 4 * s->nzmr; return s->czrs;
    case ORC_G_CXSALC: *n = 4 * s->nx; return s->cxsalc;
    case ORC_G_CYSALC: *n = 4 * s->nym; return s->cysalc;
    case ORC_G_CZSALC: *n = 4 * s->nzm; return s->czsalc;
    default: *n = 0; return NULL;
    }
}

int orc_int_grid(orc_state *s, int id, int64_t *n, int32_t **out) {
    if (!s->c.salinity && !s->c.phasefield) return -1;
    switch (id) {
    case 0: *n = s->nx + 1; *out = s->irangs; return 0;
    case 1: *n = s->ny + 1; *out = s->jrangs; return 0;
    case 2: *n = s->nz + 1; *out = s->krangs; return 0;
    case 3: *n = s->nx + 2; *out = s->irangb; return 0;
    default: return -1;
    }
}

/* ------------------------------------------------------------------ */
/* numerics helpers                                                    */
/* ------------------------------------------------------------------ */
/* Thomas elimination without pivoting.  Row k (0-based):
 *   a[k] x[k-1] + b[k] x[k] + c[k] x[k+1] = d[k];  d is overwritten by x. */
void orc_tridiag_solve(int n, const double *a, const double *b, const double *c, double *d) {
    double *cp = malloc(sizeof(double) * n);
    double piv = b[0];
    cp[0] = c[0] / piv; d[0] = d[0] / piv;
    for (int k = 1; k < n; k++) {
        piv = b[k] - a[k] * cp[k - 1];
        cp[k] = c[k] / piv;
        d[k] = (d[k] - a[k] * d[k - 1]) / piv;
    }
    for (int k = n - 2; k >= 0; k--) d[k] -= cp[k] * d[k + 1];
    free(cp);
}

/* Naive O(n^2) DFT: X_m = sum_j x_j exp(sign * 2 pi i j m / n). */
void orc_dft_naive(const double *re, const double *im, double *ore, double *oim, int n, int sign) {
    const double pi = 2.0 * asin(1.0);
    for (int m = 0; m < n; m++) {
        long double sr = 0, si = 0;
        for (int j = 0; j < n; j++) {
            double ang = sign * 2.0 * pi * (double)(((long long)j * m) % n) / n;
            double cr = cos(ang), ci = sin(ang);
            sr += (long double)re[j] * cr - (long double)im[j] * ci;
            si += (long double)re[j] * ci + (long double)im[j] * cr;
        }
        ore[m] = (double)sr; oim[m] = (double)si;
    }
}

/* Recursive mixed-radix decimation-in-time FFT (Cooley-Tukey), unnormalised,
 * X_m = sum_j x_j exp(sign*2 pi i j m/n).  tw[t] = exp(-2 pi i t/N) for the
 * top-level size N; a level of size nl uses tw[(t mod nl) * N/nl]. */
static void fft_rec(double *re, double *im, int n, int stride, int sign, double *wr, double *wi,
                    const double *twr, const double *twi, int N) {
    if (n == 1) return;
    int p = n;
    for (int f = 2; f * f <= n; f++) if (n % f == 0) { p = f; break; }
    int m = n / p, scale = N / n;
    for (int r = 0; r < p; r++)
        for (int j = 0; j < m; j++) { wr[r * m + j] = re[(j * p + r) * stride]; wi[r * m + j] = im[(j * p + r) * stride]; }
    for (int r = 0; r < p; r++) fft_rec(wr + r * m, wi + r * m, m, 1, sign, wr + n, wi + n, twr, twi, N);
    for (int k = 0; k < m; k++)
        for (int q = 0; q < p; q++) {
            int out = k + q * m;
            double sr = 0, si = 0;
            for (int r = 0; r < p; r++) {
                int t = (int)(((long long)r * out) % n) * scale;
                double cr = twr[t], ci = sign < 0 ? twi[t] : -twi[t];
                double xr = wr[r * m + k], xi = wi[r * m + k];
                sr += xr * cr - xi * ci; si += xr * ci + xi * cr;
            }
            re[out * stride] = sr; im[out * stride] = si;
        }
}

/* twiddle tables cached per size (created once per size, under a lock) */
#define NTWCACHE 256
static struct { int n; double *re, *im; } twcache[NTWCACHE];
static int ntw = 0;
static void get_twiddles(int n, const double **twr, const double **twi) {
    const double pi = 2.0 * asin(1.0);
    for (int c = 0; c < ntw; c++) if (twcache[c].n == n) { *twr = twcache[c].re; *twi = twcache[c].im; return; }
#pragma omp critical(orc_twiddles)
    {
        int found = -1;
        for (int c = 0; c < ntw; c++) if (twcache[c].n == n) found = c;
        if (found < 0 && ntw < NTWCACHE) {
            double *a = malloc(sizeof(double) * n), *b = malloc(sizeof(double) * n);
            for (int t = 0; t < n; t++) { a[t] = cos(2.0 * pi * t / n); b[t] = -sin(2.0 * pi * t / n); }
            twcache[ntw].re = a; twcache[ntw].im = b;
            __atomic_store_n(&twcache[ntw].n, n, __ATOMIC_RELEASE);
            ntw++;
        }
    }
    for (int c = 0; c < ntw; c++) if (twcache[c].n == n) { *twr = twcache[c].re; *twi = twcache[c].im; return; }
    fprintf(stderr, "oracle: twiddle cache full\n"); abort();
}

void orc_fft(double *re, double *im, int n, int sign) {
    const double *twr, *twi;
    get_twiddles(n, &twr, &twi);
    double wr[4 * 4096 + 64], wi[4 * 4096 + 64];
    if (n > 4096) { fprintf(stderr, "oracle: fft size %d too large\n", n); abort(); }
    fft_rec(re, im, n, 1, sign, wr, wi, twr, twi, n);
}

/* ------------------------------------------------------------------ */
/* TimeMarcher state (TimeMarcher.F90:28,38-40)                        */
/* ------------------------------------------------------------------ */
void orc_set_substep(orc_state *s, double dt, int ns) {
    s->dt = dt;
    s->beta = dt / s->ren * 0.5;
    s->al = s->alm[ns - 1]; s->ga = s->gam[ns - 1]; s->ro = s->rom[ns - 1];
}

#define BASE_DECL int nx = s->nx, nxm = s->nxm, nym = s->nym, nzm = s->nzm; (void)nx; (void)nxm; (void)nym; (void)nzm;
#define B(k, j0, i0) ((((size_t)(i0)) * nym + (j0)) * nx + (k) - 1)
#define P2(j0, i0) (((size_t)(i0)) * nym + (j0))

/* ExplicitTermsVX.F90:11-131 */
void orc_explicit_terms_vx(orc_state *s) {
    BASE_DECL
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ], *temp = s->f[ORC_TEMP];
    double *qcap = s->f[ORC_QCAP];
    double udy = s->dy * 0.25, udz = s->dz * 0.25, udyq = s->dyq / s->ren, udzq = s->dzq / s->ren;
    double *idx = malloc(sizeof(double) * (nxm + 1));
    for (int kc = 1; kc <= nxm; kc++) idx[kc] = 1.0 / s->udx3m[kc];
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), im = W(ic - 1, nzm), ip = W(ic + 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jm = W(jc - 1, nym), jp = W(jc + 1, nym);
            for (int kc = 2; kc <= nxm; kc++) {
                int km = kc - 1, kp = kc + 1;
                double hxz = s->udx3c[kc] * udz * (
                    (idx[km] * vz[B(km, j0, ip)] + idx[kc] * vz[B(kc, j0, ip)]) * (vx[B(kc, j0, ip)] + vx[B(kc, j0, i0)])
                  - (idx[km] * vz[B(km, j0, i0)] + idx[kc] * vz[B(kc, j0, i0)]) * (vx[B(kc, j0, i0)] + vx[B(kc, j0, im)]));
                double hxy = s->udx3c[kc] * udy * (
                    (idx[km] * vy[B(km, jp, i0)] + idx[kc] * vy[B(kc, jp, i0)]) * (vx[B(kc, jp, i0)] + vx[B(kc, j0, i0)])
                  - (idx[km] * vy[B(km, j0, i0)] + idx[kc] * vy[B(kc, j0, i0)]) * (vx[B(kc, j0, i0)] + vx[B(kc, jm, i0)]));
                double a = vx[B(kp, j0, i0)] + vx[B(kc, j0, i0)], b = vx[B(kc, j0, i0)] + vx[B(km, j0, i0)];
                double hxx = (a * a - b * b) * s->udx3c[kc] * 0.25;
                double dzzvx = (vx[B(kc, j0, im)] - 2.0 * vx[B(kc, j0, i0)] + vx[B(kc, j0, ip)]) * udzq;
                double dyyvx = (vx[B(kc, jm, i0)] - 2.0 * vx[B(kc, j0, i0)] + vx[B(kc, jp, i0)]) * udyq;
                qcap[B(kc, j0, i0)] = -(hxx + hxy + hxz) + dyyvx + dzzvx;
            }
            if (s->c.gaxis == 1)
                for (int kc = 2; kc <= nxm; kc++) {
                    double tempit = s->c.active_t * (g_plant == 3 ? 1.0 : 0.5) * (temp[B(kc, j0, i0)] + temp[B(kc - 1, j0, i0)]);
                    qcap[B(kc, j0, i0)] += s->byct * tempit;
                }
        }
    }
    free(idx);
}

/* ExplicitTermsVY.F90:11-126 */
void orc_explicit_terms_vy(orc_state *s) {
    BASE_DECL
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ], *temp = s->f[ORC_TEMP];
    double *dph = s->f[ORC_DPH];
    double udyq = s->dyq / s->ren, udzq = s->dzq / s->ren, udy = s->dy * 0.25, udz = s->dz * 0.25;
    double Gy = (s->c.inslwn == 0 || s->c.inslws == 0) ? s->c.dpdy * s->c.dpdy / (s->ren * s->ren)
                                                        : 8.0 * s->c.dpdy * s->c.dpdy / (s->ren * s->ren);
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), imm = W(ic - 1, nzm), ipp = W(ic + 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jmm = W(jc - 1, nym), jpp = W(jc + 1, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                int kmm = s->kmv[kc], kpp = s->kpv[kc], kp = kc + 1;
                double hyx = ((vx[B(kp, j0, i0)] + vx[B(kp, jmm, i0)]) * (vy[B(kpp, j0, i0)] + vy[B(kc, j0, i0)])
                            - (vx[B(kc, j0, i0)] + vx[B(kc, jmm, i0)]) * (vy[B(kc, j0, i0)] + vy[B(kmm, j0, i0)])) * s->udx3m[kc] * 0.25;
                double p = vy[B(kc, jpp, i0)] + vy[B(kc, j0, i0)], m = vy[B(kc, jmm, i0)] + vy[B(kc, j0, i0)];
                double hyy = (p * p - m * m) * udy;
                double hyz = ((vy[B(kc, j0, ipp)] + vy[B(kc, j0, i0)]) * (vz[B(kc, j0, ipp)] + vz[B(kc, jmm, ipp)])
                            - (vy[B(kc, j0, i0)] + vy[B(kc, j0, imm)]) * (vz[B(kc, j0, i0)] + vz[B(kc, jmm, i0)])) * udz;
                double dyyvy = (vy[B(kc, jpp, i0)] - 2.0 * vy[B(kc, j0, i0)] + vy[B(kc, jmm, i0)]) * udyq;
                double dzzvy = (vy[B(kc, j0, ipp)] - 2.0 * vy[B(kc, j0, i0)] + vy[B(kc, j0, imm)]) * udzq;
                dph[B(kc, j0, i0)] = -(hyx + hyy + hyz) + dyyvy + dzzvy + Gy;
            }
            if (s->c.gaxis == 2)
                for (int kc = 1; kc <= nxm; kc++)
                    dph[B(kc, j0, i0)] += s->byct * (s->c.active_t * 0.5 * (temp[B(kc, j0, i0)] + temp[B(kc, jmm, i0)]));
        }
    }
}

/* ExplicitTermsVZ.F90:11-121 */
void orc_explicit_terms_vz(orc_state *s) {
    BASE_DECL
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ], *temp = s->f[ORC_TEMP];
    double *dq = s->f[ORC_DQ];
    double udyq = s->dyq / s->ren, udzq = s->dzq / s->ren, udy = s->dy * 0.25, udz = s->dz * 0.25;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), imm = W(ic - 1, nzm), ipp = W(ic + 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jmm = W(jc - 1, nym), jpp = W(jc + 1, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                int kmm = s->kmv[kc], kpp = s->kpv[kc], kp = kc + 1;
                double p = vz[B(kc, j0, ipp)] + vz[B(kc, j0, i0)], m = vz[B(kc, j0, imm)] + vz[B(kc, j0, i0)];
                double hzz = (p * p - m * m) * udz;
                double hzy = ((vy[B(kc, jpp, i0)] + vy[B(kc, jpp, imm)]) * (vz[B(kc, jpp, i0)] + vz[B(kc, j0, i0)])
                            - (vy[B(kc, j0, i0)] + vy[B(kc, j0, imm)]) * (vz[B(kc, j0, i0)] + vz[B(kc, jmm, i0)])) * udy;
                double hzx = ((vx[B(kp, j0, i0)] + vx[B(kp, j0, imm)]) * (vz[B(kpp, j0, i0)] + vz[B(kc, j0, i0)])
                            - (vx[B(kc, j0, i0)] + vx[B(kc, j0, imm)]) * (vz[B(kc, j0, i0)] + vz[B(kmm, j0, i0)])) * s->udx3m[kc] * 0.25;
                double dzzvz = (vz[B(kc, j0, ipp)] - 2.0 * vz[B(kc, j0, i0)] + vz[B(kc, j0, imm)]) * udzq;
                double dyyvz = (vz[B(kc, jpp, i0)] - 2.0 * vz[B(kc, j0, i0)] + vz[B(kc, jmm, i0)]) * udyq;
                dq[B(kc, j0, i0)] = -(hzx + hzy + hzz) + dyyvz + dzzvz;
            }
            if (s->c.gaxis == 3)
                for (int kc = 1; kc <= nxm; kc++)
                    dq[B(kc, j0, i0)] += s->byct * (s->c.active_t * 0.5 * (temp[B(kc, j0, i0)] + temp[B(kc, j0, imm)]));
        }
    }
}

/* ExplicitTermsTemp.F90:11-85 */
void orc_explicit_terms_temp(orc_state *s) {
    BASE_DECL
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ], *temp = s->f[ORC_TEMP];
    const double *tbp = s->f[ORC_TEMPBP], *ttp = s->f[ORC_TEMPTP];
    double *hro = s->f[ORC_HRO];
    double udz = s->dz * 0.5, udy = s->dy * 0.5, udzq = s->dzq / s->pect, udyq = s->dyq / s->pect;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), im = W(ic - 1, nzm), ip = W(ic + 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jm = W(jc - 1, nym), jp = W(jc + 1, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                double htx;
                if (kc == 1)
                    htx = (vx[B(kc + 1, j0, i0)] * (temp[B(kc + 1, j0, i0)] + temp[B(kc, j0, i0)])
                         - vx[B(kc, j0, i0)] * 2.0 * tbp[P2(j0, i0)]) * 0.5 * s->udx3m[kc];
                else if (kc == nxm)
                    htx = (vx[B(kc + 1, j0, i0)] * 2.0 * ttp[P2(j0, i0)]
                         - vx[B(kc, j0, i0)] * (temp[B(kc, j0, i0)] + temp[B(kc - 1, j0, i0)])) * 0.5 * s->udx3m[kc];
                else
                    htx = (vx[B(kc + 1, j0, i0)] * (temp[B(kc + 1, j0, i0)] + temp[B(kc, j0, i0)])
                         - vx[B(kc, j0, i0)] * (temp[B(kc, j0, i0)] + temp[B(kc - 1, j0, i0)])) * 0.5 * s->udx3m[kc];
                double htz = (vz[B(kc, j0, ip)] * (temp[B(kc, j0, ip)] + temp[B(kc, j0, i0)])
                            - vz[B(kc, j0, i0)] * (temp[B(kc, j0, i0)] + temp[B(kc, j0, im)])) * udz;
                double hty = (vy[B(kc, jp, i0)] * (temp[B(kc, jp, i0)] + temp[B(kc, j0, i0)])
                            - vy[B(kc, j0, i0)] * (temp[B(kc, j0, i0)] + temp[B(kc, jm, i0)])) * udy;
                double dzzt = (temp[B(kc, j0, ip)] - 2.0 * temp[B(kc, j0, i0)] + temp[B(kc, j0, im)]) * udzq;
                double dyyt = (temp[B(kc, jp, i0)] - 2.0 * temp[B(kc, j0, i0)] + temp[B(kc, jm, i0)]) * udyq;
                hro[B(kc, j0, i0)] = -(htx + hty + htz) + dyyt + dzzt;
            }
        }
    }
}

/* Solve the normalised constant-coefficient system column by column:
 * rows r = r0..r1 (1-based) with amk, apk, ackl_b per row; identity rows
 * outside.  Used by SolveImpEqnUpdate_{X,YZ,Temp,Sal}. */
static void solve_columns(orc_state *s, double *q, double *rhs, int lx, int ny, int nzz,
                          int n, const double *am, const double *ac, const double *ap, double betadx,
                          int r0, int r1, int upd0, int upd1) {
    double *a = malloc(sizeof(double) * (n + 1)), *b = malloc(sizeof(double) * (n + 1)), *c = malloc(sizeof(double) * (n + 1));
    double *sc = malloc(sizeof(double) * (n + 1));
    for (int k = 1; k <= n; k++) { a[k - 1] = 0; b[k - 1] = 1; c[k - 1] = 0; sc[k] = 0; }
    for (int k = r0; k <= r1; k++) {
        double ackl_b = 1.0 / (1.0 - ac[k] * betadx);
        a[k - 1] = -am[k] * betadx * ackl_b;
        c[k - 1] = -ap[k] * betadx * ackl_b;
        sc[k] = 1.0 / (1.0 - ac[k] * betadx);
    }
    /* rows outside r0..r1 are the identity with zero rhs (SolveImpEqnUpdate_X.F90:26-28,50-51) */
    apply_threads();
#pragma omp parallel
    {
        double *col = malloc(sizeof(double) * (n + 1));
#pragma omp for
        for (int ic = 0; ic < nzz; ic++)
            for (int jc = 0; jc < ny; jc++) {
                size_t base = ((size_t)ic * ny + jc) * lx;
                for (int k = 1; k <= n; k++) col[k - 1] = (k >= r0 && k <= r1) ? rhs[base + k - 1] * sc[k] : 0.0;
                orc_tridiag_solve(n, a, b, c, col);
                for (int k = 1; k <= n; k++) rhs[base + k - 1] = col[k - 1];
                for (int k = upd0; k <= upd1; k++) q[base + k - 1] += col[k - 1];
            }
        free(col);
    }
    free(a); free(b); free(c); free(sc);
}

/* SolveImpEqnUpdate_X_ibm (src/ibm/SolveImpEqnUpdate_X_ibm.F90:12-86),
 * _YZ_ibm (:12-75), _Temp_ibm (:12-100) and _Sal_ibm (:12-76): one
 * tridiagonal per column whose rows follow the mask: liquid rows are the
 * normalised rows of the plain solve; solid rows the identity with rhs -q
 * (+ Tb for T, Tb = 1); upper / lower boundary points the rows (1, -d) /
 * (-d, 1) with rhs d q(k+1) - q(k) / d q(k-1) - q(k) (+ (1-d) Tb for T),
 * d taken from the list in loop order (ic, jc, kc).  S treats solid rows as
 * liquid and takes d = 1.  vx: rows 1 and nx are the identity with rhs -vx,
 * and the update covers 2..nxm.  slot: 0 vx, 1 vy, 2 vz, 3 T, 4 S. */
static void solve_columns_ibm(orc_state *s, int slot, double *q, const double *rhs, int lx, int ny, int nzz, int n,
                              const double *am, const double *ac, const double *ap, double betadx) {
    const int32_t *mask = s->ibm_mask[slot];
    const double *dist = s->ibm_dist[slot];
    const double tb = 1.0;
    double *a = malloc(sizeof(double) * (lx + 1)), *b = malloc(sizeof(double) * (lx + 1));
    double *c = malloc(sizeof(double) * (lx + 1)), *f = malloc(sizeof(double) * (lx + 1));
    int64_t nd = 0;
    const int nsolve = slot == 0 ? lx : n;          /* vx: nx rows; others nxm (nxmr) */
    for (int ic = 0; ic < nzz; ic++)
        for (int jc = 0; jc < ny; jc++) {
            size_t base = ((size_t)ic * ny + jc) * lx;
            for (int k = 0; k < nsolve; k++) { a[k] = 0.0; b[k] = 1.0; c[k] = 0.0; f[k] = 0.0; }
            for (int kc = 1; kc <= n; kc++) {
                int km = kc - 1 < 1 ? 1 : kc - 1;
                int kp = slot == 0 ? kc + 1 : (kc + 1 > n ? n : kc + 1);
                int m = mask[base + kc - 1];
                double qk = q[base + kc - 1];
                if (slot == 4 ? (m % 2 == 0) : (m == 2)) {
                    double ackl_b = 1.0 / (1.0 - ac[kc] * betadx);
                    a[kc - 1] = -am[kc] * betadx * ackl_b;
                    c[kc - 1] = -ap[kc] * betadx * ackl_b;
                    f[kc - 1] = rhs[base + kc - 1] * ackl_b;
                } else if (m == 0) {
                    f[kc - 1] = -qk + (slot == 3 ? tb : 0.0);
                } else if (m == 1 || m == -1) {
                    double d = slot == 4 ? 1.0 : dist[nd];
                    if (slot != 4 && nd >= s->ibm_ndist[slot]) d = 0.0;      /* list too short: caught by orc_set_ibm */
                    double qn = q[base + (m == 1 ? kp : km) - 1];
                    if (m == 1) c[kc - 1] = -d; else a[kc - 1] = -d;
                    f[kc - 1] = d * qn - qk + (slot == 3 ? (1.0 - d) * tb : 0.0);
                    nd++;
                }
            }
            if (slot == 0) {          /* identity wall rows, SolveImpEqnUpdate_X_ibm.F90:66-70 */
                a[0] = 0.0; b[0] = 1.0; c[0] = 0.0; f[0] = -q[base];
                a[lx - 1] = 0.0; c[lx - 1] = 0.0; f[lx - 1] = -q[base + lx - 1];
            }
            a[0] = 0.0; c[nsolve - 1] = 0.0;     /* dgttrf takes the sub- / super-diagonals of rows 2.. / ..n-1 */
            orc_tridiag_solve(nsolve, a, b, c, f);
            int u0 = slot == 0 ? 2 : 1, u1 = n;
            for (int kc = u0; kc <= u1; kc++) q[base + kc - 1] += f[kc - 1];
        }
    free(a); free(b); free(c); free(f);
}

int orc_set_ibm(orc_state *s, int field, const int32_t *mask, const double *dist, int64_t ndist) {
    int slot = field == ORC_VX ? 0 : field == ORC_VY ? 1 : field == ORC_VZ ? 2 : field == ORC_TEMP ? 3 : field == ORC_SAL ? 4 : -1;
    if (slot < 0 || (slot == 4 && !s->c.salinity)) return 1;
    size_t n = slot == 4 ? nref(s) : nbase(s);
    int64_t nb = 0;
    for (size_t q = 0; q < n; q++) if (mask[q] == 1 || mask[q] == -1) nb++;
    if (slot != 4 && ndist < nb) return 2;
    s->ibm_mask[slot] = track(s, sizeof(int32_t) * n);
    memcpy(s->ibm_mask[slot], mask, sizeof(int32_t) * n);
    s->ibm_dist[slot] = track(s, sizeof(double) * (size_t)(ndist > 0 ? ndist : 1));
    if (ndist > 0) memcpy(s->ibm_dist[slot], dist, sizeof(double) * (size_t)ndist);
    s->ibm_ndist[slot] = ndist;
    return 0;
}

/* SolveImpEqnUpdate_X_pf (phasefield.F90:821-872) and SolveImpEqnUpdate_YZ_pf
 * (:875-930): one tridiagonal per column, in which a row whose interpolated
 * phase field exceeds pf_direct_force = 0.9 (phasefield.F90:29) becomes the
 * identity with right-hand side -q, so that q is forced to zero there.
 * axis 'x': rows 1 and nx are identity rows with rhs -vx, rows 2..nxm test
 * 0.5 (phic(kc) + phic(kc+1)); axis 'y'/'z': rows 1..nxm test the average
 * of phic with its y / z neighbour below.  q(upd0..upd1) += solution; rhs is
 * left as assembled. */
static void solve_columns_pf(orc_state *s, double *q, const double *rhs, char axis,
                             const double *am, const double *ac, const double *ap, double betadx) {
    BASE_DECL
    const double *phic = s->f[ORC_PHIC];
    const int n = axis == 'x' ? nx : nxm;
    apply_threads();
#pragma omp parallel
    {
        double *a = malloc(sizeof(double) * (n + 1)), *b = malloc(sizeof(double) * (n + 1));
        double *c = malloc(sizeof(double) * (n + 1)), *f = malloc(sizeof(double) * (n + 1));
#pragma omp for
        for (int ic = 1; ic <= nzm; ic++) {
            int i0 = W(ic, nzm), imm = W(ic - 1, nzm);
            for (int jc = 1; jc <= nym; jc++) {
                int j0 = W(jc, nym), jmm = W(jc - 1, nym);
                for (int k = 0; k < n; k++) { a[k] = 0.0; b[k] = 1.0; c[k] = 0.0; }
                int k0 = 1, k1 = n;
                if (axis == 'x') {
                    f[0] = -q[B(1, j0, i0)];
                    f[nx - 1] = -q[B(nx, j0, i0)];
                    k0 = 2; k1 = nxm;
                }
                for (int kc = k0; kc <= k1; kc++) {
                    double philoc;
                    if (axis == 'x') philoc = 0.5 * (phic[B(kc, j0, i0)] + phic[B(kc + 1, j0, i0)]);
                    else if (axis == 'y') philoc = 0.5 * (phic[B(kc, j0, i0)] + phic[B(kc, jmm, i0)]);
                    else philoc = 0.5 * (phic[B(kc, j0, i0)] + phic[B(kc, j0, imm)]);
                    if (philoc > 0.9) {
                        f[kc - 1] = -q[B(kc, j0, i0)];
                    } else {
                        double ackl_b = 1.0 / (1.0 - ac[kc] * betadx);
                        if (kc > 1) a[kc - 1] = -am[kc] * betadx * ackl_b;
                        if (kc < n) c[kc - 1] = -ap[kc] * betadx * ackl_b;
                        f[kc - 1] = rhs[B(kc, j0, i0)] * ackl_b;
                    }
                }
                orc_tridiag_solve(n, a, b, c, f);
                for (int kc = k0; kc <= k1; kc++) q[B(kc, j0, i0)] += f[kc - 1];
            }
        }
        free(a); free(b); free(c); free(f);
    }
}

/* ImplicitAndUpdateVX.F90:14-89 + SolveImpEqnUpdate_X.F90:12-69 */
void orc_implicit_and_update_vx(orc_state *s) {
    BASE_DECL
    double *vx = s->f[ORC_VX], *rhs = s->f[ORC_RHS], *rux = s->f[ORC_RUX];
    const double *qcap = s->f[ORC_QCAP], *pr = s->f[ORC_PR];
    double alre = s->al / s->ren;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym);
            for (int kc = 2; kc <= nxm; kc++) {
                int km = kc - 1, kp = kc + 1;
                double udx3 = s->al * s->udx3c[kc];
                double dxxvx = vx[B(kp, j0, i0)] * s->ap3ck[kc] + vx[B(kc, j0, i0)] * s->ac3ck[kc] + vx[B(km, j0, i0)] * s->am3ck[kc];
                double dxp = (pr[B(kc, j0, i0)] - pr[B(km, j0, i0)]) * udx3;
                if (g_plant == 1) dxp = -dxp;
                rhs[B(kc, j0, i0)] = (s->ga * qcap[B(kc, j0, i0)] + s->ro * rux[B(kc, j0, i0)] + alre * dxxvx - dxp) * s->dt;
                rux[B(kc, j0, i0)] = qcap[B(kc, j0, i0)];
            }
        }
    }
    if (s->ibm_mask[0]) solve_columns_ibm(s, 0, vx, rhs, nx, nym, nzm, nxm, s->am3ck, s->ac3ck, s->ap3ck, s->beta * s->al);
    else if (s->c.phasefield) solve_columns_pf(s, vx, rhs, 'x', s->am3ck, s->ac3ck, s->ap3ck, s->beta * s->al);
    else solve_columns(s, vx, rhs, nx, nym, nzm, nx, s->am3ck, s->ac3ck, s->ap3ck, s->beta * s->al, 2, nxm, 2, nxm);
    for (size_t c = 0; c < (size_t)nym * nzm; c++) { vx[c * nx + 0] = 0.0; vx[c * nx + nx - 1] = 0.0; }
}

/* ImplicitAndUpdateVY.F90:14-95 + SolveImpEqnUpdate_YZ.F90:12-60 */
void orc_implicit_and_update_vy(orc_state *s) {
    BASE_DECL
    double *vy = s->f[ORC_VY], *rhs = s->f[ORC_RHS], *ruy = s->f[ORC_RUY];
    const double *dph = s->f[ORC_DPH], *pr = s->f[ORC_PR];
    double alre = s->al / s->ren, udy = s->dy * s->al;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jmm = W(jc - 1, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                int kmm = s->kmv[kc], kpp = s->kpv[kc];
                double amm = s->am3sk[kc], acc = s->ac3sk[kc], app = s->ap3sk[kc], dxxvy;
                if (kc == 1)
                    dxxvy = vy[B(kpp, j0, i0)] * app + vy[B(kc, j0, i0)] * acc - (acc + app) * s->c.xminusu * s->c.inslws;
                else if (kc == nxm)
                    dxxvy = -(acc + amm) * s->c.xplusu * s->c.inslwn + vy[B(kc, j0, i0)] * acc + vy[B(kmm, j0, i0)] * amm;
                else
                    dxxvy = vy[B(kpp, j0, i0)] * app + vy[B(kc, j0, i0)] * acc + vy[B(kmm, j0, i0)] * amm;
                double dyp = (pr[B(kc, j0, i0)] - pr[B(kc, jmm, i0)]) * udy;
                rhs[B(kc, j0, i0)] = (s->ga * dph[B(kc, j0, i0)] + s->ro * ruy[B(kc, j0, i0)] + alre * dxxvy - dyp) * s->dt;
                ruy[B(kc, j0, i0)] = dph[B(kc, j0, i0)];
            }
        }
    }
    if (s->ibm_mask[1]) solve_columns_ibm(s, 1, vy, rhs, nx, nym, nzm, nxm, s->am3sk, s->ac3sk, s->ap3sk, s->beta * s->al);
    else if (s->c.phasefield) solve_columns_pf(s, vy, rhs, 'y', s->am3sk, s->ac3sk, s->ap3sk, s->beta * s->al);
    else solve_columns(s, vy, rhs, nx, nym, nzm, nxm, s->am3sk, s->ac3sk, s->ap3sk, s->beta * s->al, 1, nxm, 1, nxm);
}

/* ImplicitAndUpdateVZ.F90:14-80 + SolveImpEqnUpdate_YZ.F90:12-60 */
void orc_implicit_and_update_vz(orc_state *s) {
    BASE_DECL
    double *vz = s->f[ORC_VZ], *rhs = s->f[ORC_RHS], *ruz = s->f[ORC_RUZ];
    const double *dq = s->f[ORC_DQ], *pr = s->f[ORC_PR];
    double alre = (g_plant == 6 ? 1.0 : s->al) / s->ren;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), imm = W(ic - 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                int kmm = s->kmv[kc], kpp = s->kpv[kc];
                double dxxvz = vz[B(kpp, j0, i0)] * s->ap3sk[kc] + vz[B(kc, j0, i0)] * s->ac3sk[kc] + vz[B(kmm, j0, i0)] * s->am3sk[kc];
                double dzp = (pr[B(kc, j0, i0)] - pr[B(kc, j0, imm)]) * s->dz * s->al;
                rhs[B(kc, j0, i0)] = (s->ga * dq[B(kc, j0, i0)] + s->ro * ruz[B(kc, j0, i0)] + alre * dxxvz - dzp) * s->dt;
                ruz[B(kc, j0, i0)] = dq[B(kc, j0, i0)];
            }
        }
    }
    if (s->ibm_mask[2]) solve_columns_ibm(s, 2, vz, rhs, nx, nym, nzm, nxm, s->am3sk, s->ac3sk, s->ap3sk, s->beta * s->al);
    else if (s->c.phasefield) solve_columns_pf(s, vz, rhs, 'z', s->am3sk, s->ac3sk, s->ap3sk, s->beta * s->al);
    else solve_columns(s, vz, rhs, nx, nym, nzm, nxm, s->am3sk, s->ac3sk, s->ap3sk, s->beta * s->al, 1, nxm, 1, nxm);
}

/* ImplicitAndUpdateTemp.F90:13-77 + SolveImpEqnUpdate_Temp.F90:11-63 */
void orc_implicit_and_update_temp(orc_state *s) {
    BASE_DECL
    double *temp = s->f[ORC_TEMP], *rhs = s->f[ORC_RHS], *rut = s->f[ORC_RUTEMP];
    const double *hro = s->f[ORC_HRO], *tbp = s->f[ORC_TEMPBP], *ttp = s->f[ORC_TEMPTP];
    double alpec = s->al / s->pect;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                double dxxt;
                if (kc == 1)
                    dxxt = temp[B(kc + 1, j0, i0)] * s->ap3ssk[kc] + temp[B(kc, j0, i0)] * s->ac3ssk[kc]
                         - (s->ap3ssk[kc] + s->ac3ssk[kc]) * tbp[P2(j0, i0)] * s->c.tfixs;
                else if (kc == nxm)
                    dxxt = temp[B(kc, j0, i0)] * s->ac3ssk[kc] + temp[B(kc - 1, j0, i0)] * s->am3ssk[kc]
                         - (s->am3ssk[kc] + s->ac3ssk[kc]) * ttp[P2(j0, i0)] * s->c.tfixn;
                else
                    dxxt = temp[B(kc + 1, j0, i0)] * s->ap3ssk[kc] + temp[B(kc, j0, i0)] * s->ac3ssk[kc]
                         + temp[B(kc - 1, j0, i0)] * s->am3ssk[kc];
                rhs[B(kc, j0, i0)] = (s->ga * hro[B(kc, j0, i0)] + s->ro * rut[B(kc, j0, i0)] + alpec * dxxt) * s->dt;
                rut[B(kc, j0, i0)] = hro[B(kc, j0, i0)];
            }
        }
    }
    if (s->ibm_mask[3] && !s->c.phasefield)      /* ImplicitAndUpdateTemp.F90:69-70 */
        solve_columns_ibm(s, 3, temp, rhs, nx, nym, nzm, nxm, s->am3ssk, s->ac3ssk, s->ap3ssk, 0.5 * s->al * s->dt / s->pect);
    else
        solve_columns(s, temp, rhs, nx, nym, nzm, nxm, s->am3ssk, s->ac3ssk, s->ap3ssk, 0.5 * s->al * s->dt / s->pect, 1, nxm, 1, nxm);
}

/* CalcLocalDivergence, pressure.F90:141-162 */
void orc_calc_local_divergence(orc_state *s) {
    BASE_DECL
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ];
    double *dph = s->f[ORC_DPH];
    double usdtal = 1.0 / (s->dt * s->al);
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), ip = W(ic + 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jp = W(jc + 1, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                double dqcap = (vz[B(kc, j0, ip)] - vz[B(kc, j0, i0)]) * s->dz
                             + (vy[B(kc, jp, i0)] - vy[B(kc, j0, i0)]) * s->dy
                             + (vx[B(kc + 1, j0, i0)] - vx[B(kc, j0, i0)]) * s->udx3m[kc];
                dph[B(kc, j0, i0)] = dqcap * usdtal;
            }
        }
    }
}

/* SolvePressureCorrection + SolveTridiagonalPressure, pressure.F90:165-236 */
void orc_solve_pressure_correction(orc_state *s) {
    BASE_DECL
    double *dph = s->f[ORC_DPH];
    int nymh = nym / 2 + 1;
    size_t nspec = (size_t)nxm * nymh * nzm;
    double *cr = malloc(sizeof(double) * nspec), *ci = malloc(sizeof(double) * nspec);
#define S3(k, m, i) ((((size_t)(i)) * nymh + (m)) * nxm + (k))
    apply_threads();
    /* forward y r2c (pressure.F90:174) */
#pragma omp parallel
    {
        double *lr = malloc(sizeof(double) * (nym > nzm ? nym : nzm)), *li = malloc(sizeof(double) * (nym > nzm ? nym : nzm));
#pragma omp for
        for (int i = 0; i < nzm; i++)
            for (int k = 1; k <= nxm; k++) {
                for (int j = 0; j < nym; j++) { lr[j] = dph[B(k, j, i)]; li[j] = 0.0; }
                orc_fft(lr, li, nym, -1);
                for (int m = 0; m < nymh; m++) { cr[S3(k - 1, m, i)] = lr[m]; ci[S3(k - 1, m, i)] = li[m]; }
            }
        /* forward z c2c + normalisation (pressure.F90:176-179) */
#pragma omp for
        for (int m = 0; m < nymh; m++)
            for (int k = 0; k < nxm; k++) {
                for (int i = 0; i < nzm; i++) { lr[i] = cr[S3(k, m, i)]; li[i] = ci[S3(k, m, i)]; }
                orc_fft(lr, li, nzm, -1);
                for (int i = 0; i < nzm; i++) {
                    cr[S3(k, m, i)] = lr[i] / (double)(nzm * nym);
                    ci[S3(k, m, i)] = li[i] / (double)(nzm * nym);
                }
            }
        free(lr); free(li);
    }
    /* per-mode tridiagonal (pressure.F90:209-233) */
#pragma omp parallel
    {
        double *a = malloc(sizeof(double) * nxm), *b = malloc(sizeof(double) * nxm), *c = malloc(sizeof(double) * nxm);
        double *dr = malloc(sizeof(double) * nxm), *di = malloc(sizeof(double) * nxm);
#pragma omp for
        for (int i = 1; i <= nzm; i++)
            for (int j = 1; j <= nymh; j++) {
                for (int k = 1; k <= nxm; k++) {
                    double sb = 1.0 / (s->acphk[k] - s->ak2[j] - s->ak1[i]);
                    dr[k - 1] = cr[S3(k - 1, j - 1, i - 1)] * sb;
                    di[k - 1] = ci[S3(k - 1, j - 1, i - 1)] * sb;
                    c[k - 1] = k < nxm ? s->apphk[k] * sb : 0.0;
                    a[k - 1] = k > 1 ? s->amphk[k] * sb : 0.0;
                    b[k - 1] = 1.0 + 1.0e-15;
                }
                if (i == 1 && j == 1) { b[0] = 1.0; c[0] = 0.0; dr[0] = 0.0; di[0] = 0.0; }  /* gauge */
                orc_tridiag_solve(nxm, a, b, c, dr);
                orc_tridiag_solve(nxm, a, b, c, di);
                for (int k = 1; k <= nxm; k++) { cr[S3(k - 1, j - 1, i - 1)] = dr[k - 1]; ci[S3(k - 1, j - 1, i - 1)] = di[k - 1]; }
            }
        free(a); free(b); free(c); free(dr); free(di);
    }
    /* inverse z c2c, inverse y c2r (pressure.F90:189-195), unnormalised */
#pragma omp parallel
    {
        int L = nym > nzm ? nym : nzm;
        double *lr = malloc(sizeof(double) * L), *li = malloc(sizeof(double) * L);
#pragma omp for
        for (int m = 0; m < nymh; m++)
            for (int k = 0; k < nxm; k++) {
                for (int i = 0; i < nzm; i++) { lr[i] = cr[S3(k, m, i)]; li[i] = ci[S3(k, m, i)]; }
                orc_fft(lr, li, nzm, +1);
                for (int i = 0; i < nzm; i++) { cr[S3(k, m, i)] = lr[i]; ci[S3(k, m, i)] = li[i]; }
            }
#pragma omp for
        for (int i = 0; i < nzm; i++)
            for (int k = 1; k <= nxm; k++) {
                for (int m = 0; m < nymh; m++) { lr[m] = cr[S3(k - 1, m, i)]; li[m] = ci[S3(k - 1, m, i)]; }
                li[0] = 0.0;
                if (nym % 2 == 0) li[nym / 2] = 0.0;
                for (int m = nymh; m < nym; m++) { lr[m] = lr[nym - m]; li[m] = -li[nym - m]; }
                orc_fft(lr, li, nym, +1);
                for (int j = 0; j < nym; j++) dph[B(k, j, i)] = lr[j];
            }
        free(lr); free(li);
    }
#undef S3
    free(cr); free(ci);
}

/* RemoveDivergence (pressure.F90:240-267), called by CorrectVelocity.F90:24 */
void orc_correct_velocity(orc_state *s) {
    BASE_DECL
    double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ];
    const double *dph = s->f[ORC_DPH];
    double udy = s->al * s->dt * s->dy, udz = s->al * s->dt * s->dz;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), im = W(ic - 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jm = W(jc - 1, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                int km = s->kmv[kc];
                double usukm = s->al * s->dt * s->udx3c[kc];
                double locdph = dph[B(kc, j0, i0)];
                vx[B(kc, j0, i0)] -= (locdph - dph[B(km, j0, i0)]) * usukm;
                vy[B(kc, j0, i0)] -= (locdph - dph[B(kc, jm, i0)]) * udy;
                vz[B(kc, j0, i0)] -= (locdph - dph[B(kc, j0, im)]) * udz;
            }
        }
    }
}

/* CorrectPressure, pressure.F90:270-301 */
void orc_correct_pressure(orc_state *s) {
    BASE_DECL
    double *pr = s->f[ORC_PR];
    const double *d = s->f[ORC_DPH];
    double be = g_plant == 2 ? 0.0 : s->al * s->beta;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), im = W(ic - 1, nzm), ip = W(ic + 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jm = W(jc - 1, nym), jp = W(jc + 1, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                int kp = s->kpv[kc], km = s->kmv[kc];
                double c0 = d[B(kc, j0, i0)];
                pr[B(kc, j0, i0)] += c0 - be * (
                    (d[B(kc, j0, ip)] - 2.0 * c0 + d[B(kc, j0, im)]) * s->dzq +
                    (d[B(kc, jp, i0)] - 2.0 * c0 + d[B(kc, jm, i0)]) * s->dyq +
                    (d[B(kp, j0, i0)] * s->apphk[kc] + c0 * s->acphk[kc] + d[B(km, j0, i0)] * s->amphk[kc]));
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* multiresolution + salinity                                          */
/* ------------------------------------------------------------------ */
#define REF_DECL int nxr = s->nxr, nxmr = s->nxmr, nymr = s->nymr, nzmr = s->nzmr; (void)nxr; (void)nxmr; (void)nymr; (void)nzmr;
#define R(k, j0, i0) ((((size_t)(i0)) * nymr + (j0)) * nxr + (k) - 1)
#define RP2(j0, i0) (((size_t)(i0)) * nymr + (j0))
#define TP(k, j0, i0) ((((size_t)(i0)) * nym + (j0)) * (nx + 3) + (k) + 1)      /* tpdv  x -1..nx+1 */
#define TR(k, j0, i0) ((((size_t)(i0)) * nymr + (j0)) * (nxr + 3) + (k) + 1)   /* tpdvr x -1..nxr+1 */

/* interpolate_xyz_to_refined, HermiteInterpolations.F90:182-214; cvar = tpdv
 * (x -1..nx+1, periodic y/z), rvar = tpdvr(1:nxmr,...) */
static void interp_to_refined(orc_state *s, const double *cvar, double *rvar) {
    BASE_DECL REF_DECL
    apply_threads();
#pragma omp parallel for
    for (int ic = 0; ic <= nzm; ic++) {
        int i4[4];
        for (int c = 0; c < 4; c++) i4[c] = W(ic - 1 + c, nzm);
        for (int jc = 0; jc <= nym; jc++) {
            int j4[4];
            for (int b = 0; b < 4; b++) j4[b] = W(jc - 1 + b, nym);
            for (int kc = 0; kc <= nxm; kc++) {
                double qv3[4][4][4];  /* [c][b][a] */
                for (int c = 0; c < 4; c++)
                    for (int b = 0; b < 4; b++)
                        for (int a = 0; a < 4; a++) qv3[c][b][a] = cvar[TP(kc - 1 + a, j4[b], i4[c])];
                int ilo = s->krangs[ic] > 1 ? s->krangs[ic] : 1, ihi = s->krangs[ic + 1] - 1 < nzmr ? s->krangs[ic + 1] - 1 : nzmr;
                for (int icr = ilo; icr <= ihi; icr++) {
                    const double *cz = s->czrs + (size_t)(icr - 1) * 4;
                    double qv2[4][4];
                    for (int b = 0; b < 4; b++)
                        for (int a = 0; a < 4; a++)
                            qv2[b][a] = qv3[0][b][a] * cz[0] + qv3[1][b][a] * cz[1] + qv3[2][b][a] * cz[2] + qv3[3][b][a] * cz[3];
                    int jlo = s->jrangs[jc] > 1 ? s->jrangs[jc] : 1, jhi = s->jrangs[jc + 1] - 1 < nymr ? s->jrangs[jc + 1] - 1 : nymr;
                    for (int jcr = jlo; jcr <= jhi; jcr++) {
                        const double *cy = s->cyrs + (size_t)(jcr - 1) * 4;
                        double qv1[4];
                        for (int a = 0; a < 4; a++) qv1[a] = qv2[0][a] * cy[0] + qv2[1][a] * cy[1] + qv2[2][a] * cy[2] + qv2[3][a] * cy[3];
                        int klo = s->irangs[kc] > 1 ? s->irangs[kc] : 1, khi = s->irangs[kc + 1] - 1 < nxmr ? s->irangs[kc + 1] - 1 : nxmr;
                        for (int kcr = klo; kcr <= khi; kcr++) {
                            const double *cx = s->cxrs + (size_t)(kcr - 1) * 4;
                            rvar[TR(kcr, jcr - 1, icr - 1)] = qv1[0] * cx[0] + qv1[1] * cx[1] + qv1[2] * cx[2] + qv1[3] * cx[3];
                        }
                    }
                }
            }
        }
    }
}

/* InterpVelMgrd.F90:1-243 */
void orc_interp_vel_mgrd(orc_state *s) {
    BASE_DECL REF_DECL
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ];
    double *vxr = s->f[ORC_VXR], *vyr = s->f[ORC_VYR], *vzr = s->f[ORC_VZR];
    double *tpdv = s->tpdv, *tpdvr = s->tpdvr;
    int sS = 1 - 2 * s->c.inslws, sN = 1 - 2 * s->c.inslwn;
    size_t ntp = (size_t)(nx + 3) * nym * nzm, ntr = (size_t)(nxr + 3) * nymr * nzmr;
    memset(vxr, 0, sizeof(double) * nref(s)); memset(vyr, 0, sizeof(double) * nref(s)); memset(vzr, 0, sizeof(double) * nref(s));

    /* vx: gradient at cell centres (:37-51), interpolate (:55), integrate in x (:58-66) */
    memset(tpdv, 0, sizeof(double) * ntp); memset(tpdvr, 0, sizeof(double) * ntr);
    for (int i = 0; i < nzm; i++)
        for (int j = 0; j < nym; j++) {
            for (int kc = 1; kc <= nxm; kc++) tpdv[TP(kc, j, i)] = (vx[B(kc + 1, j, i)] - vx[B(kc, j, i)]) * s->udx3m[kc];
            tpdv[TP(0, j, i)] = sS * tpdv[TP(1, j, i)];
            tpdv[TP(nx, j, i)] = sN * tpdv[TP(nxm, j, i)];
        }
    interp_to_refined(s, tpdv, tpdvr);
    for (int i = 0; i < nzmr; i++)
        for (int j = 0; j < nymr; j++) {
            vxr[R(1, j, i)] = 0.0;
            for (int kcr = 1; kcr <= nxmr; kcr++) {
                double lxa = s->xcr[kcr + 1] - s->xcr[kcr];
                vxr[R(kcr + 1, j, i)] = vxr[R(kcr, j, i)] + tpdvr[TR(kcr, j, i)] * lxa;
            }
        }

    /* vy: gradient (:75-86), interpolate (:90), first-guess plane j=1 (:93-123), integrate in y (:128-135) */
    memset(tpdv, 0, sizeof(double) * ntp); memset(tpdvr, 0, sizeof(double) * ntr);
    for (int i = 0; i < nzm; i++)
        for (int jc = 1; jc <= nym; jc++) {
            int j = W(jc, nym), jp = W(jc + 1, nym);
            for (int kc = 1; kc <= nxm; kc++) tpdv[TP(kc, j, i)] = (vy[B(kc, jp, i)] - vy[B(kc, j, i)]) * s->dy;
            tpdv[TP(0, j, i)] = sS * tpdv[TP(1, j, i)];
            tpdv[TP(nx, j, i)] = sN * tpdv[TP(nxm, j, i)];
        }
    interp_to_refined(s, tpdv, tpdvr);
    {
        /* vyxzc(-1:nxm+2, z) with wall ghosts (:97-106) */
        int lx = nxm + 4;
        double *vyxzc = calloc((size_t)lx * nzm, sizeof(double));
#define VC(k, i0) vyxzc[(size_t)(i0) * lx + (k) + 1]
        for (int i = 0; i < nzm; i++) {
            for (int kc = 1; kc <= nxm; kc++) VC(kc, i) = vy[B(kc, 0, i)];
            VC(0, i) = sS * VC(1, i);
            VC(nx, i) = sN * VC(nxm, i);
        }
        for (int ic = 0; ic <= nzm; ic++) {
            int i4[4];
            for (int c = 0; c < 4; c++) i4[c] = W(ic - 1 + c, nzm);
            for (int kc = 0; kc <= nxm; kc++) {
                double qv2[4][4];   /* [c][a] */
                for (int c = 0; c < 4; c++) for (int a = 0; a < 4; a++) qv2[c][a] = VC(kc - 1 + a, i4[c]);
                int ilo = s->krangs[ic] > 1 ? s->krangs[ic] : 1, ihi = s->krangs[ic + 1] - 1 < nzmr ? s->krangs[ic + 1] - 1 : nzmr;
                for (int icr = ilo; icr <= ihi; icr++) {
                    const double *cz = s->czrs + (size_t)(icr - 1) * 4;   /* czvy = czvx (CreateMgrdStencil.F90:65) */
                    double qv1[4];
                    for (int a = 0; a < 4; a++) qv1[a] = qv2[0][a] * cz[0] + qv2[1][a] * cz[1] + qv2[2][a] * cz[2] + qv2[3][a] * cz[3];
                    int klo = s->irangs[kc] > 1 ? s->irangs[kc] : 1, khi = s->irangs[kc + 1] - 1 < nxmr ? s->irangs[kc + 1] - 1 : nxmr;
                    for (int kcr = klo; kcr <= khi; kcr++) {
                        const double *cx = s->cxrs + (size_t)(kcr - 1) * 4;   /* cxvy = cxrs */
                        vyr[R(kcr, 0, icr - 1)] = qv1[0] * cx[0] + qv1[1] * cx[1] + qv1[2] * cx[2] + qv1[3] * cx[3];
                    }
                }
            }
        }
#undef VC
        free(vyxzc);
    }
    {
        double lya = 1.0 / s->dyr;
        /* integrate along y; the value carried past nymr lands in the halo and is discarded */
        for (int i = 0; i < nzmr; i++)
            for (int jcr = 1; jcr < nymr; jcr++)
                for (int kcr = 1; kcr <= nxmr; kcr++)
                    vyr[R(kcr, jcr, i)] = vyr[R(kcr, jcr - 1, i)] + tpdvr[TR(kcr, jcr - 1, i)] * lya;
    }

    /* vz: gradient (:161-172), interpolate (:176), first-guess plane i=1 (:179-209), integrate in z (:215-221) */
    memset(tpdv, 0, sizeof(double) * ntp); memset(tpdvr, 0, sizeof(double) * ntr);
    for (int ic = 1; ic <= nzm; ic++) {
        int i = W(ic, nzm), ip = W(ic + 1, nzm);
        for (int j = 0; j < nym; j++) {
            for (int kc = 1; kc <= nxm; kc++) tpdv[TP(kc, j, i)] = (vz[B(kc, j, ip)] - vz[B(kc, j, i)]) * s->dz;
            tpdv[TP(0, j, i)] = sS * tpdv[TP(1, j, i)];
            tpdv[TP(nx, j, i)] = sN * tpdv[TP(nxm, j, i)];
        }
    }
    interp_to_refined(s, tpdv, tpdvr);
    if (1 < nzm) {   /* guard ic0 < xend(3) at one rank (InterpVelMgrd.F90:184) */
        int lx = nxm + 4;
        double *vzxyc = calloc((size_t)lx * nym, sizeof(double));
#define VC(k, j0) vzxyc[(size_t)(j0) * lx + (k) + 1]
        for (int j = 0; j < nym; j++) {
            for (int kc = 1; kc <= nxm; kc++) VC(kc, j) = vz[B(kc, j, 0)];
            VC(0, j) = sS * VC(1, j);
            VC(nx, j) = sN * VC(nxm, j);
        }
        for (int jc = 0; jc <= nym; jc++) {
            int j4[4];
            for (int b = 0; b < 4; b++) j4[b] = W(jc - 1 + b, nym);
            for (int kc = 0; kc <= nxm; kc++) {
                double qv2[4][4];   /* [b][a] */
                for (int b = 0; b < 4; b++) for (int a = 0; a < 4; a++) qv2[b][a] = VC(kc - 1 + a, j4[b]);
                int jlo = s->jrangs[jc] > 1 ? s->jrangs[jc] : 1, jhi = s->jrangs[jc + 1] - 1 < nymr ? s->jrangs[jc + 1] - 1 : nymr;
                for (int jcr = jlo; jcr <= jhi; jcr++) {
                    const double *cy = s->cyrs + (size_t)(jcr - 1) * 4;   /* cyvz = cyvx = cyrs */
                    double qv1[4];
                    for (int a = 0; a < 4; a++) qv1[a] = qv2[0][a] * cy[0] + qv2[1][a] * cy[1] + qv2[2][a] * cy[2] + qv2[3][a] * cy[3];
                    int klo = s->irangs[kc] > 1 ? s->irangs[kc] : 1, khi = s->irangs[kc + 1] - 1 < nxmr ? s->irangs[kc + 1] - 1 : nxmr;
                    for (int kcr = klo; kcr <= khi; kcr++) {
                        const double *cx = s->cxrs + (size_t)(kcr - 1) * 4;
                        vzr[R(kcr, jcr - 1, 0)] = qv1[0] * cx[0] + qv1[1] * cx[1] + qv1[2] * cx[2] + qv1[3] * cx[3];
                    }
                }
            }
        }
#undef VC
        free(vzxyc);
    }
    {
        double lza = 1.0 / s->dzr;
        for (int icr = 1; icr < nzmr; icr++)
            for (int j = 0; j < nymr; j++)
                for (int kcr = 1; kcr <= nxmr; kcr++)
                    vzr[R(kcr, j, icr)] = vzr[R(kcr, j, icr - 1)] + tpdvr[TR(kcr, j, icr - 1)] * lza;
    }
}

/* CalcSalStats + CalcDissipationSal, salinity.F90:510-577 and :611-690 (non-IBM
 * branches): the 6 refined-grid profiles of means.h5, each of nxmr values, in
 * the reference's write order (:583-597): Sbar, Srms, vxS, vyS, vzS, chiS.
 * Sums in the reference's loop order (one rank), then :574-581. */
void orc_sal_stats(orc_state *s, double *out) {
    REF_DECL
    const double *sal = s->f[ORC_SAL], *vxr = s->f[ORC_VXR], *vyr = s->f[ORC_VYR], *vzr = s->f[ORC_VZR];
    const double *sbp = s->f[ORC_SALBP], *stp = s->f[ORC_SALTP];
    for (int q = 0; q < 6 * nxmr; q++) out[q] = 0.0;
    double *Sb = out - 1, *Sr = out + nxmr - 1, *XS = out + 2 * nxmr - 1, *YS = out + 3 * nxmr - 1;
    double *ZS = out + 4 * nxmr - 1, *CH = out + 5 * nxmr - 1;        /* 1-based views */
    const double inyzmr = 1.0 / nymr / nzmr;
    for (int ic = 1; ic <= nzmr; ic++) {          /* :544-552 */
        int i0 = W(ic, nzmr);
        for (int jc = 1; jc <= nymr; jc++) {
            int j0 = W(jc, nymr);
            for (int k = 1; k <= nxmr; k++) {
                double v = sal[R(k, j0, i0)];
                Sb[k] = Sb[k] + v;
                Sr[k] = Sr[k] + v * v;
            }
        }
    }
    for (int ic = 1; ic <= nzmr; ic++) {          /* :555-563 */
        int i0 = W(ic, nzmr), ip = W(ic + 1, nzmr);
        for (int jc = 1; jc <= nymr; jc++) {
            int j0 = W(jc, nymr), jp = W(jc + 1, nymr);
            for (int k = 1; k <= nxmr; k++) {
                double v = sal[R(k, j0, i0)];
                XS[k] = XS[k] + 0.5 * (vxr[R(k, j0, i0)] + vxr[R(k + 1, j0, i0)]) * v;
                YS[k] = YS[k] + 0.5 * (vyr[R(k, j0, i0)] + vyr[R(k, jp, i0)]) * v;
                ZS[k] = ZS[k] + 0.5 * (vzr[R(k, j0, i0)] + vzr[R(k, j0, ip)]) * v;
            }
        }
    }
    for (int ic = 1; ic <= nzmr; ic++) {          /* CalcDissipationSal, :668-690 */
        int i0 = W(ic, nzmr), ip = W(ic + 1, nzmr), im = W(ic - 1, nzmr);
        for (int jc = 1; jc <= nymr; jc++) {
            int j0 = W(jc, nymr), jp = W(jc + 1, nymr), jm = W(jc - 1, nymr);
            for (int k = 1; k <= nxmr; k++) {
                double a = (sal[R(k, j0, ip)] - sal[R(k, j0, im)]) * 0.5 * s->dzr;
                CH[k] = CH[k] + a * a;
                double b = (sal[R(k, jp, i0)] - sal[R(k, jm, i0)]) * 0.5 * s->dyr;
                CH[k] = CH[k] + b * b;
            }
            double t1 = 0.5 * s->dxr / s->g3rmr[1];
            double w = (sal[R(2, j0, i0)] - sal[R(1, j0, i0)] + 2.0 * s->c.sfixs * (sal[R(1, j0, i0)] - sbp[RP2(j0, i0)])) * t1;
            CH[1] = CH[1] + w * w;
            for (int k = 2; k <= nxmr - 1; k++) {
                double tk = 0.5 * s->dxr / s->g3rmr[k];
                double d = (sal[R(k + 1, j0, i0)] - sal[R(k - 1, j0, i0)]) * tk;
                CH[k] = CH[k] + d * d;
            }
            double tn = 0.5 * s->dxr / s->g3rmr[nxmr];
            w = (sal[R(nxmr, j0, i0)] - sal[R(nxmr - 1, j0, i0)] + 2.0 * s->c.sfixn * (stp[RP2(j0, i0)] - sal[R(nxmr, j0, i0)])) * tn;
            CH[nxmr] = CH[nxmr] + w * w;
        }
    }
    for (int k = 1; k <= nxmr; k++) {             /* :574-581 */
        Sb[k] = Sb[k] * inyzmr;
        Sr[k] = sqrt(Sr[k] * inyzmr);
        XS[k] = XS[k] * inyzmr;
        YS[k] = YS[k] * inyzmr;
        ZS[k] = ZS[k] * inyzmr;
        CH[k] = CH[k] / s->pecs * inyzmr;
    }
}

/* CalcPhiStats, phasefield.F90:750-798: phibar and phirms of the refined phase
 * field (nxmr values each), sums in the reference's loop order, then :781-784. */
void orc_phi_stats(orc_state *s, double *out) {
    REF_DECL
    const double *phi = s->f[ORC_PHI];
    double *pb = out - 1, *pr = out + nxmr - 1;
    for (int q = 0; q < 2 * nxmr; q++) out[q] = 0.0;
    const double inyzm = 1.0 / nymr / nzmr;
    for (int ic = 1; ic <= nzmr; ic++) {
        int i0 = W(ic, nzmr);
        for (int jc = 1; jc <= nymr; jc++) {
            int j0 = W(jc, nymr);
            for (int k = 1; k <= nxmr; k++) {
                double v = phi[R(k, j0, i0)];
                pb[k] = pb[k] + v;
                pr[k] = pr[k] + v * v;
            }
        }
    }
    for (int k = 1; k <= nxmr; k++) {
        pb[k] = pb[k] * inyzm;
        pr[k] = sqrt(pr[k] * inyzm);
    }
}

/* ExplicitSalinity, salinity.F90:268-366 (non-IBM branch) */
void orc_explicit_salinity(orc_state *s) {
    REF_DECL
    const double *sal = s->f[ORC_SAL], *vxr = s->f[ORC_VXR], *vyr = s->f[ORC_VYR], *vzr = s->f[ORC_VZR];
    const double *sbp = s->f[ORC_SALBP], *stp = s->f[ORC_SALTP];
    double *hsal = s->f[ORC_HSAL];
    double udyr = 0.5 * s->dyr, udzr = 0.5 * s->dzr, udyrq = s->dyqr / s->pecs, udzrq = s->dzqr / s->pecs;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzmr; ic++) {
        int i0 = W(ic, nzmr), im = W(ic - 1, nzmr), ip = W(ic + 1, nzmr);
        for (int jc = 1; jc <= nymr; jc++) {
            int j0 = W(jc, nymr), jm = W(jc - 1, nymr), jp = W(jc + 1, nymr);
            for (int kc = 1; kc <= nxmr; kc++) {
                int km = kc - 1, kp = kc + 1;
                double hsx;
                if (kc == 1)
                    hsx = (vxr[R(kp, j0, i0)] * (sal[R(kp, j0, i0)] + sal[R(kc, j0, i0)])
                         - vxr[R(kc, j0, i0)] * 2.0 * sbp[RP2(j0, i0)]) * s->udx3mr[kc] * 0.5;
                else if (kc == nxmr)
                    hsx = (vxr[R(kp, j0, i0)] * 2.0 * stp[RP2(j0, i0)]
                         - vxr[R(kc, j0, i0)] * (sal[R(kc, j0, i0)] + sal[R(km, j0, i0)])) * s->udx3mr[kc] * 0.5;
                else
                    hsx = (vxr[R(kp, j0, i0)] * (sal[R(kp, j0, i0)] + sal[R(kc, j0, i0)])
                         - vxr[R(kc, j0, i0)] * (sal[R(kc, j0, i0)] + sal[R(km, j0, i0)])) * s->udx3mr[kc] * 0.5;
                double hsy = (vyr[R(kc, jp, i0)] * (sal[R(kc, jp, i0)] + sal[R(kc, j0, i0)])
                            - vyr[R(kc, j0, i0)] * (sal[R(kc, j0, i0)] + sal[R(kc, jm, i0)])) * udyr;
                double hsz = (vzr[R(kc, j0, ip)] * (sal[R(kc, j0, ip)] + sal[R(kc, j0, i0)])
                            - vzr[R(kc, j0, i0)] * (sal[R(kc, j0, i0)] + sal[R(kc, j0, im)])) * udzr;
                double dyys = (sal[R(kc, jp, i0)] - 2.0 * sal[R(kc, j0, i0)] + sal[R(kc, jm, i0)]) * udyrq;
                double dzzs = (sal[R(kc, j0, ip)] - 2.0 * sal[R(kc, j0, i0)] + sal[R(kc, j0, im)]) * udzrq;
                hsal[R(kc, j0, i0)] = -(hsx + hsy + hsz) + dyys + dzzs;
            }
        }
    }
}

/* AddSalBuoyancy, salinity.F90:495-506, into the component picked by gAxis (TimeMarcher.F90:54-62) */
void orc_add_sal_buoyancy(orc_state *s) {
    BASE_DECL
    double *rkv = s->c.gaxis == 1 ? s->f[ORC_QCAP] : s->c.gaxis == 2 ? s->f[ORC_DPH] : s->f[ORC_DQ];
    const double *salc = s->f[ORC_SALC];
    for (int i = 0; i < nzm; i++)
        for (int j = 0; j < nym; j++)
            for (int kc = 1; kc <= nxm; kc++) rkv[B(kc, j, i)] -= s->bycs * salc[B(kc, j, i)];
}

/* ImplicitSalinity + SolveImpEqnUpdate_Sal, salinity.F90:369-454 */
void orc_implicit_salinity(orc_state *s) {
    REF_DECL
    double *sal = s->f[ORC_SAL], *rusal = s->f[ORC_RUSAL], *rhsr = s->f[ORC_RHSR];
    const double *hsal = s->f[ORC_HSAL], *sbp = s->f[ORC_SALBP], *stp = s->f[ORC_SALTP];
    double alpec = s->al / s->pecs;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzmr; ic++) {
        int i0 = W(ic, nzmr);
        for (int jc = 1; jc <= nymr; jc++) {
            int j0 = W(jc, nymr);
            for (int kc = 1; kc <= nxmr; kc++) {
                double dxxs;
                if (kc == 1)
                    dxxs = sal[R(kc + 1, j0, i0)] * s->ap3sskr[kc] + sal[R(kc, j0, i0)] * s->ac3sskr[kc]
                         - (s->ap3sskr[kc] + s->ac3sskr[kc]) * sbp[RP2(j0, i0)] * s->c.sfixs;
                else if (kc == nxmr)
                    dxxs = sal[R(kc, j0, i0)] * s->ac3sskr[kc] + sal[R(kc - 1, j0, i0)] * s->am3sskr[kc]
                         - (s->am3sskr[kc] + s->ac3sskr[kc]) * stp[RP2(j0, i0)] * s->c.sfixn;
                else
                    dxxs = sal[R(kc + 1, j0, i0)] * s->ap3sskr[kc] + sal[R(kc, j0, i0)] * s->ac3sskr[kc]
                         + sal[R(kc - 1, j0, i0)] * s->am3sskr[kc];
                rhsr[R(kc, j0, i0)] = (s->ga * hsal[R(kc, j0, i0)] + s->ro * rusal[R(kc, j0, i0)] + alpec * dxxs) * s->dt;
                rusal[R(kc, j0, i0)] = hsal[R(kc, j0, i0)];
            }
        }
    }
    if (s->ibm_mask[4])            /* salinity.F90:403-404 */
        solve_columns_ibm(s, 4, sal, rhsr, nxr, nymr, nzmr, nxmr, s->am3sskr, s->ac3sskr, s->ap3sskr, 0.5 * s->al * s->dt / s->pecs);
    else
        solve_columns(s, sal, rhsr, nxr, nymr, nzmr, nxmr, s->am3sskr, s->ac3sskr, s->ap3sskr,
                  0.5 * s->al * s->dt / s->pecs, 1, nxmr, 1, nxmr);
}

/* InterpSalMultigrid (salinity.F90:458-491) via interpolate_xyz_to_coarse_fast
 * (HermiteInterpolations.F90:256-312).  Only the fast path is restated: it
 * is taken whenever the refined grid brackets the coarse one. */
/* interpolate_xyz_to_coarse, HermiteInterpolations.F90:218-250: the general
 * refined -> coarse loop, taken when the refined cell centres do not bracket
 * the coarse ones (xmr(1) >= xm(1) or xmr(nxmr) <= xm(nxm)).  Every refined
 * cell kcr = 0..nxmr (y, z: 0..n, the periodic halo cell 0 included) spreads
 * its 4x4x4 block tpdvr(kcr-1:kcr+2, ...) to the coarse points it owns
 * (irangr(kcr) <= kc < irangr(kcr+1)) with the cxphic / cyphic / czphic
 * weights.  tpdvr holds the x ghosts at 0 and nxr, zero beyond them.
 * out = v (acc = 0) or out += scale * v (acc = 1), kc = 1..nxm. */
static void to_coarse_general(orc_state *s, const double *tpdvr, double *out, double scale, int acc) {
    BASE_DECL REF_DECL
    for (int icr = 0; icr <= nzmr; icr++) {
        int ilo = s->krangr[icr] > 1 ? s->krangr[icr] : 1, ihi = s->krangr[icr + 1] - 1 < nzm ? s->krangr[icr + 1] - 1 : nzm;
        for (int jcr = 0; jcr <= nymr; jcr++) {
            int jlo = s->jrangr[jcr] > 1 ? s->jrangr[jcr] : 1, jhi = s->jrangr[jcr + 1] - 1 < nym ? s->jrangr[jcr + 1] - 1 : nym;
            for (int kcr = 0; kcr <= nxmr; kcr++) {
                int klo = s->irangr[kcr] > 1 ? s->irangr[kcr] : 1, khi = s->irangr[kcr + 1] - 1 < nxm ? s->irangr[kcr + 1] - 1 : nxm;
                if (ilo > ihi || jlo > jhi || klo > khi) continue;
                double qv3[4][4][4];
                for (int c = 0; c < 4; c++)
                    for (int b = 0; b < 4; b++)
                        for (int a = 0; a < 4; a++) {
                            int kk = kcr - 1 + a;
                            qv3[c][b][a] = (kk < -1 || kk > nxr + 1) ? 0.0
                                         : tpdvr[TR(kk, W(jcr - 1 + b, nymr), W(icr - 1 + c, nzmr))];
                        }
                for (int ic = ilo; ic <= ihi; ic++) {
                    const double *cz = s->czphic + (size_t)(ic - 1) * 4;
                    double qv2[4][4];
                    for (int b = 0; b < 4; b++)
                        for (int a = 0; a < 4; a++)
                            qv2[b][a] = qv3[0][b][a] * cz[0] + qv3[1][b][a] * cz[1] + qv3[2][b][a] * cz[2] + qv3[3][b][a] * cz[3];
                    for (int jc = jlo; jc <= jhi; jc++) {
                        const double *cy = s->cyphic + (size_t)(jc - 1) * 4;
                        double qv1[4];
                        for (int a = 0; a < 4; a++) qv1[a] = qv2[0][a] * cy[0] + qv2[1][a] * cy[1] + qv2[2][a] * cy[2] + qv2[3][a] * cy[3];
                        for (int kc = klo; kc <= khi; kc++) {
                            const double *cx = s->cxphic + (size_t)(kc - 1) * 4;
                            double v = qv1[0] * cx[0] + qv1[1] * cx[1] + qv1[2] * cx[2] + qv1[3] * cx[3];
                            size_t o = B(kc, jc - 1, ic - 1);
                            if (acc) out[o] += scale * v; else out[o] = v;
                        }
                    }
                }
            }
        }
    }
}

static int brackets(orc_state *s) { return s->xmr[1] < s->xm[1] && s->xmr[s->nxmr] > s->xm[s->nxm]; }

void orc_interp_sal_multigrid(orc_state *s) {
    BASE_DECL REF_DECL
    const double *sal = s->f[ORC_SAL], *sbp = s->f[ORC_SALBP], *stp = s->f[ORC_SALTP];
    double *salc = s->f[ORC_SALC], *tpdvr = s->tpdvr;
    memset(salc, 0, sizeof(double) * nbase(s));
    memset(tpdvr, 0, sizeof(double) * (size_t)(nxr + 3) * nymr * nzmr);
    for (int i = 0; i < nzmr; i++)
        for (int j = 0; j < nymr; j++) {
            for (int k = 1; k <= nxmr; k++) tpdvr[TR(k, j, i)] = sal[R(k, j, i)];
            tpdvr[TR(0, j, i)] = s->c.sfixs == 1 ? 2.0 * sbp[RP2(j, i)] - sal[R(1, j, i)] : sal[R(1, j, i)];
            tpdvr[TR(nxr, j, i)] = s->c.sfixn == 1 ? 2.0 * stp[RP2(j, i)] - sal[R(nxmr, j, i)] : sal[R(nxmr, j, i)];
        }
    if (!brackets(s)) {          /* salinity.F90:487-488 */
        to_coarse_general(s, tpdvr, salc, 1.0, 0);
        return;
    }
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int icr = s->krangb[ic];
        for (int jc = 1; jc <= nym; jc++) {
            int jcr = s->jrangb[jc];
            for (int kc = 1; kc <= nxm; kc++) {
                int kcr = s->irangb[kc];
                double qv3[4][4][4];
                for (int c = 0; c < 4; c++)
                    for (int b = 0; b < 4; b++)
                        for (int a = 0; a < 4; a++)
                            qv3[c][b][a] = tpdvr[TR(kcr - 2 + a, W(jcr - 2 + b, nymr), W(icr - 2 + c, nzmr))];
                const double *cz = s->czsalc + (size_t)(ic - 1) * 4, *cy = s->cysalc + (size_t)(jc - 1) * 4, *cx = s->cxsalc + (size_t)(kc - 1) * 4;
                double qv2[4][4], qv1[4];
                for (int b = 0; b < 4; b++)
                    for (int a = 0; a < 4; a++)
                        qv2[b][a] = qv3[0][b][a] * cz[0] + qv3[1][b][a] * cz[1] + qv3[2][b][a] * cz[2] + qv3[3][b][a] * cz[3];
                for (int a = 0; a < 4; a++) qv1[a] = qv2[0][a] * cy[0] + qv2[1][a] * cy[1] + qv2[2][a] * cy[2] + qv2[3][a] * cy[3];
                salc[B(kc, W(jc, nym), W(ic, nzm))] = qv1[0] * cx[0] + qv1[1] * cx[1] + qv1[2] * cx[2] + qv1[3] * cx[3];
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* phase field (src/phasefield.F90)                                          */
/* ------------------------------------------------------------------------ */
void orc_pf_scalars(orc_state *s, double out[4]) {
    out[0] = s->pf_d; out[1] = s->pf_eps;
    out[2] = s->ren * pow(1.51044285 * s->pf_eps, 2);     /* pf_eta, phasefield.F90:531 */
    out[3] = s->pf_d / (s->pf_eps * s->pf_eps);            /* pf_B, phasefield.F90:357 */
}

/* ExplicitPhase, phasefield.F90:345-383 */
void orc_explicit_phase(orc_state *s) {
    REF_DECL
    const double *phi = s->f[ORC_PHI], *tempr = s->f[ORC_TEMPR];
    double *hphi = s->f[ORC_HPHI];
    double pf_b = s->pf_d / (s->pf_eps * s->pf_eps);
    double udzrq = s->pf_d * s->dzqr, udyrq = s->pf_d * s->dyqr;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzmr; ic++) {
        int i0 = W(ic, nzmr), im = W(ic - 1, nzmr), ip = W(ic + 1, nzmr);
        for (int jc = 1; jc <= nymr; jc++) {
            int j0 = W(jc, nymr), jm = W(jc - 1, nymr), jp = W(jc + 1, nymr);
            for (int kc = 1; kc <= nxmr; kc++) {
                double pc = phi[R(kc, j0, i0)];
                double dyyp = (phi[R(kc, jp, i0)] - 2.0 * pc + phi[R(kc, jm, i0)]) * udyrq;
                double dzzp = (phi[R(kc, j0, ip)] - 2.0 * pc + phi[R(kc, j0, im)]) * udzrq;
                double nlphi = pf_b * pc * (1.0 - pc) * (1.0 - 2.0 * pc + s->pf_a * (tempr[R(kc, j0, i0)] - s->pf_tm));
                hphi[R(kc, j0, i0)] = dyyp + dzzp - nlphi;
            }
        }
    }
}

/* AddVolumePenalty, phasefield.F90:525-566 */
void orc_add_volume_penalty(orc_state *s) {
    BASE_DECL
    const double *phic = s->f[ORC_PHIC], *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ];
    double *qcap = s->f[ORC_QCAP], *dph = s->f[ORC_DPH], *dq = s->f[ORC_DQ];
    double pf_eta = s->ren * pow(1.51044285 * s->pf_eps, 2), heta = 0.5 / pf_eta;
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), imm = W(ic - 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jmm = W(jc - 1, nym);
            for (int kc = 2; kc <= nxm; kc++)
                qcap[B(kc, j0, i0)] -= (phic[B(kc, j0, i0)] + phic[B(kc - 1, j0, i0)]) * vx[B(kc, j0, i0)] * heta;
            for (int kc = 1; kc <= nxm; kc++)
                dph[B(kc, j0, i0)] -= (phic[B(kc, j0, i0)] + phic[B(kc, jmm, i0)]) * vy[B(kc, j0, i0)] * heta;
            for (int kc = 1; kc <= nxm; kc++)
                dq[B(kc, j0, i0)] -= (phic[B(kc, j0, i0)] + phic[B(kc, j0, imm)]) * vz[B(kc, j0, i0)] * heta;
        }
    }
}

/* ImplicitPhase + SolveImpEqnUpdate_Phi, phasefield.F90:386-466; rhsr is
 * left holding phi(l+1) - phi(l) for AddLatentHeat */
void orc_implicit_phase(orc_state *s) {
    REF_DECL
    double *phi = s->f[ORC_PHI], *ruphi = s->f[ORC_RUPHI], *rhsr = s->f[ORC_RHSR];
    const double *hphi = s->f[ORC_HPHI];
    double alpec = s->al * s->pf_d;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzmr; ic++) {
        int i0 = W(ic, nzmr);
        for (int jc = 1; jc <= nymr; jc++) {
            int j0 = W(jc, nymr);
            for (int kc = 1; kc <= nxmr; kc++) {
                double dxxp;
                if (kc == 1)
                    dxxp = phi[R(kc + 1, j0, i0)] * s->ap3spkr[kc] + phi[R(kc, j0, i0)] * s->ac3spkr[kc];
                else if (kc == nxmr)
                    dxxp = phi[R(kc, j0, i0)] * s->ac3spkr[kc] + phi[R(kc - 1, j0, i0)] * s->am3spkr[kc];
                else
                    dxxp = phi[R(kc + 1, j0, i0)] * s->ap3spkr[kc] + phi[R(kc, j0, i0)] * s->ac3spkr[kc]
                         + phi[R(kc - 1, j0, i0)] * s->am3spkr[kc];
                rhsr[R(kc, j0, i0)] = (s->ga * hphi[R(kc, j0, i0)] + s->ro * ruphi[R(kc, j0, i0)] + alpec * dxxp) * s->dt;
                ruphi[R(kc, j0, i0)] = hphi[R(kc, j0, i0)];
            }
        }
    }
    solve_columns(s, phi, rhsr, nxr, nymr, nzmr, nxmr, s->am3spkr, s->ac3spkr, s->ap3spkr,
                  0.5 * s->al * s->dt * s->pf_d, 1, nxmr, 1, nxmr);
}

/* Refined field with x ghosts (tpdvr) to the coarse cell centres through the
 * fast loop of interpolate_xyz_to_coarse_fast, "phi" stencils
 * (HermiteInterpolations.F90:256-312, cxphic/irangs): out = v (acc = 0) or
 * out += scale * v (acc = 1) for kc = 1..nxm. */
static void pf_to_coarse(orc_state *s, const double *tpdvr, double *out, double scale, int acc) {
    BASE_DECL REF_DECL
    if (!brackets(s)) {          /* phasefield.F90:489-492 and :701-723 */
        to_coarse_general(s, tpdvr, out, scale, acc);
        return;
    }
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int icr = s->krangs[ic];
        for (int jc = 1; jc <= nym; jc++) {
            int jcr = s->jrangs[jc];
            for (int kc = 1; kc <= nxm; kc++) {
                int kcr = s->irangs[kc];
                double qv3[4][4][4];
                for (int c = 0; c < 4; c++)
                    for (int b = 0; b < 4; b++)
                        for (int a = 0; a < 4; a++)
                            qv3[c][b][a] = tpdvr[TR(kcr - 2 + a, W(jcr - 2 + b, nymr), W(icr - 2 + c, nzmr))];
                const double *cz = s->czphic + (size_t)(ic - 1) * 4, *cy = s->cyphic + (size_t)(jc - 1) * 4, *cx = s->cxphic + (size_t)(kc - 1) * 4;
                double qv2[4][4], qv1[4];
                for (int b = 0; b < 4; b++)
                    for (int a = 0; a < 4; a++)
                        qv2[b][a] = qv3[0][b][a] * cz[0] + qv3[1][b][a] * cz[1] + qv3[2][b][a] * cz[2] + qv3[3][b][a] * cz[3];
                for (int a = 0; a < 4; a++) qv1[a] = qv2[0][a] * cy[0] + qv2[1][a] * cy[1] + qv2[2][a] * cy[2] + qv2[3][a] * cy[3];
                double v = qv1[0] * cx[0] + qv1[1] * cx[1] + qv1[2] * cx[2] + qv1[3] * cx[3];
                size_t o = B(kc, W(jc, nym), W(ic, nzm));
                if (acc) out[o] += scale * v; else out[o] = v;
            }
        }
    }
}

/* copy a refined field into tpdvr with the zero-gradient x ghosts
 * tpdvr(0) = f(1), tpdvr(nxr) = f(nxmr) (phasefield.F90:479-487, :674-683) */
static void fill_neumann_ghosts(orc_state *s, const double *f) {
    REF_DECL
    double *tpdvr = s->tpdvr;
    memset(tpdvr, 0, sizeof(double) * (size_t)(nxr + 3) * nymr * nzmr);
    for (int i = 0; i < nzmr; i++)
        for (int j = 0; j < nymr; j++) {
            for (int k = 1; k <= nxmr; k++) tpdvr[TR(k, j, i)] = f[R(k, j, i)];
            tpdvr[TR(0, j, i)] = f[R(1, j, i)];
            tpdvr[TR(nxr, j, i)] = f[R(nxmr, j, i)];
        }
}

/* AddLatentHeat, phasefield.F90:651-725 (fast path): hro += pf_S d(phi)/dt
 * on the coarse grid, d(phi) from rhsr after ImplicitPhase */
void orc_add_latent_heat(orc_state *s) {
    fill_neumann_ghosts(s, s->f[ORC_RHSR]);
    pf_to_coarse(s, s->tpdvr, s->f[ORC_HRO], s->pf_s / (s->al * s->dt), 1);
}

/* InterpPhiMultigrid, phasefield.F90:470-494 (fast path) */
void orc_interp_phi_multigrid(orc_state *s) {
    memset(s->f[ORC_PHIC], 0, sizeof(double) * nbase(s));
    fill_neumann_ghosts(s, s->f[ORC_PHI]);
    pf_to_coarse(s, s->tpdvr, s->f[ORC_PHIC], 1.0, 0);
}

/* InterpTempMultigrid, phasefield.F90:498-520 */
void orc_interp_temp_multigrid(orc_state *s) {
    BASE_DECL REF_DECL
    const double *temp = s->f[ORC_TEMP], *tbp = s->f[ORC_TEMPBP], *ttp = s->f[ORC_TEMPTP];
    double *tpdv = s->tpdv, *tpdvr = s->tpdvr, *tempr = s->f[ORC_TEMPR];
    memset(tpdv, 0, sizeof(double) * (size_t)(nx + 3) * nym * nzm);
    memset(tpdvr, 0, sizeof(double) * (size_t)(nxr + 3) * nymr * nzmr);
    memset(tempr, 0, sizeof(double) * nref(s));
    for (int i = 0; i < nzm; i++)
        for (int j = 0; j < nym; j++) {
            tpdv[TP(0, j, i)] = (1.0 - 2.0 * s->c.tfixs) * temp[B(1, j, i)] + 2.0 * s->c.tfixs * tbp[P2(j, i)];
            tpdv[TP(nx, j, i)] = (1.0 - 2.0 * s->c.tfixn) * temp[B(nxm, j, i)] + 2.0 * s->c.tfixn * ttp[P2(j, i)];
            for (int k = 1; k <= nxm; k++) tpdv[TP(k, j, i)] = temp[B(k, j, i)];
        }
    interp_to_refined(s, tpdv, tpdvr);
    for (int i = 0; i < nzmr; i++)
        for (int j = 0; j < nymr; j++)
            for (int k = 1; k <= nxmr; k++) tempr[R(k, j, i)] = tpdvr[TR(k, j, i)];
}

/* ------------------------------------------------------------------------ */
/* moisture (src/moisture.F90): specific humidity on the base grid          */
/* ------------------------------------------------------------------------ */
void orc_moist_scalars(orc_state *s, double out[5]) {
    out[0] = s->q_alpha; out[1] = s->q_beta; out[2] = s->q_gamma; out[3] = s->q_tau; out[4] = s->pecq;
}

/* SetHumidityBCs, moisture.F90:129-148: saturated walls, humbp =
 * exp(alpha_q tempbp), humtp = exp(alpha_q (temptp - beta_q)); zero at the
 * top with the unsaturated-top option qfixN = 2 */
void orc_set_humidity_bcs(orc_state *s) {
    BASE_DECL
    const double *tbp = s->f[ORC_TEMPBP], *ttp = s->f[ORC_TEMPTP];
    double *hbp = s->f[ORC_HUMBP], *htp = s->f[ORC_HUMTP];
    for (int i0 = 0; i0 < nzm; i0++)
        for (int j0 = 0; j0 < nym; j0++) {
            hbp[P2(j0, i0)] = exp(s->q_alpha * tbp[P2(j0, i0)]);
            htp[P2(j0, i0)] = s->c.q_fixn == 2 ? 0.0 : exp(s->q_alpha * (ttp[P2(j0, i0)] - s->q_beta));
        }
}

/* ExplicitHumidity, moisture.F90:220-292: advection in flux form with the
 * wall values, kappa_q times the y/z second differences, and the
 * condensation sink -(q - qsat)/tau_q where q > qsat */
void orc_explicit_humidity(orc_state *s) {
    BASE_DECL
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ], *q = s->f[ORC_HUMID];
    const double *qs = s->f[ORC_QSAT], *hbp = s->f[ORC_HUMBP], *htp = s->f[ORC_HUMTP];
    double *rk = s->f[ORC_HHUMID];
    const double udy = 0.5 * s->dy, udz = 0.5 * s->dz, itau = 1.0 / s->q_tau;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), im = W(ic - 1, nzm), ip = W(ic + 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jm = W(jc - 1, nym), jp = W(jc + 1, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                double hqx;
                if (kc == 1)
                    hqx = (vx[B(kc + 1, j0, i0)] * (q[B(kc + 1, j0, i0)] + q[B(kc, j0, i0)])
                         - vx[B(kc, j0, i0)] * 2.0 * hbp[P2(j0, i0)]) * 0.5 * s->udx3m[kc];
                else if (kc == nxm)
                    hqx = (vx[B(kc + 1, j0, i0)] * 2.0 * htp[P2(j0, i0)]
                         - vx[B(kc, j0, i0)] * (q[B(kc, j0, i0)] + q[B(kc - 1, j0, i0)])) * 0.5 * s->udx3m[kc];
                else
                    hqx = (vx[B(kc + 1, j0, i0)] * (q[B(kc + 1, j0, i0)] + q[B(kc, j0, i0)])
                         - vx[B(kc, j0, i0)] * (q[B(kc, j0, i0)] + q[B(kc - 1, j0, i0)])) * 0.5 * s->udx3m[kc];
                double hqy = (vy[B(kc, jp, i0)] * (q[B(kc, jp, i0)] + q[B(kc, j0, i0)])
                            - vy[B(kc, j0, i0)] * (q[B(kc, j0, i0)] + q[B(kc, jm, i0)])) * udy;
                double hqz = (vz[B(kc, j0, ip)] * (q[B(kc, j0, ip)] + q[B(kc, j0, i0)])
                            - vz[B(kc, j0, i0)] * (q[B(kc, j0, i0)] + q[B(kc, j0, im)])) * udz;
                double dyyq = (q[B(kc, jp, i0)] - 2.0 * q[B(kc, j0, i0)] + q[B(kc, jm, i0)]) * s->dyq;
                double dzzq = (q[B(kc, j0, ip)] - 2.0 * q[B(kc, j0, i0)] + q[B(kc, j0, im)]) * s->dzq;
                double r = s->kappa_q * (dyyq + dzzq) - (hqx + hqy + hqz);
                if (q[B(kc, j0, i0)] > qs[B(kc, j0, i0)]) r = r + (qs[B(kc, j0, i0)] - q[B(kc, j0, i0)]) * itau;
                rk[B(kc, j0, i0)] = r;
            }
        }
    }
}

/* AddCondensation, moisture.F90:296-318: the latent heat of condensing
 * vapour, gamma_q/tau_q (q - qsat) where q > qsat, into the T explicit term */
void orc_add_condensation(orc_state *s) {
    BASE_DECL
    const double *q = s->f[ORC_HUMID], *qs = s->f[ORC_QSAT];
    double *hro = s->f[ORC_HRO];
    const double gtau = s->q_gamma / s->q_tau;
    for (int i0 = 0; i0 < nzm; i0++)
        for (int j0 = 0; j0 < nym; j0++)
            for (int kc = 1; kc <= nxm; kc++) {
                const size_t c = B(kc, j0, i0);
                if (q[c] > qs[c]) hro[c] = hro[c] + (g_plant == 7 ? -gtau : gtau) * (q[c] - qs[c]);
            }
}

/* UpdateSaturation, moisture.F90:321-333: qsat = exp(alpha_q (T - beta_q x)),
 * T being the buoyancy field of this model */
void orc_update_saturation(orc_state *s) {
    BASE_DECL
    const double *temp = s->f[ORC_TEMP];
    double *qs = s->f[ORC_QSAT];
    for (int i0 = 0; i0 < nzm; i0++)
        for (int j0 = 0; j0 < nym; j0++)
            for (int kc = 1; kc <= nxm; kc++)
                qs[B(kc, j0, i0)] = exp(s->q_alpha * (temp[B(kc, j0, i0)] - s->q_beta * s->xm[kc]));
}

/* ImplicitHumidity, moisture.F90:336-373, and SolveImpEqnUpdate_Humidity,
 * :377-418: the x-diffusion with the wall values (qfixS / qfixN), then the
 * Crank-Nicolson solve with the T coefficients and pecq */
void orc_implicit_humidity(orc_state *s) {
    BASE_DECL
    double *q = s->f[ORC_HUMID], *rhs = s->f[ORC_RHS], *hk = s->f[ORC_RUHUMID];
    const double *rk = s->f[ORC_HHUMID], *hbp = s->f[ORC_HUMBP], *htp = s->f[ORC_HUMTP];
    const double alpec = s->al / s->pecq;
    apply_threads();
#pragma omp parallel for
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                double dxxq;
                if (kc == 1)
                    dxxq = q[B(kc + 1, j0, i0)] * s->ap3ssk[kc] + q[B(kc, j0, i0)] * s->ac3ssk[kc]
                         - (s->ap3ssk[kc] + s->ac3ssk[kc]) * hbp[P2(j0, i0)] * s->q_fixs;
                else if (kc == nxm)
                    dxxq = -(s->am3ssk[kc] + s->ac3ssk[kc]) * htp[P2(j0, i0)] * s->q_fixn
                         + q[B(kc, j0, i0)] * s->ac3ssk[kc] + q[B(kc - 1, j0, i0)] * s->am3ssk[kc];
                else
                    dxxq = q[B(kc + 1, j0, i0)] * s->ap3ssk[kc] + q[B(kc, j0, i0)] * s->ac3ssk[kc]
                         + q[B(kc - 1, j0, i0)] * s->am3ssk[kc];
                const double a = g_plant == 8 ? s->al / s->pect : alpec;
                rhs[B(kc, j0, i0)] = (s->ga * rk[B(kc, j0, i0)] + s->ro * hk[B(kc, j0, i0)] + a * dxxq) * s->dt;
                hk[B(kc, j0, i0)] = rk[B(kc, j0, i0)];
            }
        }
    }
    solve_columns(s, q, rhs, nx, nym, nzm, nxm, s->am3ssk, s->ac3ssk, s->ap3ssk, 0.5 * s->al * s->dt / s->pecq, 1, nxm, 1, nxm);
}

/* CalcMoistStats, moisture.F90:422-505: the 7 wall-parallel profiles qbar,
 * qsbar, qrms, qrel, vxq, vyq, vzq (nxm each, in the order of means.h5),
 * summed in the reference's loop order */
void orc_moist_stats(orc_state *s, double *out) {
    BASE_DECL
    const double *q = s->f[ORC_HUMID], *qs = s->f[ORC_QSAT];
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ];
    double *qbar = out, *qsbar = out + nxm, *qrms = out + 2 * nxm, *qrel = out + 3 * nxm;
    double *vxq = out + 4 * nxm, *vyq = out + 5 * nxm, *vzq = out + 6 * nxm;
    for (int a = 0; a < 7 * nxm; a++) out[a] = 0.0;
    const double inyzm = 1.0 / nym / nzm;
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), ip = W(ic + 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jp = W(jc + 1, nym);
            for (int k = 1; k <= nxm; k++) {
                const double h = q[B(k, j0, i0)];
                qbar[k - 1] = qbar[k - 1] + h;
                qsbar[k - 1] = qsbar[k - 1] + qs[B(k, j0, i0)];
                qrms[k - 1] = qrms[k - 1] + h * h;
                qrel[k - 1] = qrel[k - 1] + h / qs[B(k, j0, i0)];
                vxq[k - 1] = vxq[k - 1] + 0.5 * (vx[B(k, j0, i0)] + vx[B(k + 1, j0, i0)]) * h;
                vyq[k - 1] = vyq[k - 1] + 0.5 * (vy[B(k, j0, i0)] + vy[B(k, jp, i0)]) * h;
                vzq[k - 1] = vzq[k - 1] + 0.5 * (vz[B(k, j0, i0)] + vz[B(k, j0, ip)]) * h;
            }
        }
    }
    for (int k = 0; k < nxm; k++) {
        qbar[k] = qbar[k] * inyzm; qsbar[k] = qsbar[k] * inyzm; qrms[k] = sqrt(qrms[k] * inyzm);
        qrel[k] = qrel[k] * inyzm; vxq[k] = vxq[k] * inyzm; vyq[k] = vyq[k] * inyzm; vzq[k] = vzq[k] * inyzm;
    }
}

/* TimeMarcher.F90:13-165 (no IBM) */
void orc_time_marcher(orc_state *s, double dt) {
    for (int ns = 1; ns <= s->c.nsst; ns++) {
        orc_set_substep(s, dt, ns);
        orc_explicit_terms_vx(s);
        orc_explicit_terms_vy(s);
        orc_explicit_terms_vz(s);
        orc_explicit_terms_temp(s);
        if (s->c.salinity) {
            orc_explicit_salinity(s);
            if (s->c.active_s == 1) orc_add_sal_buoyancy(s);
        }
        if (s->c.phasefield) {
            orc_explicit_phase(s);
            orc_add_volume_penalty(s);
            orc_implicit_phase(s);
            orc_add_latent_heat(s);
        }
        if (s->c.moist) {                       /* :79-81 */
            orc_explicit_humidity(s);
            orc_add_condensation(s);
        }
        orc_implicit_and_update_vx(s);
        orc_implicit_and_update_vy(s);
        orc_implicit_and_update_vz(s);
        orc_implicit_and_update_temp(s);
        if (s->c.salinity) orc_implicit_salinity(s);
        if (s->c.moist) orc_implicit_humidity(s);     /* :97 */
        orc_calc_local_divergence(s);
        orc_solve_pressure_correction(s);
        orc_correct_velocity(s);
        orc_correct_pressure(s);
        if (s->c.salinity) {
            orc_interp_vel_mgrd(s);
            orc_interp_sal_multigrid(s);
        }
        if (s->c.phasefield) {
            orc_interp_temp_multigrid(s);
            orc_interp_phi_multigrid(s);
        }
        if (s->c.moist) orc_update_saturation(s);     /* :154 */
    }
}

/* CalcMaxCFL.F90:11-72 */
void orc_calc_max_cfl(orc_state *s, double *cflm, double *cflmr) {
    BASE_DECL
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ];
    double m = 0.00000001;
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), ip = W(ic + 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jp = W(jc + 1, nym);
            for (int k = 1; k <= nxm; k++) {
                double q = fabs((vz[B(k, j0, i0)] + vz[B(k, j0, ip)]) * 0.5 * s->dz)
                         + fabs((vy[B(k, j0, i0)] + vy[B(k, jp, i0)]) * 0.5 * s->dy)
                         + fabs((vx[B(k, j0, i0)] + vx[B(k + 1, j0, i0)]) * 0.5 * s->udx3m[k]);
                if (q > m) m = q;
            }
        }
    }
    *cflm = m;
    if (s->c.salinity) {
        REF_DECL
        const double *vxr = s->f[ORC_VXR], *vyr = s->f[ORC_VYR], *vzr = s->f[ORC_VZR];
        double mr = 1.0e-8;
        for (int ic = 1; ic <= nzmr; ic++) {
            int i0 = W(ic, nzmr), ip = W(ic + 1, nzmr);
            for (int jc = 1; jc <= nymr; jc++) {
                int j0 = W(jc, nymr), jp = W(jc + 1, nymr);
                for (int k = 1; k <= nxmr; k++) {
                    double q = fabs((vzr[R(k, j0, i0)] + vzr[R(k, j0, ip)]) * 0.5 * s->dzr)
                             + fabs((vyr[R(k, j0, i0)] + vyr[R(k, jp, i0)]) * 0.5 * s->dyr)
                             + fabs((vxr[R(k, j0, i0)] + vxr[R(k + 1, j0, i0)]) * 0.5 * s->udx3mr[k]);
                    if (q > mr) mr = q;
                }
            }
        }
        *cflmr = mr;
    } else *cflmr = m;
}

/* CheckDivergence.F90:11-87 */
void orc_check_divergence(orc_state *s, double *qmax, double *qmaxr) {
    BASE_DECL
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ];
    double m = -1.0e300;
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), ip = W(ic + 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jp = W(jc + 1, nym);
            for (int kc = 1; kc <= nxm; kc++) {
                double d = (vz[B(kc, j0, ip)] - vz[B(kc, j0, i0)]) * s->dz + (vy[B(kc, jp, i0)] - vy[B(kc, j0, i0)]) * s->dy
                         + (vx[B(kc + 1, j0, i0)] - vx[B(kc, j0, i0)]) * s->udx3m[kc];
                if (fabs(d) > m) m = fabs(d);
            }
        }
    }
    *qmax = m;
    if (s->c.salinity) {
        REF_DECL
        const double *vxr = s->f[ORC_VXR], *vyr = s->f[ORC_VYR], *vzr = s->f[ORC_VZR];
        double mr = -1.0e300;
        for (int ic = 1; ic <= nzmr; ic++) {
            int i0 = W(ic, nzmr), ip = W(ic + 1, nzmr);
            for (int jc = 1; jc <= nymr; jc++) {
                int j0 = W(jc, nymr), jp = W(jc + 1, nymr);
                for (int kc = 1; kc <= nxmr; kc++) {
                    double d = (vzr[R(kc, j0, ip)] - vzr[R(kc, j0, i0)]) * s->dzr + (vyr[R(kc, jp, i0)] - vyr[R(kc, j0, i0)]) * s->dyr
                             + (vxr[R(kc + 1, j0, i0)] - vxr[R(kc, j0, i0)]) * s->udx3mr[kc];
                    if (fabs(d) > mr) mr = fabs(d);
                }
            }
        }
        *qmaxr = mr;
    } else *qmaxr = m;
}

/* CalcMeanProfiles.F90:11-229 : the 15 wall-parallel mean profiles, each of
 * nxm values, in the order the reference writes them to means.h5
 * (:230-283): Tbar, vybar, vzbar, vxT, vyT, vzT, Trms, vxrms, vyrms, vzrms,
 * vxvy, vxvz, vyvz, chiT, epsilon.  Sums in the reference's loop order (one
 * rank), then the normalisations of :87-101 and :219-223. */
void orc_mean_profiles(orc_state *s, double *out) {
    BASE_DECL
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ], *temp = s->f[ORC_TEMP];
    const double *tbp = s->f[ORC_TEMPBP], *ttp = s->f[ORC_TEMPTP];
    double *Tbar = out, *vybar = out + nxm, *vzbar = out + 2 * nxm, *vxT = out + 3 * nxm, *vyT = out + 4 * nxm;
    double *vzT = out + 5 * nxm, *Trms = out + 6 * nxm, *vxrms = out + 7 * nxm, *vyrms = out + 8 * nxm;
    double *vzrms = out + 9 * nxm, *vxvy = out + 10 * nxm, *vxvz = out + 11 * nxm, *vyvz = out + 12 * nxm;
    double *chiT = out + 13 * nxm, *eps = out + 14 * nxm;
    for (int q = 0; q < 15 * nxm; q++) out[q] = 0.0;
    double *T = Tbar - 1, *VYB = vybar - 1, *VZB = vzbar - 1, *VXT = vxT - 1, *VYT = vyT - 1, *VZT = vzT - 1;
    double *TR = Trms - 1, *VXR = vxrms - 1, *VYR = vyrms - 1, *VZR = vzrms - 1, *XY = vxvy - 1, *XZ = vxvz - 1;
    double *YZ = vyvz - 1, *CH = chiT - 1, *EP = eps - 1;       /* 1-based views */
    const double inym = 1.0 / nym, inzm = 1.0 / nzm, dy = s->dy, dz = s->dz, dx = s->dx;
    /* :44-73 */
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), ip = W(ic + 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jp = W(jc + 1, nym);
            for (int k = 1; k <= nxm; k++) {
                int kp = k + 1;
                double t = temp[B(k, j0, i0)];
                double vxk = vx[B(k, j0, i0)], vxp = vx[B(kp, j0, i0)];
                double vyk = vy[B(k, j0, i0)], vyp = vy[B(k, jp, i0)];
                double vzk = vz[B(k, j0, i0)], vzp = vz[B(k, j0, ip)];
                T[k] = T[k] + t;
                VYB[k] = VYB[k] + 0.5 * (vyk + vyp);
                VZB[k] = VZB[k] + 0.5 * (vzk + vzp);
                VXT[k] = VXT[k] + 0.5 * (vxk + vxp) * t;
                VYT[k] = VYT[k] + 0.5 * (vyk + vyp) * t;
                VZT[k] = VZT[k] + 0.5 * (vzk + vzp) * t;
                TR[k] = TR[k] + t * t;
                VXR[k] = VXR[k] + 0.5 * (vxk * vxk + vxp * vxp);
                VYR[k] = VYR[k] + 0.5 * (vyk * vyk + vyp * vyp);
                VZR[k] = VZR[k] + 0.5 * (vzk * vzk + vzp * vzp);
                XY[k] = XY[k] + 0.25 * (vxk + vxp) * (vyk + vyp);
                XZ[k] = XZ[k] + 0.25 * (vxk + vxp) * (vzk + vzp);
                YZ[k] = YZ[k] + 0.25 * (vyk + vyp) * (vzk + vzp);
            }
        }
    }
    /* :89-101 */
    for (int k = 1; k <= nxm; k++) {
        T[k] = T[k] * inym * inzm;
        VYB[k] = VYB[k] * inym * inzm;
        VZB[k] = VZB[k] * inym * inzm;
        VXT[k] = VXT[k] * inym * inzm;
        VYT[k] = VYT[k] * inym * inzm;
        VZT[k] = VZT[k] * inym * inzm;
        TR[k] = sqrt(TR[k] * inym * inzm);
        VXR[k] = sqrt(VXR[k] * inym * inzm);
        VYR[k] = sqrt(VYR[k] * inym * inzm);
        VZR[k] = sqrt(VZR[k] * inym * inzm);
        XY[k] = XY[k] * inym * inzm;
        XZ[k] = XZ[k] * inym * inzm;
        YZ[k] = YZ[k] * inym * inzm;
    }
    /* chiT, :104-127 (tdx = 0.5*dx/g3rm(k)) */
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), ip = W(ic + 1, nzm), im = W(ic - 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jp = W(jc + 1, nym), jm = W(jc - 1, nym);
            for (int k = 1; k <= nxm; k++) {
                double a = (temp[B(k, j0, ip)] - temp[B(k, j0, im)]) * 0.5 * dz;
                double b = (temp[B(k, jp, i0)] - temp[B(k, jm, i0)]) * 0.5 * dy;
                CH[k] = CH[k] + a * a + b * b;
            }
            double tdx = 0.5 * dx / s->g3rm[1];
            double w = (temp[B(2, j0, i0)] - temp[B(1, j0, i0)] + 2.0 * s->c.tfixs * (temp[B(1, j0, i0)] - tbp[P2(j0, i0)])) * tdx;
            CH[1] = CH[1] + w * w;
            for (int k = 2; k <= nxm - 1; k++) {
                tdx = 0.5 * dx / s->g3rm[k];
                double d = (temp[B(k + 1, j0, i0)] - temp[B(k - 1, j0, i0)]) * tdx;
                CH[k] = CH[k] + d * d;
            }
            tdx = 0.5 * dx / s->g3rm[nxm];
            w = (temp[B(nxm, j0, i0)] - temp[B(nxm - 1, j0, i0)] + 2.0 * s->c.tfixn * (ttp[P2(j0, i0)] - temp[B(nxm, j0, i0)])) * tdx;
            CH[nxm] = CH[nxm] + w * w;
        }
    }
    /* epsilon, :165-215 */
    for (int ic = 1; ic <= nzm; ic++) {
        int i0 = W(ic, nzm), ip = W(ic + 1, nzm), im = W(ic - 1, nzm);
        for (int jc = 1; jc <= nym; jc++) {
            int j0 = W(jc, nym), jp = W(jc + 1, nym), jm = W(jc - 1, nym);
            for (int k = 1; k <= nxm; k++) {
                int kp = k + 1;
                double a = (vz[B(k, j0, ip)] - vz[B(k, j0, i0)]) * dz;
                double b1 = vy[B(k, jp, ip)] - vy[B(k, jp, im)], b2 = vy[B(k, j0, ip)] - vy[B(k, j0, im)];
                double c1 = vx[B(kp, j0, ip)] - vx[B(kp, j0, im)], c2 = vx[B(k, j0, ip)] - vx[B(k, j0, im)];
                double d1 = vz[B(k, jp, ip)] - vz[B(k, jm, ip)], d2 = vz[B(k, jp, i0)] - vz[B(k, jm, i0)];
                double e = (vy[B(k, jp, i0)] - vy[B(k, j0, i0)]) * dy;
                double f1 = vx[B(kp, jp, i0)] - vx[B(kp, jm, i0)], f2 = vx[B(k, jp, i0)] - vx[B(k, jm, i0)];
                double g = (vx[B(kp, j0, i0)] - vx[B(k, j0, i0)]) * s->udx3m[k];
                EP[k] = EP[k] + a * a + 0.125 * dz * dz * (b1 * b1 + b2 * b2) + 0.125 * dz * dz * (c1 * c1 + c2 * c2)
                        + 0.125 * dy * dy * (d1 * d1 + d2 * d2) + e * e + 0.125 * dy * dy * (f1 * f1 + f2 * f2) + g * g;
            }
            double tdx = 0.125 * (dx / s->g3rm[1]) * (dx / s->g3rm[1]);
            double z1 = vz[B(2, j0, ip)] - vz[B(1, j0, ip)] + s->c.inslws * 2.0 * (vz[B(1, j0, ip)] - 0.0);
            double z2 = vz[B(2, j0, i0)] - vz[B(1, j0, i0)] + s->c.inslws * 2.0 * (vz[B(1, j0, i0)] - 0.0);
            double y1 = vy[B(2, jp, i0)] - vy[B(1, jp, i0)] + s->c.inslws * 2.0 * (vy[B(1, jp, i0)] - s->c.xminusu);
            double y2 = vy[B(2, j0, i0)] - vy[B(1, j0, i0)] + s->c.inslws * 2.0 * (vy[B(1, j0, i0)] - s->c.xminusu);
            EP[1] = EP[1] + tdx * (z1 * z1 + z2 * z2) + tdx * (y1 * y1 + y2 * y2);
            for (int k = 2; k <= nxm - 1; k++) {
                tdx = 0.125 * (dx / s->g3rm[k]) * (dx / s->g3rm[k]);
                double p1 = vz[B(k + 1, j0, ip)] - vz[B(k - 1, j0, ip)], p2 = vz[B(k + 1, j0, i0)] - vz[B(k - 1, j0, i0)];
                double q1 = vy[B(k + 1, jp, i0)] - vy[B(k - 1, jp, i0)], q2 = vy[B(k + 1, j0, i0)] - vy[B(k - 1, j0, i0)];
                EP[k] = EP[k] + tdx * (p1 * p1 + p2 * p2) + tdx * (q1 * q1 + q2 * q2);
            }
            tdx = 0.125 * (dx / s->g3rm[nxm]) * (dx / s->g3rm[nxm]);
            z1 = vz[B(nxm, j0, ip)] - vz[B(nxm - 1, j0, ip)] + s->c.inslwn * 2.0 * (0.0 - vz[B(nxm, j0, ip)]);
            z2 = vz[B(nxm, j0, i0)] - vz[B(nxm - 1, j0, i0)] + s->c.inslwn * 2.0 * (0.0 - vz[B(nxm, j0, i0)]);
            y1 = vy[B(nxm, jp, i0)] - vy[B(nxm - 1, jp, i0)] + s->c.inslwn * 2.0 * (s->c.xplusu - vy[B(nxm, jp, i0)]);
            y2 = vy[B(nxm, j0, i0)] - vy[B(nxm - 1, j0, i0)] + s->c.inslwn * 2.0 * (s->c.xplusu - vy[B(nxm, j0, i0)]);
            EP[nxm] = EP[nxm] + tdx * (z1 * z1 + z2 * z2) + tdx * (y1 * y1 + y2 * y2);
        }
    }
    /* :219-223 */
    for (int k = 1; k <= nxm; k++) {
        CH[k] = CH[k] * inym * inzm / s->pect;
        EP[k] = EP[k] * inym * inzm / s->ren;
    }
}

/* GlobalQuantities.F90:11-102 : Nu_vol (:79), mean kinetic energy with the same
 * g3rc weights (:55-58), rms velocity (:91) and mean temperature (:89). */
void orc_global_quantities(orc_state *s, double out[4]) {
    BASE_DECL
    const double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ], *temp = s->f[ORC_TEMP];
    long double anusin = 0, tempm = 0, e = 0;
    double vol = 1.0 / (s->c.alx3 * s->dx * (double)nzm * (double)nym);
    for (int i = 0; i < nzm; i++)
        for (int j = 0; j < nym; j++)
            for (int kc = 1; kc <= nxm; kc++) {
                double fac2 = s->g3rc[kc];
                double vxcen = g_plant == 5 ? vx[B(kc, j, i)] : (vx[B(kc, j, i)] + vx[B(kc + 1, j, i)]) * 0.5, t = temp[B(kc, j, i)];
                anusin += (long double)(t * vxcen * fac2);
                tempm += (long double)(t * fac2);
                e += (long double)(fac2 * (vz[B(kc, j, i)] * vz[B(kc, j, i)] + (g_plant == 4 ? 0.0 : vy[B(kc, j, i)] * vy[B(kc, j, i)]) +
                                           vx[B(kc, j, i)] * vx[B(kc, j, i)]));
            }
    out[0] = 1.0 + sqrt(s->c.prat * fabs(s->c.rayt)) * (double)anusin * vol;
    out[1] = 0.5 * (double)e * vol;
    out[2] = sqrt((double)e * vol) * sqrt(fabs(s->c.rayt) / s->c.prat);
    out[3] = (double)tempm * vol;
}

/* ------------------------------------------------------------------------ */
/* restart onto a new grid: ReadFlowInterp.F90:90-141 and                   */
/* src/multires/IC_interpolation/ (CreateOldGrid, CreateNewInputStencil,    */
/* CreateNewSalStencil, InterpInputVel, InterpInputSal, InterpInputPhi)     */
/* ------------------------------------------------------------------------ */
/* CreateOldGrid.F90:11-115, one wall-normal grid of the old run: the options
 * of CreateGrid with the old stretching str3o (option 3 with parameter 1.0
 * for both grids).  The old base grid takes option 5 when the *current*
 * istr3 is 5 (CreateOldGrid.F90:77); an old option 5 without it leaves the
 * reference's grid unset, which is refused here.  c(1:Nm+1), m(1:Nm+1). */
static int old_xgrid(orc_state *s, int istr3o, double str3o, double *c, double *m, int Nm, int refined) {
    double L = s->c.alx3;
    int opt = istr3o;
    if (!refined && s->c.istr3 == 5) opt = 5;
    else if (!refined && istr3o == 5) return -1;
    switch (opt) {
    case 0: uniform_grid(c, m, Nm, L); break;
    case 1: centre_focus_grid(c, m, Nm, L, str3o); break;
    case 2: natural_BL_grid(c, m, Nm, L); break;
    case 3: sym_natural_BL_grid(c, m, Nm, L, 1.0); break;
    case 4: tanh_grid(c, m, Nm, L, str3o); break;
    case 5: scallop_grid(c, m, Nm, L, s->c.dpdy, 0.5); break;
    case 6: cheb_grid(c, m, Nm, L, str3o, s->pi); break;
    case 7: asym_cheb_grid(c, m, Nm, L, str3o, s->pi); break;
    case 8: cosine_grid(c, m, Nm, L, s->pi); break;
    default: return -1;
    }
    m[Nm + 1] = 2 * c[Nm + 1] - m[Nm];     /* xmo(nxo) = 2 xco(nxo) - xmo(nxmo) */
    return 0;
}

/* old -> new tables along one axis: idx(0:n_old+1) and the weights cx(4, n_new) */
typedef struct { int *idx; double *w; } axis_tab;
static axis_tab make_axis(const double *xold, int n_old, const double *xnew, int n_new, double L, char axis) {
    axis_tab t;
    t.idx = malloc(sizeof(int) * (size_t)(n_old + 2));
    t.w = malloc(sizeof(double) * 4 * (size_t)n_new);
    interpolation_indices(t.idx, xold, n_old, xnew, n_new, L);
    construct_stencil(t.w, xold, n_old, xnew, n_new, L, t.idx, axis);
    return t;
}
static void free_axis(axis_tab *t) { free(t->idx); free(t->w); }

/* interpolate_xyz_old_to_new(_ref), HermiteInterpolations.F90:314-380: every
 * old cell (kc = 0..nxo-1, jc = 0..nyo, ic = 0..nzo with the periodic halo
 * cell 0) spreads its 4x4x4 block ovar(kc-1:kc+2, ...) to the new centres it
 * owns.  ovar: (nzo, nyo, nxo+3) with x index -1..nxo+1 (ghosts set by the
 * caller); nvar: (nzn, nyn, ldn), x index 1..nxn at offset 0. */
static void old_to_new(const double *ovar, int nxo_m, int nyo, int nzo, const axis_tab *tx, const axis_tab *ty,
                       const axis_tab *tz, double *nvar, int nxn, int nyn, int nzn, int ldn) {
    int lo = nxo_m + 4;          /* x index -1..nxo+1 with nxo = nxo_m + 1 */
#define OV(k, j, i) ovar[((size_t)W((i), nzo) * nyo + W((j), nyo)) * lo + (k) + 1]
    for (int ic = 0; ic <= nzo; ic++) {
        int ilo = tz->idx[ic] > 1 ? tz->idx[ic] : 1, ihi = tz->idx[ic + 1] - 1 < nzn ? tz->idx[ic + 1] - 1 : nzn;
        for (int jc = 0; jc <= nyo; jc++) {
            int jlo = ty->idx[jc] > 1 ? ty->idx[jc] : 1, jhi = ty->idx[jc + 1] - 1 < nyn ? ty->idx[jc + 1] - 1 : nyn;
            for (int kc = 0; kc <= nxo_m; kc++) {
                int klo = tx->idx[kc] > 1 ? tx->idx[kc] : 1, khi = tx->idx[kc + 1] - 1 < nxn ? tx->idx[kc + 1] - 1 : nxn;
                if (ilo > ihi || jlo > jhi || klo > khi) continue;
                double qv3[4][4][4];
                for (int c = 0; c < 4; c++)
                    for (int b = 0; b < 4; b++)
                        for (int a = 0; a < 4; a++) qv3[c][b][a] = OV(kc - 1 + a, jc - 1 + b, ic - 1 + c);
                for (int icr = ilo; icr <= ihi; icr++) {
                    const double *cz = tz->w + (size_t)(icr - 1) * 4;
                    double qv2[4][4];
                    for (int b = 0; b < 4; b++)
                        for (int a = 0; a < 4; a++)
                            qv2[b][a] = qv3[0][b][a] * cz[0] + qv3[1][b][a] * cz[1] + qv3[2][b][a] * cz[2] + qv3[3][b][a] * cz[3];
                    for (int jcr = jlo; jcr <= jhi; jcr++) {
                        const double *cy = ty->w + (size_t)(jcr - 1) * 4;
                        double qv1[4];
                        for (int a = 0; a < 4; a++) qv1[a] = qv2[0][a] * cy[0] + qv2[1][a] * cy[1] + qv2[2][a] * cy[2] + qv2[3][a] * cy[3];
                        for (int kcr = klo; kcr <= khi; kcr++) {
                            const double *cx = tx->w + (size_t)(kcr - 1) * 4;
                            nvar[((size_t)(icr - 1) * nyn + (jcr - 1)) * ldn + kcr - 1] =
                                qv1[0] * cx[0] + qv1[1] * cx[1] + qv1[2] * cx[2] + qv1[3] * cx[3];
                        }
                    }
                }
            }
        }
    }
#undef OV
}

/* the old grids (CreateOldGrid) and the old -> new cell-centre tables of the
 * base grid (CreateNewInputStencil.F90) or of the refined grid
 * (CreateNewSalStencil.F90) */
typedef struct {
    int nxm, nym, nzm;                 /* old cell counts */
    double *xc, *xm, *yc, *ym, *zm;    /* old grid, 1-based (xm to nxm+1; ym, zm 0..n+1) */
    axis_tab tx, ty, tz;               /* old centres -> new centres */
} old_grid;

static void free_old(old_grid *g) {
    free(g->xc); free(g->xm); free(g->yc); free(g->ym); free(g->zm);
    free_axis(&g->tx); free_axis(&g->ty); free_axis(&g->tz);
}

static int make_old(orc_state *s, const orc_old_grid *og, int refined, old_grid *g) {
    int nx = refined ? og->nxro : og->nxo, ny = refined ? og->nyro : og->nyo, nz = refined ? og->nzro : og->nzo;
    g->nxm = nx - 1; g->nym = ny - 1; g->nzm = nz - 1;
    if (g->nxm < 2 || g->nym < 1 || g->nzm < 1) return -1;
    g->xc = calloc((size_t)nx + 2, sizeof(double)); g->xm = calloc((size_t)nx + 2, sizeof(double));
    g->yc = calloc((size_t)ny + 2, sizeof(double)); g->ym = calloc((size_t)ny + 2, sizeof(double));
    g->zm = calloc((size_t)nz + 2, sizeof(double));
    double *zc = calloc((size_t)nz + 2, sizeof(double));
    if (old_xgrid(s, refined ? og->istr3ro : og->istr3o, og->str3o, g->xc, g->xm, g->nxm, refined)) { free(zc); return -2; }
    uniform_grid(g->yc, g->ym, g->nym, s->c.ylen);
    uniform_grid(zc, g->zm, g->nzm, s->c.zlen);
    free(zc);
    const double *xn = refined ? s->xmr : s->xm, *yn = refined ? s->ymr : s->ym, *zn = refined ? s->zmr : s->zm;
    int nxn = refined ? s->nxmr : s->nxm, nyn = refined ? s->nymr : s->nym, nzn = refined ? s->nzmr : s->nzm;
    g->tx = make_axis(g->xm, g->nxm, xn, nxn, s->c.alx3, 'x');
    g->ty = make_axis(g->ym, g->nym, yn, nyn, s->c.ylen, 'y');
    g->tz = make_axis(g->zm, g->nzm, zn, nzn, s->c.zlen, 'z');
    return 0;
}

/* InterpInputVel.F90:1-350.  vxo, vyo, vzo, tempo: the old fields as the
 * checkpoint holds them, (nzo-1, nyo-1, nxo) x fastest.  The gradients of
 * the old velocity are interpolated to the new cell centres and integrated
 * (vx from the wall, vy / vz from a first-guess plane j = 1 / i = 1 made by
 * 2-D Hermite interpolation of the old plane); T is interpolated with its
 * wall ghosts (:280-330).  Returns 0, or nonzero for an old grid option the
 * reference cannot build. */
int orc_interp_input_vel(orc_state *s, const orc_old_grid *og, const double *vxo, const double *vyo,
                         const double *vzo, const double *tempo) {
    BASE_DECL
    old_grid g;
    if (make_old(s, og, 0, &g)) return 1;
    int nxmo = g.nxm, nymo = g.nym, nzmo = g.nzm, nxo = nxmo + 1, lo = nxmo + 4;
    int sS = 1 - 2 * s->c.inslws, sN = 1 - 2 * s->c.inslwn;
    double *ov = calloc((size_t)lo * nymo * nzmo, sizeof(double));    /* x -1..nxo+1 */
    double *tp = calloc((size_t)nx * nym * nzm, sizeof(double));      /* new centres, x 1..nxm */
#define O3(a, k, j, i) (a)[((size_t)(i) * nymo + (j)) * nxo + (k) - 1]
#define OE(k, j, i) ov[((size_t)(i) * nymo + (j)) * lo + (k) + 1]
    double *vx = s->f[ORC_VX], *vy = s->f[ORC_VY], *vz = s->f[ORC_VZ], *temp = s->f[ORC_TEMP];
    memset(vx, 0, sizeof(double) * nbase(s)); memset(vy, 0, sizeof(double) * nbase(s)); memset(vz, 0, sizeof(double) * nbase(s));
    double dyo = 1.0 / (g.yc[2] - g.yc[1]), dzo = (double)nzmo / s->c.zlen;   /* dzo = 1/(zco(2)-zco(1)) */

    /* vx: dvx/dx at the old centres (:49-64), interpolated (:68), integrated from the wall (:71-79) */
    memset(ov, 0, sizeof(double) * (size_t)lo * nymo * nzmo);
    for (int i = 0; i < nzmo; i++)
        for (int j = 0; j < nymo; j++) {
            for (int kc = 1; kc <= nxmo; kc++)
                OE(kc, j, i) = (O3(vxo, kc + 1, j, i) - O3(vxo, kc, j, i)) / (g.xc[kc + 1] - g.xc[kc]);
            OE(0, j, i) = sS * OE(1, j, i);
            OE(nxo, j, i) = sN * OE(nxmo, j, i);
        }
    old_to_new(ov, nxmo, nymo, nzmo, &g.tx, &g.ty, &g.tz, tp, nxm, nym, nzm, nx);
    for (int i = 0; i < nzm; i++)
        for (int j = 0; j < nym; j++) {
            vx[B(1, j, i)] = 0.0;
            for (int kc = 1; kc <= nxm; kc++)
                vx[B(kc + 1, j, i)] = vx[B(kc, j, i)] + tp[((size_t)i * nym + j) * nx + kc - 1] * (s->xc[kc + 1] - s->xc[kc]);
        }

    /* vy: dvy/dy (:104-118), interpolated (:122), first-guess plane at the
     * first y face from the old plane jc0 = 1 (:125-156), integrated along y
     * (:161-168; the value carried past nym lands in the halo) */
    memset(ov, 0, sizeof(double) * (size_t)lo * nymo * nzmo);
    for (int i = 0; i < nzmo; i++)
        for (int j = 0; j < nymo; j++) {
            int jp = (j + 1) % nymo;
            for (int kc = 1; kc <= nxmo; kc++) OE(kc, j, i) = (O3(vyo, kc, jp, i) - O3(vyo, kc, j, i)) * dyo;
            OE(0, j, i) = sS * OE(1, j, i);
            OE(nxo, j, i) = sN * OE(nxmo, j, i);
        }
    old_to_new(ov, nxmo, nymo, nzmo, &g.tx, &g.ty, &g.tz, tp, nxm, nym, nzm, nx);
    {
        /* vyxzc(-1:nxmo+2, z) with the wall ghosts (:132-141) */
        double *pl = calloc((size_t)lo * nzmo, sizeof(double));
#define PL(k, i0) pl[(size_t)W((i0), nzmo) * lo + (k) + 1]
        for (int i = 1; i <= nzmo; i++) {
            for (int kc = 1; kc <= nxmo; kc++) PL(kc, i) = O3(vyo, kc, 0, i - 1);
            PL(0, i) = sS * PL(1, i);
            PL(nxo, i) = sN * PL(nxmo, i);
        }
        for (int ic = 0; ic <= nzmo; ic++) {
            int ilo = g.tz.idx[ic] > 1 ? g.tz.idx[ic] : 1, ihi = g.tz.idx[ic + 1] - 1 < nzm ? g.tz.idx[ic + 1] - 1 : nzm;
            for (int kc = 0; kc <= nxmo; kc++) {
                int klo = g.tx.idx[kc] > 1 ? g.tx.idx[kc] : 1, khi = g.tx.idx[kc + 1] - 1 < nxm ? g.tx.idx[kc + 1] - 1 : nxm;
                double qv2[4][4];   /* [c][a] = vyxzc(kc-1+a, ic-1+c) */
                for (int c = 0; c < 4; c++) for (int a = 0; a < 4; a++) qv2[c][a] = PL(kc - 1 + a, ic - 1 + c);
                for (int icr = ilo; icr <= ihi; icr++) {
                    const double *cz = g.tz.w + (size_t)(icr - 1) * 4;      /* czvy = czvx */
                    double qv1[4];
                    for (int a = 0; a < 4; a++) qv1[a] = qv2[0][a] * cz[0] + qv2[1][a] * cz[1] + qv2[2][a] * cz[2] + qv2[3][a] * cz[3];
                    for (int kcr = klo; kcr <= khi; kcr++) {
                        const double *cx = g.tx.w + (size_t)(kcr - 1) * 4;  /* cxvy */
                        vy[B(kcr, 0, icr - 1)] = qv1[0] * cx[0] + qv1[1] * cx[1] + qv1[2] * cx[2] + qv1[3] * cx[3];
                    }
                }
            }
        }
#undef PL
        free(pl);
        double lya = 1.0 / s->dy;
        for (int i = 0; i < nzm; i++)
            for (int j = 1; j < nym; j++)
                for (int kc = 1; kc <= nxm; kc++)
                    vy[B(kc, j, i)] = vy[B(kc, j - 1, i)] + tp[((size_t)i * nym + j - 1) * nx + kc - 1] * lya;
    }

    /* vz: dvz/dz (:196-210), interpolated (:214), first-guess plane at the
     * first z face from the old plane ic0 = 1 (:217-248), integrated along z (:253-261) */
    memset(ov, 0, sizeof(double) * (size_t)lo * nymo * nzmo);
    for (int i = 0; i < nzmo; i++) {
        int ip = (i + 1) % nzmo;
        for (int j = 0; j < nymo; j++) {
            for (int kc = 1; kc <= nxmo; kc++) OE(kc, j, i) = (O3(vzo, kc, j, ip) - O3(vzo, kc, j, i)) * dzo;
            OE(0, j, i) = sS * OE(1, j, i);
            OE(nxo, j, i) = sN * OE(nxmo, j, i);
        }
    }
    old_to_new(ov, nxmo, nymo, nzmo, &g.tx, &g.ty, &g.tz, tp, nxm, nym, nzm, nx);
    {
        double *pl = calloc((size_t)lo * nymo, sizeof(double));
#define PL(k, j0) pl[(size_t)W((j0), nymo) * lo + (k) + 1]
        for (int j = 1; j <= nymo; j++) {
            for (int kc = 1; kc <= nxmo; kc++) PL(kc, j) = O3(vzo, kc, j - 1, 0);
            PL(0, j) = sS * PL(1, j);
            PL(nxo, j) = sN * PL(nxmo, j);
        }
        for (int jc = 0; jc <= nymo; jc++) {
            int jlo = g.ty.idx[jc] > 1 ? g.ty.idx[jc] : 1, jhi = g.ty.idx[jc + 1] - 1 < nym ? g.ty.idx[jc + 1] - 1 : nym;
            for (int kc = 0; kc <= nxmo; kc++) {
                int klo = g.tx.idx[kc] > 1 ? g.tx.idx[kc] : 1, khi = g.tx.idx[kc + 1] - 1 < nxm ? g.tx.idx[kc + 1] - 1 : nxm;
                double qv2[4][4];   /* [b][a] = vzxyc(kc-1+a, jc-1+b) */
                for (int b = 0; b < 4; b++) for (int a = 0; a < 4; a++) qv2[b][a] = PL(kc - 1 + a, jc - 1 + b);
                for (int jcr = jlo; jcr <= jhi; jcr++) {
                    const double *cy = g.ty.w + (size_t)(jcr - 1) * 4;      /* cyvz = cyvx */
                    double qv1[4];
                    for (int a = 0; a < 4; a++) qv1[a] = qv2[0][a] * cy[0] + qv2[1][a] * cy[1] + qv2[2][a] * cy[2] + qv2[3][a] * cy[3];
                    for (int kcr = klo; kcr <= khi; kcr++) {
                        const double *cx = g.tx.w + (size_t)(kcr - 1) * 4;  /* cxvz */
                        vz[B(kcr, jcr - 1, 0)] = qv1[0] * cx[0] + qv1[1] * cx[1] + qv1[2] * cx[2] + qv1[3] * cx[3];
                    }
                }
            }
        }
#undef PL
        free(pl);
        double lza = 1.0 / s->dz;
        for (int i = 1; i < nzm; i++)
            for (int j = 0; j < nym; j++)
                for (int kc = 1; kc <= nxm; kc++)
                    vz[B(kc, j, i)] = vz[B(kc, j, i - 1)] + tp[((size_t)(i - 1) * nym + j) * nx + kc - 1] * lza;
    }

    /* T with the wall ghosts of its boundary values (:280-330) */
    double tup, tlo;
    if (s->c.rayt > 0) {
        if (s->c.inslwn == 0) { tup = 0.0; tlo = 1.0; } else { tup = -0.5; tlo = 0.5; }
    } else { tup = 0.5; tlo = -0.5; }
    if (s->c.phasefield) {
        if (s->c.rayt > 0) { tup = 0.0; tlo = 1.0; } else { tup = 1.0; tlo = 0.0; }
    }
    memset(ov, 0, sizeof(double) * (size_t)lo * nymo * nzmo);
    for (int i = 0; i < nzmo; i++)
        for (int j = 0; j < nymo; j++) {
            for (int kc = 1; kc <= nxo; kc++) OE(kc, j, i) = O3(tempo, kc, j, i);
            OE(0, j, i) = s->c.tfixs == 1 ? 2.0 * tlo - OE(1, j, i) : OE(1, j, i);
            OE(nxo, j, i) = s->c.tfixn == 1 ? 2.0 * tup - OE(nxmo, j, i) : OE(nxmo, j, i);
        }
    memset(temp, 0, sizeof(double) * nbase(s));
    old_to_new(ov, nxmo, nymo, nzmo, &g.tx, &g.ty, &g.tz, temp, nxm, nym, nzm, nx);
#undef O3
#undef OE
    free(ov); free(tp);
    free_old(&g);
    return 0;
}

/* InterpInputSal.F90 (which = 0) and InterpInputPhi.F90 (which = 1): the old
 * refined scalar (nzro-1, nyro-1, nxro) with its wall ghosts, Dirichlet from
 * the sign of Ra_S (phase field: 1 below, 0 above) or zero-gradient for phi,
 * interpolated to the new refined cell centres (CreateNewSalStencil.F90). */
static int interp_input_ref(orc_state *s, const orc_old_grid *og, const double *fo, int which) {
    REF_DECL
    old_grid g;
    if (make_old(s, og, 1, &g)) return 1;
    int nxmo = g.nxm, nymo = g.nym, nzmo = g.nzm, nxo = nxmo + 1, lo = nxmo + 4;
    double *ov = calloc((size_t)lo * nymo * nzmo, sizeof(double));
    double sup, slo;
    if (s->c.rays > 0) { sup = 0.5; slo = -0.5; } else { sup = -0.5; slo = 0.5; }
    if (s->c.phasefield) { sup = 0.0; slo = 1.0; }
    for (int i = 0; i < nzmo; i++)
        for (int j = 0; j < nymo; j++) {
            double *row = ov + ((size_t)i * nymo + j) * lo + 1;
            const double *src = fo + ((size_t)i * nymo + j) * nxo;
            for (int kc = 1; kc <= nxo; kc++) row[kc] = src[kc - 1];
            if (which == 0) {
                row[0] = s->c.sfixs == 1 ? 2.0 * slo - row[1] : row[1];
                row[nxo] = s->c.sfixn == 1 ? 2.0 * sup - row[nxmo] : row[nxmo];
            } else {
                row[0] = row[1];
                row[nxo] = row[nxmo];
            }
        }
    double *f = s->f[which == 0 ? ORC_SAL : ORC_PHI];
    memset(f, 0, sizeof(double) * nref(s));
    old_to_new(ov, nxmo, nymo, nzmo, &g.tx, &g.ty, &g.tz, f, nxmr, nymr, nzmr, nxr);
    free(ov);
    free_old(&g);
    return 0;
}
/* InterpInputHum, moisture.F90:539-582: the old humidity (nzo-1, nyo-1, nxo)
 * on the base grid with Dirichlet (qfixS / qfixN) or zero-gradient ghosts;
 * the wall values are those of the first column of humbp / humtp (the
 * routine assumes them uniform, :557-559, after SetHumidityBCs). */
int orc_interp_input_humid(orc_state *s, const orc_old_grid *og, const double *qo) {
    BASE_DECL
    if (!s->c.moist) return 1;
    old_grid g;
    if (make_old(s, og, 0, &g)) return 1;
    int nxmo = g.nxm, nymo = g.nym, nzmo = g.nzm, nxo = nxmo + 1, lo = nxmo + 4;
    double *ov = calloc((size_t)lo * nymo * nzmo, sizeof(double));
    const double qup = s->f[ORC_HUMTP][0], qlo = s->f[ORC_HUMBP][0];
    for (int i = 0; i < nzmo; i++)
        for (int j = 0; j < nymo; j++) {
            double *row = ov + ((size_t)i * nymo + j) * lo + 1;
            const double *src = qo + ((size_t)i * nymo + j) * nxo;
            for (int kc = 1; kc <= nxo; kc++) row[kc] = src[kc - 1];
            row[0] = s->q_fixs == 1 ? 2.0 * qlo - row[1] : row[1];
            row[nxo] = s->q_fixn == 1 ? 2.0 * qup - row[nxmo] : row[nxmo];
        }
    double *q = s->f[ORC_HUMID];
    memset(q, 0, sizeof(double) * nbase(s));
    old_to_new(ov, nxmo, nymo, nzmo, &g.tx, &g.ty, &g.tz, q, nxm, nym, nzm, nx);
    memset(s->f[ORC_RUHUMID], 0, sizeof(double) * nbase(s));
    free(ov);
    free_old(&g);
    return 0;
}
int orc_interp_input_sal(orc_state *s, const orc_old_grid *og, const double *salo) { return interp_input_ref(s, og, salo, 0); }
int orc_interp_input_phi(orc_state *s, const orc_old_grid *og, const double *phio) { return interp_input_ref(s, og, phio, 1); }
