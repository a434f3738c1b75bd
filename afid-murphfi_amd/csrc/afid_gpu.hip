// afid_gpu.hip -- context, set-up and C ABI of the MI355X RK3 substep.
//
// Build (gfx950 only):
//   hipcc --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared
//         -Xclang -target-feature -Xclang -packed-fp32-ops
//         -I include -o afid-murphfi_amd/libafid_gpu.so afid-murphfi_amd/csrc/afid_gpu.hip
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "afid_gpu.h"
#include "afid_setup.hpp"
#include "afid_params.hpp"
#include "afid_stencils.hpp"
#include "afid_kernels.hpp"
#include "afid_poisson.hpp"
#include "afid_phase.hpp"
#include "afid_means.hpp"
#include "afid_restart.hpp"
#include "afid_ibm.hpp"
#include "afid_moist.hpp"

using namespace afid;

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct TimerRec {
    std::string name;
    hipEvent_t a, b;
};

}  // namespace

// Refined -> coarse tables of one transfer (k_interp_coarse_tile), per axis
// and coarse index: the first refined source and 4 weights (FP64, and FP32
// for the mixed form), with the tile's LDS plan.
struct CoarseTab {
    const int *x_src = nullptr, *y_src = nullptr, *z_src = nullptr;
    const double *x_w = nullptr, *y_w = nullptr, *z_w = nullptr;
    const float *x_wf = nullptr, *y_wf = nullptr, *z_wf = nullptr;
    int xr = 0, yr = 0;          // largest refined x / y span of a tile
    size_t lds = 0;
    bool general = false;        // interpolate_xyz_to_coarse (refined grid does not bracket)
};

struct afid_ctx {
    afid_config c{};
    std::string err;
    hipStream_t stream = nullptr;
    bool own_stream = true;            // false: the stream belongs to an ensemble
    int device = 0;
    // sizes
    int nxm, nx, ny, nz, ldx, ldt, nymh;
    int nxmr = 0, nxr = 0, nyr = 0, nzr = 0, ldxr = 0;
    bool sal = false;
    // physics
    double ren, pect, pecs, byct, bycs, Gy, rrho;
    double dx, dy, dz, dxr = 0, dyr = 0, dzr = 0;
    double gam[3], rom[3], alm[3];
    // substep state
    double dt = 0, al = 0, ga = 0, ro = 0, beta = 0;
    int ns = 0;
    // device memory
    std::vector<DevBuf> bufs;
    int64_t bytes = 0;
    double *f[AFID_NFIELDS] = {nullptr};
    float *g[AFID_NFIELDS] = {nullptr};   // FP32 storage of a field (mixed form)
    double *rhs4[4] = {nullptr};       // rhs for vx, vy, vz, temp
    double *alt[4] = {nullptr};        // second buffers of vx, vy, vz, temp (fused column step)
    double *sal_alt = nullptr;
    float *sal_alt_f = nullptr;
    double *tpdv = nullptr;
    void *spec = nullptr;              // cpx<double> (FP64 form) or cpx<float> (mixed form)
    // phase field
    bool pf = false;
    double pf_d = 0, pf_eps = 0, pf_b = 0, heta = 0, pf_r0 = 0.1;
    std::vector<double> h_ap3spkr, h_ac3spkr, h_am3spkr;
    const double *d_ap3spkr = nullptr, *d_ac3spkr = nullptr, *d_am3spkr = nullptr;
    double *phi_alt = nullptr;
    float *phi_alt_f = nullptr;                 // second phi buffer of the mixed form
    CoarseTab tab_sal, tab_phi;                 // InterpSalMultigrid / InterpPhiMultigrid + AddLatentHeat
    // moisture (moisture.F90:15-23): the parameters after ReadMoistParameters'
    // derivations, the wall flags after SetHumidityBCs, the second humidity
    // buffer of the fused column kernel
    bool moist = false;
    double q_alpha = 0, q_beta = 0, q_gamma = 0, q_tau = 0, pecq = 0, kappa_q = 0;
    int q_fixn = 1, q_fixs = 1;
    double *humid_alt = nullptr;
    double *d_moist_part = nullptr, *d_moist_out = nullptr;    // CalcMoistStats partial sums and profiles
    // arithmetic form
    bool mixed = false;
    double div_tol = 1e-10;
    int max_sweeps = 3;
    unsigned long long *slots = nullptr;   // refinement residuals of the last substep (bit patterns)
    // metrics
    std::map<std::string, std::vector<double>> host1d;   // for afid_grid_array
    const double *d_xm = nullptr, *d_xmr = nullptr;
    const double *d_udx3m, *d_udx3c, *d_rudx3m, *d_ap3ck, *d_ac3ck, *d_am3ck, *d_ap3sk, *d_ac3sk, *d_am3sk;
    const double *d_ap3ssk, *d_ac3ssk, *d_am3ssk, *d_amphk, *d_acphk, *d_apphk, *d_g3rc;
    // CalcMeanProfiles: dx/g3rm per cell, the per-group partial sums and the profiles
    const double *d_sdx = nullptr, *d_sdxr = nullptr;
    double *d_mean_part = nullptr, *d_mean_out = nullptr, *d_sal_part = nullptr, *d_sal_out = nullptr;
    double *d_phi_part = nullptr, *d_phi_out = nullptr;
    int mean_groups = 0, sal_groups = 0, phi_groups = 0;
    bool col_four_pass = false;         // base columns through the four-pass kernel (AFID_COLUMN_FOUR_PASS=1)
    bool vyz_tile = false;              // vyr / vzr through the LDS tile kernel (AFID_VYZ_TILE=1)
    // mixed form, gAxis = 1: the FP32 vx explicit term and its history hold
    // the buoyancy minus a per-row FP64 reference bref (that of the
    // conductive profile between the mean wall temperatures), which the FP64
    // RHS adds back; recomputed at the first substep after a wall upload
    // immersed boundary: per solve (vx, vy, vz, T, S) the row mask and the
    // per-cell boundary distances on the device (null: the plain solve)
    signed char *ibm_mask[5] = {nullptr};
    double *ibm_dist[5] = {nullptr};
    std::vector<double> h_bref;
    double *d_bref = nullptr;
    double wall_mean[2] = {0.0, 0.0};
    bool bref_dirty = false;
    // afid_time_marcher_graph: two RK3 steps captured as one graph (the
    // ping-pong buffers swap an odd number of times per step, so two steps
    // bring every pointer back), valid for graph_dt and the pointers it
    // captured
    hipGraphExec_t graph = nullptr;
    double graph_dt = 0.0;
    double *graph_ptrs[AFID_NFIELDS] = {nullptr};
    const double *d_udx3mr = nullptr, *d_ap3sskr = nullptr, *d_ac3sskr = nullptr, *d_am3sskr = nullptr, *d_lxa = nullptr;
    const int *d_xs_src = nullptr, *d_ys_src = nullptr, *d_zs_src = nullptr;
    const double *d_xs_w = nullptr, *d_ys_w = nullptr, *d_zs_w = nullptr;
    const float *f_xs_w = nullptr, *f_ys_w = nullptr, *f_zs_w = nullptr;
    const float *f_lxa = nullptr;
    // host copies of the 1-D coefficients for the tridiagonal factors
    std::vector<double> h_ap3ck, h_ac3ck, h_am3ck, h_ap3sk, h_ac3sk, h_am3sk, h_ap3ssk, h_ac3ssk, h_am3ssk;
    std::vector<double> h_ap3sskr, h_ac3sskr, h_am3sskr;
    // tridiagonal factors: [ns][kind] -> 4 arrays of npad
    int chunk[MAT_NKIND] = {0}, nrows[MAT_NKIND] = {0};
    double *tri = nullptr;          // device, 3*4*4*npadmax doubles
    int npadmax = 0;
    double fact_dt = -1.0;
    std::vector<double> tri_host;
    // Poisson
    FftPlan<double> plan_y, plan_z;
    FftPlan<float> plan_yf, plan_zf;
    int Ly = 1, Lz = 1, lds_y = 0, lds_z = 0;
    int fftz_maxv = FFT_MAXV;           // values per thread and stage of k_fft_z
    bool fft_z16 = true;                // FP32 z transform may take 16 values per thread (AFID_FFT_Z16=0: no)
    PoissonFactors<double> pois{};
    PoissonFactors<float> pff{};
    int pchunk = 1;
    // LDS plan of the transfer kernels
    int vxr_xc = 0, vxr_cj = 0, vxr_ci = 0, vy_cs = 0, vy_co = 0, vz_cs = 0, vz_co = 0;
    size_t lds_vxr = 0, lds_vy = 0, lds_vz = 0;
    // reductions
    unsigned long long *redmax = nullptr;
    double *redsum = nullptr, *redout = nullptr;
    long long nred = 0;
    // timing
    bool timing = false;
    std::vector<TimerRec> timers;
    std::vector<std::pair<std::string, double>> last_timing;
};

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
#define HIPCHK(ctx, expr)                                                              \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);           \
            return AFID_E_HIP;                                                         \
        }                                                                              \
    } while (0)

static int set_err(afid_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

static int dalloc(afid_ctx *ctx, void **p, size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        ctx->err = "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e);
        return AFID_E_MEM;
    }
    ctx->bufs.push_back({*p, bytes});
    ctx->bytes += int64_t(bytes);
    e = hipMemsetAsync(*p, 0, bytes, ctx->stream);
    if (e != hipSuccess) { ctx->err = "hipMemset failed"; return AFID_E_HIP; }
    return AFID_OK;
}

template <typename T>
static int upload_vec(afid_ctx *ctx, const std::vector<T> &v, const T **out) {
    void *p = nullptr;
    int rc = dalloc(ctx, &p, sizeof(T) * v.size());
    if (rc) return rc;
    HIPCHK(ctx, hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    *out = static_cast<const T *>(p);
    return AFID_OK;
}

static std::vector<double> farr_range(const FArr &a, int lo, int hi) {
    std::vector<double> v;
    for (int i = lo; i <= hi; ++i) v.push_back(a(i));
    return v;
}

static bool is_refined(int f) {
    return (f >= AFID_SAL && f <= AFID_VZR) || f == AFID_SALBP || f == AFID_SALTP ||
           f == AFID_PHI || f == AFID_RUPHI || f == AFID_HPHI || f == AFID_TEMPR;
}
static bool is_plane(int f) {
    return f == AFID_TEMPBP || f == AFID_TEMPTP || f == AFID_SALBP || f == AFID_SALTP || f == AFID_HUMBP || f == AFID_HUMTP;
}
static bool is_moist_field(int f) { return f >= AFID_HUMID && f <= AFID_HUMTP; }

static KP make_kp(const afid_ctx *c) {
    KP p{};
    p.nxm = c->nxm; p.nx = c->nx; p.ny = c->ny; p.nz = c->nz; p.ldx = c->ldx;
    p.plane = (long long)c->ny * c->ldx; p.nymh = c->nymh; p.ldt = c->ldt;
    p.nxmr = c->nxmr; p.nxr = c->nxr; p.nyr = c->nyr; p.nzr = c->nzr; p.ldxr = c->ldxr;
    p.planer = (long long)c->nyr * c->ldxr;
    p.dt = c->dt; p.al = c->al; p.ga = c->ga; p.ro = c->ro; p.beta = c->beta;
    p.ren = c->ren; p.pect = c->pect; p.pecs = c->pecs; p.byct = c->byct; p.bycs = c->bycs; p.Gy = c->Gy;
    p.dx = c->dx; p.dy = c->dy; p.dz = c->dz; p.dxq = c->dx * c->dx; p.dyq = c->dy * c->dy; p.dzq = c->dz * c->dz;
    p.dxr = c->dxr; p.dyr = c->dyr; p.dzr = c->dzr; p.dxqr = c->dxr * c->dxr; p.dyqr = c->dyr * c->dyr; p.dzqr = c->dzr * c->dzr;
    p.inslws = c->c.inslws; p.inslwn = c->c.inslwn; p.tfixs = c->c.tfixs; p.tfixn = c->c.tfixn;
    p.sfixs = c->c.sfixs; p.sfixn = c->c.sfixn; p.active_t = c->c.active_t; p.active_s = c->sal ? c->c.active_s : 0; p.gaxis = c->c.gaxis;
    p.xplusu = c->c.xplusu; p.xminusu = c->c.xminusu;
    p.vx = c->f[AFID_VX]; p.vy = c->f[AFID_VY]; p.vz = c->f[AFID_VZ]; p.temp = c->f[AFID_TEMP]; p.pr = c->f[AFID_PR];
    p.rux = c->f[AFID_RUX]; p.ruy = c->f[AFID_RUY]; p.ruz = c->f[AFID_RUZ]; p.rut = c->f[AFID_RUTEMP];
    p.qcap = c->f[AFID_QCAP]; p.dph = c->f[AFID_DPH]; p.dq = c->f[AFID_DQ]; p.hro = c->f[AFID_HRO];
    p.rhsx = c->rhs4[0]; p.rhsy = c->rhs4[1]; p.rhsz = c->rhs4[2]; p.rhst = c->rhs4[3];
    p.salc = c->f[AFID_SALC]; p.tpdv = c->tpdv;
    p.tempbp = c->f[AFID_TEMPBP]; p.temptp = c->f[AFID_TEMPTP];
    p.sal = c->f[AFID_SAL]; p.rusal = c->f[AFID_RUSAL]; p.hsal = c->f[AFID_HSAL]; p.rhsr = c->f[AFID_RHSR];
    p.vxr = c->f[AFID_VXR]; p.vyr = c->f[AFID_VYR]; p.vzr = c->f[AFID_VZR];
    p.salbp = c->f[AFID_SALBP]; p.saltp = c->f[AFID_SALTP];
    p.udx3m = c->d_udx3m; p.udx3c = c->d_udx3c; p.rudx3m = c->d_rudx3m;
    p.ap3ck = c->d_ap3ck; p.ac3ck = c->d_ac3ck; p.am3ck = c->d_am3ck;
    p.ap3sk = c->d_ap3sk; p.ac3sk = c->d_ac3sk; p.am3sk = c->d_am3sk;
    p.ap3ssk = c->d_ap3ssk; p.ac3ssk = c->d_ac3ssk; p.am3ssk = c->d_am3ssk;
    p.amphk = c->d_amphk; p.acphk = c->d_acphk; p.apphk = c->d_apphk; p.g3rc = c->d_g3rc;
    p.udx3mr = c->d_udx3mr; p.ap3sskr = c->d_ap3sskr; p.ac3sskr = c->d_ac3sskr; p.am3sskr = c->d_am3sskr; p.lxa = c->d_lxa;
    p.xs_src = c->d_xs_src; p.ys_src = c->d_ys_src; p.zs_src = c->d_zs_src;
    p.xs_w = c->d_xs_w; p.ys_w = c->d_ys_w; p.zs_w = c->d_zs_w;
    p.redmax = c->redmax; p.redsum = c->redsum;
    p.salf = c->g[AFID_SAL]; p.rusalf = c->g[AFID_RUSAL]; p.vxrf = c->g[AFID_VXR]; p.vyrf = c->g[AFID_VYR]; p.vzrf = c->g[AFID_VZR];
    p.ruxf = c->g[AFID_RUX]; p.ruyf = c->g[AFID_RUY]; p.ruzf = c->g[AFID_RUZ]; p.rutf = c->g[AFID_RUTEMP]; p.salcf = c->g[AFID_SALC];
    p.xs_wf = c->f_xs_w; p.ys_wf = c->f_ys_w; p.zs_wf = c->f_zs_w;
    p.lxaf = c->f_lxa;
    p.phi = c->f[AFID_PHI]; p.ruphi = c->f[AFID_RUPHI]; p.hphi = c->f[AFID_HPHI]; p.tempr = c->f[AFID_TEMPR];
    p.phic = c->f[AFID_PHIC]; p.qlat = c->f[AFID_HRO];
    p.phif = c->g[AFID_PHI]; p.ruphif = c->g[AFID_RUPHI]; p.temprf = c->g[AFID_TEMPR]; p.rhsrf = c->g[AFID_RHSR];
    p.dphf = c->g[AFID_DPH];
    p.pf_d = c->pf_d; p.pf_b = c->pf_b; p.pf_a = c->c.pf_a; p.pf_tm = c->c.pf_tm; p.pf_s = c->c.pf_s; p.heta = c->heta;
    p.pf = c->pf ? 1 : 0;
    p.bref = c->d_bref;
    p.ap3spkr = c->d_ap3spkr; p.ac3spkr = c->d_ac3spkr; p.am3spkr = c->d_am3spkr;
    p.humid = c->f[AFID_HUMID]; p.ruhumid = c->f[AFID_RUHUMID]; p.hhumid = c->f[AFID_HHUMID]; p.qsat = c->f[AFID_QSAT];
    p.humbp = c->f[AFID_HUMBP]; p.humtp = c->f[AFID_HUMTP]; p.xm = c->d_xm;
    p.moist = c->moist ? 1 : 0;
    if (c->moist) {
        p.q_alpha = c->q_alpha; p.q_beta = c->q_beta; p.q_gtau = c->q_gamma / c->q_tau; p.q_itau = 1.0 / c->q_tau;
        p.kappa_q = c->kappa_q; p.pecq = c->pecq; p.q_fixn = c->q_fixn; p.q_fixs = c->q_fixs;
    }
    return p;
}

static dim3 point_grid(int nxl, int ny, int nz) { return dim3((nxl + 63) / 64, (ny + 3) / 4, nz); }
static const dim3 kPointBlock(64, 4, 1);
// max-reductions loop over z inside the kernel: few blocks, few atomics
static dim3 red_grid(int nxl, int ny, int nz) { return dim3((nxl + 63) / 64, (ny + 3) / 4, std::min(nz, 8)); }

static void tstart(afid_ctx *c, const char *name) {
    if (!c->timing) return;
    TimerRec r;
    r.name = name;
    (void)hipEventCreate(&r.a);
    (void)hipEventCreate(&r.b);
    (void)hipEventRecord(r.a, c->stream);
    c->timers.push_back(r);
}
static void tstop(afid_ctx *c) {
    if (!c->timing || c->timers.empty()) return;
    (void)hipEventRecord(c->timers.back().b, c->stream);
}

static int check_launch(afid_ctx *c, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        c->err = std::string("launch of ") + what + " failed: " + hipGetErrorString(e);
        return AFID_E_HIP;
    }
    return AFID_OK;
}

// arrays per tridiagonal kind and substep: L, E, G, S (factors) and A, C
// (normalised off-diagonals, for masked solves)
constexpr int kTriArrays = 8;     // l, e, g, s, a, c, ln, sn

// ---------------------------------------------------------------------------
// tridiagonal factors (SolveImpEqnUpdate_X/YZ/Temp/Sal/Phi matrices, computed
// on the host for each (dt, substep) and cached while dt is unchanged)
// ---------------------------------------------------------------------------
static void lu_rows(const std::vector<double> &am, const std::vector<double> &ac, const std::vector<double> &ap,
                    double bd, int n, int r0, int r1, int npad, double *L, double *E, double *G, double *S,
                    double *A, double *C, double *LN, double *SN) {
    // rows r (0-based) in [r0, r1) carry coefficients; others are identity rows with zero RHS
    std::vector<double> a(size_t(npad), 0.0), cc(size_t(npad), 0.0), s(size_t(npad), 0.0);
    for (int r = r0; r < r1; ++r) {
        const double sb = 1.0 / (1.0 - ac[size_t(r)] * bd);
        a[size_t(r)] = -am[size_t(r)] * bd * sb;
        cc[size_t(r)] = -ap[size_t(r)] * bd * sb;
        s[size_t(r)] = 1.0 / (1.0 - ac[size_t(r)] * bd);
    }
    double u_prev = 1.0;
    for (int r = 0; r < npad; ++r) {
        if (r >= n) { L[r] = 0; E[r] = 0; G[r] = 0; S[r] = 0; A[r] = 0; C[r] = 0; LN[r] = 0; SN[r] = 0; continue; }
        const double l = r > 0 ? a[size_t(r)] / u_prev : 0.0;
        const double u = 1.0 - l * (r > 0 ? cc[size_t(r - 1)] : 0.0);
        L[r] = l; E[r] = cc[size_t(r)] / u; G[r] = 1.0 / u; S[r] = s[size_t(r)];
        A[r] = a[size_t(r)]; C[r] = cc[size_t(r)];
        LN[r] = l * u_prev / u;             // g_r l_r / g_{r-1}
        SN[r] = s[size_t(r)] / u;           // g_r s_r
        u_prev = u;
    }
}

static int ensure_factors(afid_ctx *c) {
    if (c->fact_dt == c->dt) return AFID_OK;
    const int np = c->npadmax;
    c->tri_host.assign(size_t(3 * MAT_NKIND * kTriArrays * np), 0.0);
    for (int s = 0; s < c->c.nsst && s < 3; ++s) {
        const double al = c->alm[s], beta = c->dt / c->ren * 0.5;
        for (int kind = 0; kind < MAT_NKIND; ++kind) {
            if (kind == MAT_S && !c->sal) continue;
            if (kind == MAT_P && !c->pf) continue;
            if (kind == MAT_Q && !c->moist) continue;
            double *base = c->tri_host.data() + size_t((s * MAT_NKIND + kind) * kTriArrays * np);
            double *L = base, *E = base + np, *G = base + 2 * np, *S = base + 3 * np, *A = base + 4 * np, *C = base + 5 * np;
            double *LN = base + 6 * np, *SN = base + 7 * np;
            switch (kind) {
            case MAT_X:   // SolveImpEqnUpdate_X.F90:22-54: identity rows 1 and nx
                lu_rows(c->h_am3ck, c->h_ac3ck, c->h_ap3ck, beta * al, c->nx, 1, c->nxm, np, L, E, G, S, A, C, LN, SN);
                break;
            case MAT_YZ:  // SolveImpEqnUpdate_YZ.F90:25-44
                lu_rows(c->h_am3sk, c->h_ac3sk, c->h_ap3sk, beta * al, c->nxm, 0, c->nxm, np, L, E, G, S, A, C, LN, SN);
                break;
            case MAT_T:   // SolveImpEqnUpdate_Temp.F90:25-46
                lu_rows(c->h_am3ssk, c->h_ac3ssk, c->h_ap3ssk, 0.5 * al * c->dt / c->pect, c->nxm, 0, c->nxm, np, L, E, G, S, A, C, LN, SN);
                break;
            case MAT_S:   // salinity.F90:418-440
                lu_rows(c->h_am3sskr, c->h_ac3sskr, c->h_ap3sskr, 0.5 * al * c->dt / c->pecs, c->nxmr, 0, c->nxmr, np, L, E, G, S, A, C, LN, SN);
                break;
            case MAT_P:   // SolveImpEqnUpdate_Phi, phasefield.F90:427-443
                lu_rows(c->h_am3spkr, c->h_ac3spkr, c->h_ap3spkr, 0.5 * al * c->dt * c->pf_d, c->nxmr, 0, c->nxmr, np, L, E, G, S, A, C, LN, SN);
                break;
            case MAT_Q:   // SolveImpEqnUpdate_Humidity, moisture.F90:382-396
                lu_rows(c->h_am3ssk, c->h_ac3ssk, c->h_ap3ssk, 0.5 * al * c->dt / c->pecq, c->nxm, 0, c->nxm, np, L, E, G, S, A, C, LN, SN);
                break;
            }
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->tri, c->tri_host.data(), sizeof(double) * c->tri_host.size(), hipMemcpyHostToDevice));
    c->fact_dt = c->dt;
    return AFID_OK;
}

static TriFactors factors(const afid_ctx *c, int kind) {
    const int np = c->npadmax;
    const double *base = c->tri + size_t(((c->ns - 1) * MAT_NKIND + kind) * kTriArrays * np);
    TriFactors f;
    f.l = base; f.e = base + np; f.g = base + 2 * np; f.s = base + 3 * np; f.a = base + 4 * np; f.c = base + 5 * np;
    f.ln = base + 6 * np; f.sn = base + 7 * np;
    f.n = c->nrows[kind]; f.chunk = c->chunk[kind];
    return f;
}

template <int CH>
static void launch_tri(afid_ctx *c, const TriFactors &f, double *rhs, double *q, int ldx, long long plane, int ny,
                       long long ncols, int u0, int u1, int z0, int z1, int keep) {
    const unsigned nb = unsigned((ncols + 3) / 4);
    hipLaunchKernelGGL(k_tri_update<CH>, dim3(nb), dim3(256), 0, c->stream, f, rhs, q, ldx, plane, ny, ncols, u0, u1, z0, z1, keep);
}

// keep = 1: leave the increment in rhs, as the reference's rhs array holds it
static int run_tri(afid_ctx *c, int kind, double *rhs, double *q, int keep = 1) {
    TriFactors f = factors(c, kind);
    const bool ref = kind == MAT_S || kind == MAT_P;
    const int ldx = ref ? c->ldxr : c->ldx, ny = ref ? c->nyr : c->ny, nz = ref ? c->nzr : c->nz;
    const long long plane = (long long)ny * ldx, ncols = (long long)ny * nz;
    int u0 = 0, u1 = f.n, z0 = -1, z1 = -1;
    if (kind == MAT_X) { u0 = 1; u1 = c->nxm; z0 = 0; z1 = c->nx - 1; }
    switch (f.chunk) {
#define AFID_TRI_CASE(CH) case CH: launch_tri<CH>(c, f, rhs, q, ldx, plane, ny, ncols, u0, u1, z0, z1, keep); break;
        AFID_TRI_CASE(1) AFID_TRI_CASE(2) AFID_TRI_CASE(3) AFID_TRI_CASE(4) AFID_TRI_CASE(5) AFID_TRI_CASE(6)
        AFID_TRI_CASE(7) AFID_TRI_CASE(8) AFID_TRI_CASE(9) AFID_TRI_CASE(10) AFID_TRI_CASE(11) AFID_TRI_CASE(12)
        AFID_TRI_CASE(13) AFID_TRI_CASE(14) AFID_TRI_CASE(15) AFID_TRI_CASE(16)
#undef AFID_TRI_CASE
    default: return set_err(c, AFID_E_ARG, "wall-normal size above 1024 points is not supported");
    }
    return check_launch(c, "k_tri_update");
}

// ---------------------------------------------------------------------------
// set-up
// ---------------------------------------------------------------------------
template <typename R>
static int build_plan(afid_ctx *c, int n, FftPlan<R> &pl) {
    pl = FftPlan<R>{};
    pl.n = n;
    int m = n;
    while (m % 8 == 0 && m > 1) { pl.rad[pl.nst++] = 8; m /= 8; }
    while (m % 4 == 0 && m > 1) { pl.rad[pl.nst++] = 4; m /= 4; }
    while (m % 2 == 0 && m > 1) { pl.rad[pl.nst++] = 2; m /= 2; }
    while (m % 3 == 0 && m > 1) { pl.rad[pl.nst++] = 3; m /= 3; }
    if (m != 1 || pl.nst > 12) return set_err(c, AFID_E_ARG, "periodic size " + std::to_string(n) + " must factor into 2 and 3");
    std::vector<cpx<R>> tw((size_t)(n));
    for (int t = 0; t < n; ++t) {
        tw[size_t(t)].x = R(std::cos(2.0 * kPi * t / n));
        tw[size_t(t)].y = R(-std::sin(2.0 * kPi * t / n));
    }
    const cpx<R> *d = nullptr;
    int rc = upload_vec(c, tw, &d);
    pl.tw = d;
    return rc;
}

// lines per transform workgroup (a power of two): L lines of N complex
// values within the LDS budget and the per-thread register budget
// (L*N <= FFT_MAXV*FFT_THREADS), at most cap
template <typename R>
static int pick_lines(int N, int maxlds, int cap, int maxv = FFT_MAXV) {
    int L = 1;
    while (2 * L * line_stride<R>(N) * int(sizeof(cpx<R>)) <= maxlds && 2 * L * N <= maxv * FFT_THREADS && 2 * L <= cap)
        L *= 2;
    return L;
}

static int setup(afid_ctx *c) {
    const afid_config &g = c->c;
    if (g.nxm < 2 || g.nym < 1 || g.nzm < 1) return set_err(c, AFID_E_ARG, "grid sizes must be nxm >= 2, nym >= 1, nzm >= 1");
    if (g.nsst != 1 && g.nsst != 3) return set_err(c, AFID_E_ARG, "nsst must be 1 (Adams-Bashforth) or 3 (RK3)");
    if (g.precision != AFID_FP64 && g.precision != AFID_MIXED) return set_err(c, AFID_E_ARG, "precision must be AFID_FP64 or AFID_MIXED");
    c->mixed = g.precision == AFID_MIXED;
    c->div_tol = g.div_tol > 0.0 ? g.div_tol : 1e-10;
    c->max_sweeps = g.max_sweeps > 0 ? g.max_sweeps : 3;
    if (c->max_sweeps > 4) return set_err(c, AFID_E_ARG, "max_sweeps must be 1..4");
    c->sal = g.salinity != 0;
    if (c->sal && !g.multires) return set_err(c, AFID_E_ARG, "salinity requires the refined grid (MULTIRES=1)");
    c->pf = g.phasefield != 0;
    if (c->pf && !g.multires) return set_err(c, AFID_E_ARG, "the phase field requires the refined grid (MULTIRES=1)");
    if (c->pf && c->sal) return set_err(c, AFID_E_ARG, "phase field together with salinity is not supported");
    if (c->pf && !(g.pf_a != 0.0 && g.pf_s != 0.0)) return set_err(c, AFID_E_ARG, "phase field needs pf_A and pf_S nonzero");
    c->nxm = g.nxm; c->nx = g.nxm + 1; c->ny = g.nym; c->nz = g.nzm;
    c->ldx = (c->nx + 7) / 8 * 8;
    c->ldt = (c->nx + 3 + 7) / 8 * 8;
    c->nymh = c->ny / 2 + 1;
    if (c->sal || c->pf) {
        c->nxmr = g.nxmr; c->nxr = g.nxmr + 1; c->nyr = g.nymr; c->nzr = g.nzmr;
        c->ldxr = (c->nxr + 7) / 8 * 8;
        if (c->nxmr < 2 || c->nyr < 2 || c->nzr < 2) return set_err(c, AFID_E_ARG, "refined grid too small");
    }
    // ReadInputFile.F90:101-122
    c->rrho = std::fabs(g.rayt) * g.pras / (std::fabs(g.rays) * g.prat);
    if (c->sal && g.ffscales == 1) { c->ren = std::sqrt(std::fabs(g.rays) / g.pras); c->byct = c->rrho; c->bycs = 1.0; }
    else { c->ren = std::sqrt(std::fabs(g.rayt) / g.prat); c->byct = 1.0; c->bycs = 1.0 / c->rrho; }
    c->pect = c->ren * g.prat;
    c->pecs = c->ren * g.pras;
    if (c->pf) {       // ReadInputFile.F90:125-126, phasefield.F90:357,531
        c->pf_d = 1.2 / c->pect / g.pf_s / g.pf_a;
        c->pf_eps = g.pf_eps > 0.0 ? g.pf_eps : 1.0 / g.nxmr;
        c->pf_b = c->pf_d / (c->pf_eps * c->pf_eps);
        c->heta = 0.5 / (c->ren * std::pow(1.51044285 * c->pf_eps, 2));
        c->pf_r0 = g.pf_r0 > 0.0 ? g.pf_r0 : 0.1;
    }
    // moisture.F90:44-56, :99-110 (the humid.in values; pecq = pect / Sm)
    c->moist = g.moist != 0;
    if (c->moist) {
        if (!(g.q_sm > 0.0) || !(g.q_tau > 0.0)) return set_err(c, AFID_E_ARG, "moisture needs q_sm > 0 and q_tau > 0");
        if (g.q_fixn < 0 || g.q_fixn > 2 || g.q_fixs < 0 || g.q_fixs > 1) return set_err(c, AFID_E_ARG, "qfixN must be 0, 1 or 2 and qfixS 0 or 1");
        c->q_alpha = g.q_alpha; c->q_beta = g.q_beta;
        c->pecq = c->pect / g.q_sm;
        c->kappa_q = 1.0 / c->pecq;
        c->q_tau = g.q_ifdiff ? c->pect * g.q_tau : g.q_tau;
        c->q_gamma = g.q_gamma;
        if (c->q_gamma < 0.0) c->q_gamma = g.q_fixn == 2 ? c->q_beta : c->q_beta / (1.0 - std::exp(-c->q_alpha));
        c->q_fixn = g.q_fixn == 2 ? 1 : g.q_fixn;      // SetHumidityBCs, :139-146
        c->q_fixs = g.q_fixs;
    }
    // ExplicitTermsVY.F90:30-34
    c->Gy = (g.inslwn == 0 || g.inslws == 0) ? g.dpdy * g.dpdy / (c->ren * c->ren) : 8.0 * g.dpdy * g.dpdy / (c->ren * c->ren);
    // InitTimeMarchScheme.F90:17-49
    if (g.nsst > 1) {
        c->gam[0] = 8.0 / 15.0; c->gam[1] = 5.0 / 12.0; c->gam[2] = 3.0 / 4.0;
        c->rom[0] = 0.0; c->rom[1] = -17.0 / 60.0; c->rom[2] = -5.0 / 12.0;
    } else {
        c->gam[0] = 1.5; c->gam[1] = 0.0; c->gam[2] = 0.0;
        c->rom[0] = -0.5; c->rom[1] = 0.0; c->rom[2] = 0.0;
    }
    for (int s = 0; s < 3; ++s) c->alm[s] = c->gam[s] + c->rom[s];

    const int nx = c->nx, nxm = c->nxm;
    // CreateGrid.F90:87-176
    FArr xc, xm;
    make_wall_grid(g.istr3, nxm, g.alx3, g.str3, g.istr3 == 5 ? g.dpdy : 1.0, xc, xm);
    c->dx = double(nxm) / g.alx3; c->dy = double(c->ny) / g.ylen; c->dz = double(c->nz) / g.zlen;
    xm(nx) = 2 * xc(nx) - xm(nxm);
    FArr g3rc(1, nx), g3rm(1, nx), d3xc(1, nx), d3xm(1, nx), udx3m(1, nx), udx3c(1, nx);
    for (int k = 2; k <= nxm; ++k) {
        g3rc(k) = (xc(k + 1) - xc(k - 1)) * c->dx * 0.5;
        g3rm(k) = (xm(k + 1) - xm(k - 1)) * c->dx * 0.5;
    }
    g3rm(1) = (xm(2) - (2 * xc(1) - xm(1))) * c->dx * 0.5;
    g3rc(1) = (xc(2) - xc(1)) * c->dx;
    g3rc(nx) = (xc(nx) - xc(nxm)) * c->dx;
    for (int k = 1; k <= nxm; ++k) { d3xc(k) = (xc(k + 1) - xc(k)) * c->dx; d3xm(k) = (xm(k + 1) - xm(k)) * c->dx; }
    for (int k = 1; k <= nxm; ++k) { udx3m(k) = c->dx / d3xc(k); udx3c(k) = c->dx / g3rc(k); }
    udx3c(nx) = c->dx / g3rc(nx);
    FArr ap3ck, ac3ck, am3ck, ap3sk, ac3sk, am3sk, ap3ssk, ac3ssk, am3ssk;
    second_derivative_coeff(ap3ck, ac3ck, am3ck, xc, nx, g.alx3, 1, 1);
    FArr xmm(1, nxm);
    for (int k = 1; k <= nxm; ++k) xmm(k) = xm(k);
    second_derivative_coeff(ap3sk, ac3sk, am3sk, xmm, nxm, g.alx3, g.inslwn, g.inslws);
    second_derivative_coeff(ap3ssk, ac3ssk, am3ssk, xmm, nxm, g.alx3, g.tfixn, g.tfixs);
    // InitPressureSolver, pressure.F90:78-114
    std::vector<double> ak1((size_t)(c->nz)), ak2((size_t)(c->ny));
    {
        const int nzm = c->nz, nym = c->ny;
        for (int i = 1; i <= nzm; ++i) {
            double ao = i <= nzm / 2 + 1 ? (i - 1) * 2.0 * kPi : -(nzm - i + 1) * 2.0 * kPi;
            ak1[size_t(i - 1)] = 2.0 * (1.0 - std::cos(ao / nzm)) * std::pow(double(nzm) / g.zlen, 2);
        }
        for (int j = 1; j <= nym; ++j) {
            double ap = j <= nym / 2 + 1 ? (j - 1) * 2.0 * kPi : -(nym - j + 1) * 2.0 * kPi;
            ak2[size_t(j - 1)] = 2.0 * (1.0 - std::cos(ap / nym)) * std::pow(double(nym) / g.ylen, 2);
        }
    }
    std::vector<double> amphk((size_t)(nxm)), acphk((size_t)(nxm)), apphk((size_t)(nxm));
    for (int k = 1; k <= nxm; ++k) {
        const int km = k == 1 ? 1 : k - 1, kp = k == nxm ? k : k + 1;
        const int kmc = k - km, kpc = kp - k;
        const double a33icc = kmc * c->dx * c->dx / d3xm(km), a33icp = kpc * c->dx * c->dx / d3xm(k);
        const double ug = 1.0 / d3xc(k);
        amphk[size_t(k - 1)] = a33icc * ug;
        apphk[size_t(k - 1)] = a33icp * ug;
        acphk[size_t(k - 1)] = -(amphk[size_t(k - 1)] + apphk[size_t(k - 1)]);
    }
    std::vector<double> rudx3m;
    for (int k = 1; k <= nxm; ++k) rudx3m.push_back(1.0 / udx3m(k));   // ExplicitTermsVX.F90:30-32

    c->h_ap3ck = farr_range(ap3ck, 1, nx); c->h_ac3ck = farr_range(ac3ck, 1, nx); c->h_am3ck = farr_range(am3ck, 1, nx);
    c->h_ap3sk = farr_range(ap3sk, 1, nxm); c->h_ac3sk = farr_range(ac3sk, 1, nxm); c->h_am3sk = farr_range(am3sk, 1, nxm);
    c->h_ap3ssk = farr_range(ap3ssk, 1, nxm); c->h_ac3ssk = farr_range(ac3ssk, 1, nxm); c->h_am3ssk = farr_range(am3ssk, 1, nxm);
    auto &H = c->host1d;
    H["xc"] = farr_range(xc, 1, nx); H["xm"] = farr_range(xm, 1, nx);
    H["g3rc"] = farr_range(g3rc, 1, nx); H["g3rm"] = farr_range(g3rm, 1, nx);
    H["d3xc"] = farr_range(d3xc, 1, nxm); H["d3xm"] = farr_range(d3xm, 1, nxm);
    H["udx3c"] = farr_range(udx3c, 1, nx); H["udx3m"] = farr_range(udx3m, 1, nxm);
    H["ap3ck"] = c->h_ap3ck; H["ac3ck"] = c->h_ac3ck; H["am3ck"] = c->h_am3ck;
    H["ap3sk"] = c->h_ap3sk; H["ac3sk"] = c->h_ac3sk; H["am3sk"] = c->h_am3sk;
    H["ap3ssk"] = c->h_ap3ssk; H["ac3ssk"] = c->h_ac3ssk; H["am3ssk"] = c->h_am3ssk;
    H["ak1"] = ak1; H["ak2"] = ak2; H["amphk"] = amphk; H["acphk"] = acphk; H["apphk"] = apphk;

    int rc;
    auto up = [&](const std::vector<double> &v, const double **d) { return rc ? rc : (rc = upload_vec(c, v, d)); };
    rc = 0;
    std::vector<double> udx3m_v = farr_range(udx3m, 1, nxm), udx3c_v = farr_range(udx3c, 1, nx), g3rc_v = farr_range(g3rc, 1, nx);
    std::vector<double> xm_v = farr_range(xm, 1, nxm);
    up(xm_v, &c->d_xm);
    std::vector<double> sdx_v;
    for (int k = 1; k <= nxm; ++k) sdx_v.push_back(c->dx / g3rm(k));     // CalcMeanProfiles.F90:116-123
    up(sdx_v, &c->d_sdx);
    up(udx3m_v, &c->d_udx3m); up(udx3c_v, &c->d_udx3c); up(rudx3m, &c->d_rudx3m); up(g3rc_v, &c->d_g3rc);
    up(c->h_ap3ck, &c->d_ap3ck); up(c->h_ac3ck, &c->d_ac3ck); up(c->h_am3ck, &c->d_am3ck);
    up(c->h_ap3sk, &c->d_ap3sk); up(c->h_ac3sk, &c->d_ac3sk); up(c->h_am3sk, &c->d_am3sk);
    up(c->h_ap3ssk, &c->d_ap3ssk); up(c->h_ac3ssk, &c->d_ac3ssk); up(c->h_am3ssk, &c->d_am3ssk);
    up(amphk, &c->d_amphk); up(acphk, &c->d_acphk); up(apphk, &c->d_apphk);
    if (rc) return rc;

    // fields.  FP64 form: every field in FP64.  Mixed form: the explicit-term
    // history and salc in FP32; qcap/dq/hro and the rhs arrays of the
    // per-routine calls are not held (the mixed form runs through the fused
    // step); dph holds the pressure correction of an FP32 sweep in FP32 (the
    // inverse transform computes it in FP32, so nothing is lost).
    const size_t nb = size_t(c->nz) * c->ny * c->ldx;
    for (int f = 0; f < AFID_NFIELDS; ++f) {
        if (is_refined(f) || f == AFID_RHS || is_plane(f) || f == AFID_PHIC) continue;
        if (is_moist_field(f) && !c->moist) continue;      // humidity is FP64 in both forms
        // (hro carries the latent heat of the phase field in both forms)
        if (c->mixed && (f == AFID_QCAP || f == AFID_DQ || (f == AFID_HRO && !c->pf))) continue;
        if (c->mixed && ((f >= AFID_RUX && f <= AFID_RUTEMP) || f == AFID_SALC || f == AFID_DPH)) {
            if (f == AFID_SALC && !c->sal) continue;
            if ((rc = dalloc(c, reinterpret_cast<void **>(&c->g[f]), sizeof(float) * nb))) return rc;
            continue;
        }
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->f[f]), sizeof(double) * nb))) return rc;
    }
    if (!c->mixed) {
        for (int q = 0; q < 4; ++q)
            if ((rc = dalloc(c, reinterpret_cast<void **>(&c->rhs4[q]), sizeof(double) * nb))) return rc;
        c->f[AFID_RHS] = c->rhs4[0];
    }
    for (int q = 0; q < 4; ++q)
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->alt[q]), sizeof(double) * nb))) return rc;
    for (int f : {AFID_TEMPBP, AFID_TEMPTP})
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->f[f]), sizeof(double) * size_t(c->ny) * c->nz))) return rc;
    if (c->moist) {
        for (int f : {AFID_HUMBP, AFID_HUMTP})
            if ((rc = dalloc(c, reinterpret_cast<void **>(&c->f[f]), sizeof(double) * size_t(c->ny) * c->nz))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->humid_alt), sizeof(double) * nb))) return rc;
    }
    const int csize = c->mixed ? int(sizeof(cpx<float>)) : int(sizeof(cpx<double>));
    if ((rc = dalloc(c, &c->spec, size_t(csize) * size_t(c->nz) * c->nymh * c->nxm))) return rc;
    if ((rc = dalloc(c, reinterpret_cast<void **>(&c->slots), sizeof(unsigned long long) * 4))) return rc;

    // Poisson transforms and factors
    if (c->mixed) {
        if ((rc = build_plan(c, c->ny, c->plan_yf))) return rc;
        if ((rc = build_plan(c, c->nz, c->plan_zf))) return rc;
    } else {
        if ((rc = build_plan(c, c->ny, c->plan_y))) return rc;
        if ((rc = build_plan(c, c->nz, c->plan_z))) return rc;
    }
    int maxlds = 65536;
    {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) == hipSuccess && v > 0)
            maxlds = std::min(v, 160 * 1024);
    }
    // y: Ly complex lines = 2*Ly x-adjacent real lines (16 FP64 velocity
    // values = 128 bytes per row); z: Lz x-adjacent lines (128-byte rows)
    if (c->mixed) {
        c->Ly = pick_lines<float>(c->ny, maxlds, 8);
        // 16 values per thread and stage (k_fft_z<float, 16>) when 8 would
        // leave rows shorter than 16 lines (128 bytes)
        c->Lz = pick_lines<float>(c->nz, maxlds, 16);
        if (c->Lz < 16 && c->fft_z16) {
            c->Lz = pick_lines<float>(c->nz, maxlds, 16, 16);
            c->fftz_maxv = 16;
        }
        c->lds_y = c->Ly * line_stride<float>(c->ny) * csize;
        c->lds_z = c->Lz * line_stride<float>(c->nz) * csize;
    } else {
        c->Ly = pick_lines<double>(c->ny, maxlds, 8);
        c->Lz = pick_lines<double>(c->nz, maxlds, 8);
        c->lds_y = c->Ly * line_stride<double>(c->ny) * csize;
        c->lds_z = c->Lz * line_stride<double>(c->nz) * csize;
    }
    if (c->lds_y > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_div_fft_y<double>), hipFuncAttributeMaxDynamicSharedMemorySize, c->lds_y);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_div_fft_y<float>), hipFuncAttributeMaxDynamicSharedMemorySize, c->lds_y);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_ifft_y<double>), hipFuncAttributeMaxDynamicSharedMemorySize, c->lds_y);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_ifft_y<float>), hipFuncAttributeMaxDynamicSharedMemorySize, c->lds_y);
    }
    if (c->lds_z > 65536) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_fft_z<double>), hipFuncAttributeMaxDynamicSharedMemorySize, c->lds_z);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_fft_z<float>), hipFuncAttributeMaxDynamicSharedMemorySize, c->lds_z);
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(k_fft_z<float, 16>), hipFuncAttributeMaxDynamicSharedMemorySize, c->lds_z);
    }
    if (c->lds_y > maxlds || c->lds_z > maxlds || c->ny > FFT_MAXV * FFT_THREADS || c->nz > FFT_MAXV * FFT_THREADS)
        return set_err(c, AFID_E_ARG, "periodic size too large for one workgroup transform");
    {
        c->pchunk = (nxm + 63) / 64;
        if (c->pchunk > 16) return set_err(c, AFID_E_ARG, "nxm above 1024 is not supported");
        const int npad = 64 * c->pchunk;
        const long long nmodes = (long long)c->nymh * (c->nz / 2 + 1);
        double *fm, *fg, *fe;
        const double *dak1, *dak2;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&fm), sizeof(double) * size_t(nmodes) * npad))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&fg), sizeof(double) * size_t(nmodes) * npad))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&fe), sizeof(double) * size_t(nmodes) * npad))) return rc;
        if ((rc = upload_vec(c, ak1, &dak1))) return rc;
        if ((rc = upload_vec(c, ak2, &dak2))) return rc;
        c->pois.m = fm; c->pois.g = fg; c->pois.e = fe; c->pois.npad = npad;
        KP p = make_kp(c);
        hipLaunchKernelGGL(k_poisson_factors, dim3(unsigned((nmodes + 127) / 128)), dim3(128), 0, c->stream, p, dak1, dak2, fm, fg, fe, npad);
        if ((rc = check_launch(c, "k_poisson_factors"))) return rc;
        if (c->mixed) {
            float *tf[3];
            const double *sf[3] = {fm, fg, fe};
            for (int q = 0; q < 3; ++q) {
                if ((rc = dalloc(c, reinterpret_cast<void **>(&tf[q]), sizeof(float) * size_t(nmodes) * npad))) return rc;
                hipLaunchKernelGGL(k_to_float, dim3(1024), dim3(256), 0, c->stream, sf[q], tf[q], (long long)nmodes * npad);
                if ((rc = check_launch(c, "k_to_float"))) return rc;
            }
            c->pff.m = tf[0]; c->pff.g = tf[1]; c->pff.e = tf[2]; c->pff.npad = npad;
        }
    }

    // tridiagonal factor storage
    c->nrows[MAT_X] = nx; c->nrows[MAT_YZ] = nxm; c->nrows[MAT_T] = nxm; c->nrows[MAT_S] = c->sal ? c->nxmr : 0;
    c->nrows[MAT_P] = c->pf ? c->nxmr : 0;
    c->nrows[MAT_Q] = c->moist ? nxm : 0;
    int chmax = 1;
    for (int k = 0; k < MAT_NKIND; ++k) { c->chunk[k] = std::max(1, (c->nrows[k] + 63) / 64); chmax = std::max(chmax, c->chunk[k]); }
    if (chmax > 16) return set_err(c, AFID_E_ARG, "wall-normal size above 1024 points is not supported");
    c->npadmax = 64 * chmax;
    if ((rc = dalloc(c, reinterpret_cast<void **>(&c->tri), sizeof(double) * size_t(3 * MAT_NKIND * kTriArrays) * c->npadmax))) return rc;

    // mixed form: the reference buoyancy of the FP32 vx explicit term (zero until the walls are set)
    if (c->mixed && g.gaxis == 1 && g.active_t != 0) {
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->d_bref), sizeof(double) * size_t(c->ldx)))) return rc;
        c->h_bref.assign(size_t(c->ldx), 0.0);
    }
    // reductions
    if ((rc = dalloc(c, reinterpret_cast<void **>(&c->redmax), sizeof(unsigned long long) * 4))) return rc;
    {
        dim3 gr = point_grid(nxm, c->ny, c->nz);
        c->nred = (long long)gr.x * gr.y * gr.z;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->redsum), sizeof(double) * 3 * size_t(c->nred)))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->redout), sizeof(double) * 4))) return rc;
    }

    if (!c->sal && !c->pf) return AFID_OK;

    // ---- refined grid: CreateMgrdGrid.F90:38-161 ----
    const int nxr = c->nxr, nxmr = c->nxmr, nyr = c->nyr, nzr = c->nzr;
    FArr xcr, xmr;
    const int istr = g.istr3 == 5 ? 5 : g.istr3r;   // CreateMgrdGrid.F90:77 tests istr3 for option 5
    if (istr == 2) return set_err(c, AFID_E_ARG, "refined grid option 2 does not exist in the reference");
    make_wall_grid(istr, nxmr, g.alx3, g.str3, istr == 5 ? g.dpdy : g.pras, xcr, xmr);
    c->dxr = double(nxmr) / g.alx3; c->dyr = double(nyr) / g.ylen; c->dzr = double(nzr) / g.zlen;
    xmr(0) = 2.0 * xcr(1) - xmr(1);
    xmr(nxr) = 2.0 * xcr(nxr) - xmr(nxmr);
    FArr udx3mr(1, nxmr), g3rmr(1, nxr);
    for (int k = 2; k <= nxmr; ++k) g3rmr(k) = (xmr(k + 1) - xmr(k - 1)) * c->dxr * 0.5;
    g3rmr(1) = (xmr(2) - (2 * xcr(1) - xmr(1))) * c->dxr * 0.5;
    for (int k = 1; k <= nxmr; ++k) udx3mr(k) = c->dxr / ((xcr(k + 1) - xcr(k)) * c->dxr);
    FArr xmrr(1, nxmr);
    for (int k = 1; k <= nxmr; ++k) xmrr(k) = xmr(k);
    FArr ap3sskr, ac3sskr, am3sskr;
    second_derivative_coeff(ap3sskr, ac3sskr, am3sskr, xmrr, nxmr, g.alx3, g.sfixn, g.sfixs);
    c->h_ap3sskr = farr_range(ap3sskr, 1, nxmr); c->h_ac3sskr = farr_range(ac3sskr, 1, nxmr); c->h_am3sskr = farr_range(am3sskr, 1, nxmr);
    if (c->pf) {       // zero-gradient operator of phi, CreateMgrdGrid.F90:163
        FArr ap3spkr, ac3spkr, am3spkr;
        second_derivative_coeff(ap3spkr, ac3spkr, am3spkr, xmrr, nxmr, g.alx3, 0, 0);
        c->h_ap3spkr = farr_range(ap3spkr, 1, nxmr); c->h_ac3spkr = farr_range(ac3spkr, 1, nxmr); c->h_am3spkr = farr_range(am3spkr, 1, nxmr);
        H["ap3spkr"] = c->h_ap3spkr; H["ac3spkr"] = c->h_ac3spkr; H["am3spkr"] = c->h_am3spkr;
    }
    std::vector<double> lxa;
    for (int k = 1; k <= nxmr; ++k) lxa.push_back(xcr(k + 1) - xcr(k));   // InterpVelMgrd.F90:62
    H["xcr"] = farr_range(xcr, 1, nxr); H["xmr"] = farr_range(xmr, 1, nxr);
    H["udx3mr"] = farr_range(udx3mr, 1, nxmr); H["g3rmr"] = farr_range(g3rmr, 1, nxr);
    H["ap3sskr"] = c->h_ap3sskr; H["ac3sskr"] = c->h_ac3sskr; H["am3sskr"] = c->h_am3sskr;

    // uniform periodic grids (uniform_grid + end extensions, CreateGrid.F90:32-40, CreateMgrdGrid.F90:38-46)
    auto uni = [](int n, double L, FArr &cc, FArr &mm) {
        cc = FArr(1, n + 1); mm = FArr(0, n + 1);
        for (int i = 1; i <= n + 1; ++i) cc(i) = L * double(i - 1) / double(n);
        for (int i = 1; i <= n; ++i) mm(i) = 0.5 * (cc(i) + cc(i + 1));
    };
    FArr yc, ym, zc, zm, ycr, ymr, zcr, zmr;
    uni(c->ny, g.ylen, yc, ym); uni(c->nz, g.zlen, zc, zm); uni(nyr, g.ylen, ycr, ymr); uni(nzr, g.zlen, zcr, zmr);
    FArr xmc(1, nxm);
    for (int k = 1; k <= nxm; ++k) xmc(k) = xm(k);

    // CreateMgrdStencil.F90:17-78: coarse cell centres -> refined cell centres
    IArr irangs = interpolation_indices(xmc, nxm, xmrr, nxmr, g.alx3);
    Stencil sx = construct_stencil(xmc, nxm, xmrr, nxmr, g.alx3, irangs, 'x');
    IArr jrangs = interpolation_indices(ym, c->ny, ymr, nyr, g.ylen);
    Stencil sy = construct_stencil(ym, c->ny, ymr, nyr, g.ylen, jrangs, 'y');
    IArr krangs = interpolation_indices(zm, c->nz, zmr, nzr, g.zlen);
    Stencil sz = construct_stencil(zm, c->nz, zmr, nzr, g.zlen, krangs, 'z');
    // refined -> coarse tables (CreateMgrdStencil.F90:21-78).  Each is, per
    // coarse index, the first of 4 refined sources (1-based) and 4 weights.
    // The bracketing (fast) path (HermiteInterpolations.F90:256-312) takes
    // sources irang-2..irang+1 with the coarse -> refined index arrays; the
    // general path (:218-250, used when xmr(1) >= xm(1) or xmr(nxmr) <=
    // xm(nxm), salinity.F90:485-489, phasefield.F90:489-492, :701-723) takes
    // kcr-1..kcr+2 of the refined cell kcr owning the coarse point, with the
    // cell-centre (phic) weights on every axis.
    const bool bracket = xmr(1) < xm(1) && xmr(nxmr) > xm(nxm);
    IArr irangr = interpolation_indices(xmrr, nxmr, xmc, nxm, g.alx3);
    IArr jrangr = interpolation_indices(ymr, nyr, ym, c->ny, g.ylen);
    IArr krangr = interpolation_indices(zmr, nzr, zm, c->nz, g.zlen);
    Stencil bp = construct_stencil(xmrr, nxmr, xmc, nxm, g.alx3, irangr, 'x');   // cxphic
    Stencil byc = construct_stencil(ymr, nyr, ym, c->ny, g.ylen, jrangr, 'y');    // cyphic
    Stencil bzc = construct_stencil(zmr, nzr, zm, c->nz, g.zlen, krangr, 'z');    // czphic
    H["cxphic"] = bp.w;
    struct HostTab { std::vector<int> xs, ys, zs; std::vector<double> xw, yw, zw; };
    HostTab hs, hp;     // salinity, phase field
    if (bracket) {
        // salinity: salc at the faces along gAxis (cxsalc / cysalc / czsalc, irangb / jrangb / krangb)
        FArr xcc(1, nx), ycc(1, c->ny), zcc(1, c->nz);
        for (int k = 1; k <= nx; ++k) xcc(k) = xc(k);
        for (int j = 1; j <= c->ny; ++j) ycc(j) = yc(j);
        for (int i = 1; i <= c->nz; ++i) zcc(i) = zc(i);
        hs.xw = bp.w; hs.yw = byc.w; hs.zw = bzc.w;
        IArr irangb = irangs, jrangb = jrangs, krangb = krangs;
        if (g.gaxis == 1) {
            IArr irange = interpolation_indices(xmrr, nxmr, xcc, nx, g.alx3);
            hs.xw = construct_stencil(xmrr, nxmr, xcc, nx, g.alx3, irange, 'x').w;
            irangb = interpolation_indices(xcc, nx, xmrr, nxmr, g.alx3);
        } else if (g.gaxis == 2) {
            IArr jrange = interpolation_indices(ymr, nyr, ycc, c->ny, g.ylen);
            hs.yw = construct_stencil(ymr, nyr, ycc, c->ny, g.ylen, jrange, 'y').w;
            jrangb = interpolation_indices(ycc, c->ny, ymr, nyr, g.ylen);
        } else if (g.gaxis == 3) {
            IArr krange = interpolation_indices(zmr, nzr, zcc, c->nz, g.zlen);
            hs.zw = construct_stencil(zmr, nzr, zcc, c->nz, g.zlen, krange, 'z').w;
            krangb = interpolation_indices(zcc, c->nz, zmr, nzr, g.zlen);
        }
        for (int k = 1; k <= nxm; ++k) hs.xs.push_back(irangb(k) - 2);
        for (int j = 1; j <= c->ny; ++j) hs.ys.push_back(jrangb(j) - 2);
        for (int i = 1; i <= c->nz; ++i) hs.zs.push_back(krangb(i) - 2);
        // phase field: cxphic with irangs ("phi" branch, :283-290)
        hp.xw = bp.w; hp.yw = byc.w; hp.zw = bzc.w;
        for (int k = 1; k <= nxm; ++k) hp.xs.push_back(irangs(k) - 2);
        for (int j = 1; j <= c->ny; ++j) hp.ys.push_back(jrangs(j) - 2);
        for (int i = 1; i <= c->nz; ++i) hp.zs.push_back(krangs(i) - 2);
    } else {
        // construct_stencil's src: kcr - 1 of the refined cell kcr owning the coarse point
        hp.xs = bp.src; hp.ys = byc.src; hp.zs = bzc.src;
        hp.xw = bp.w; hp.yw = byc.w; hp.zw = bzc.w;
        hs = hp;
    }
    H["cxsalc"] = hs.xw; H["cysalc"] = hs.yw; H["czsalc"] = hs.zw;
    H["cxrs"] = sx.w; H["cyrs"] = sy.w; H["czrs"] = sz.w;

    const int *pi;
    const double *pd;
    if ((rc = upload_vec(c, sx.src, &pi))) return rc; c->d_xs_src = pi;
    if ((rc = upload_vec(c, sy.src, &pi))) return rc; c->d_ys_src = pi;
    if ((rc = upload_vec(c, sz.src, &pi))) return rc; c->d_zs_src = pi;
    if ((rc = upload_vec(c, sx.w, &pd))) return rc; c->d_xs_w = pd;
    if ((rc = upload_vec(c, sy.w, &pd))) return rc; c->d_ys_w = pd;
    if ((rc = upload_vec(c, sz.w, &pd))) return rc; c->d_zs_w = pd;
    // the refined -> coarse tables on the device (FP32 weights too in the mixed form)
    auto up_tab = [&](const HostTab &h, CoarseTab &t) {
        const int *q;
        const double *w;
        if (rc || (rc = upload_vec(c, h.xs, &q))) return; t.x_src = q;
        if ((rc = upload_vec(c, h.ys, &q))) return; t.y_src = q;
        if ((rc = upload_vec(c, h.zs, &q))) return; t.z_src = q;
        if ((rc = upload_vec(c, h.xw, &w))) return; t.x_w = w;
        if ((rc = upload_vec(c, h.yw, &w))) return; t.y_w = w;
        if ((rc = upload_vec(c, h.zw, &w))) return; t.z_w = w;
        if (c->mixed) {
            const float *f;
            std::vector<float> a(h.xw.begin(), h.xw.end()), b(h.yw.begin(), h.yw.end()), d(h.zw.begin(), h.zw.end());
            if ((rc = upload_vec(c, a, &f))) return; t.x_wf = f;
            if ((rc = upload_vec(c, b, &f))) return; t.y_wf = f;
            if ((rc = upload_vec(c, d, &f))) return; t.z_wf = f;
        }
        t.general = !bracket;
    };
    if (c->sal) up_tab(hs, c->tab_sal);
    if (c->pf) up_tab(hp, c->tab_phi);
    if (rc) return rc;
    std::vector<double> udx3mr_v = farr_range(udx3mr, 1, nxmr);
    std::vector<double> xmr_v = farr_range(xmr, 1, nxmr);
    up(xmr_v, &c->d_xmr);
    std::vector<double> sdxr_v;
    for (int k = 1; k <= nxmr; ++k) sdxr_v.push_back(c->dxr / g3rmr(k));     // CalcDissipationSal, salinity.F90:620-622
    up(sdxr_v, &c->d_sdxr);
    up(udx3mr_v, &c->d_udx3mr); up(c->h_ap3sskr, &c->d_ap3sskr); up(c->h_ac3sskr, &c->d_ac3sskr); up(c->h_am3sskr, &c->d_am3sskr);
    up(lxa, &c->d_lxa);
    if (c->pf) {
        up(c->h_ap3spkr, &c->d_ap3spkr); up(c->h_ac3spkr, &c->d_ac3spkr); up(c->h_am3spkr, &c->d_am3spkr);
    }
    if (rc) return rc;

    const size_t nr = size_t(nzr) * nyr * c->ldxr;
    if (c->mixed) {
        // FP32 refined fields; hsal/hphi and the salinity rhsr belong to the
        // per-routine calls only (the phase field keeps d phi in rhsr)
        if (c->sal) {
            for (int f : {AFID_SAL, AFID_RUSAL, AFID_VXR, AFID_VYR, AFID_VZR})
                if ((rc = dalloc(c, reinterpret_cast<void **>(&c->g[f]), sizeof(float) * nr))) return rc;
            if ((rc = dalloc(c, reinterpret_cast<void **>(&c->sal_alt_f), sizeof(float) * nr))) return rc;
        } else {
            for (int f : {AFID_PHI, AFID_RUPHI, AFID_TEMPR, AFID_RHSR})
                if ((rc = dalloc(c, reinterpret_cast<void **>(&c->g[f]), sizeof(float) * nr))) return rc;
            if ((rc = dalloc(c, reinterpret_cast<void **>(&c->phi_alt_f), sizeof(float) * nr))) return rc;
            if ((rc = dalloc(c, reinterpret_cast<void **>(&c->f[AFID_PHIC]), sizeof(double) * size_t(c->nz) * c->ny * c->ldx))) return rc;
        }
        auto upf = [&](const std::vector<double> &v, const float **d) {
            std::vector<float> w(v.begin(), v.end());
            return rc ? rc : (rc = upload_vec(c, w, d));
        };
        upf(sx.w, &c->f_xs_w); upf(sy.w, &c->f_ys_w); upf(sz.w, &c->f_zs_w);
        upf(lxa, &c->f_lxa);
        if (rc) return rc;
    } else if (c->sal) {
        for (int f = AFID_SAL; f <= AFID_VZR; ++f)
            if ((rc = dalloc(c, reinterpret_cast<void **>(&c->f[f]), sizeof(double) * nr))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->sal_alt), sizeof(double) * nr))) return rc;
    } else {
        // phase field: phi (+ second buffer), its RK arrays, tempr, rhsr; coarse phic
        for (int f : {AFID_PHI, AFID_RUPHI, AFID_HPHI, AFID_TEMPR, AFID_RHSR})
            if ((rc = dalloc(c, reinterpret_cast<void **>(&c->f[f]), sizeof(double) * nr))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->phi_alt), sizeof(double) * nr))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->f[AFID_PHIC]), sizeof(double) * size_t(c->nz) * c->ny * c->ldx))) return rc;
    }
    if (c->sal)
        for (int f : {AFID_SALBP, AFID_SALTP})
            if ((rc = dalloc(c, reinterpret_cast<void **>(&c->f[f]), sizeof(double) * size_t(nyr) * nzr))) return rc;
    if (nxmr > 1024) return set_err(c, AFID_E_ARG, "refined wall-normal size above 1024 is not supported");
    // LDS plan of the tiled transfer kernels: largest coarse block of any tile
    auto span = [](const std::vector<int> &src, int n, int tile) {
        int m = 0;
        for (int a = 0; a < n; a += tile) {
            const int b = std::min(a + tile, n) - 1;
            m = std::max(m, src[size_t(b)] - src[size_t(a)] + 4);
        }
        return m;
    };
    c->vxr_xc = span(sx.src, nxmr, 64);
    if (c->vxr_xc > 128) return set_err(c, AFID_E_ARG, "refined x coarser than the base x is not supported");
    c->vxr_cj = span(sy.src, nyr, VXR_TJ);
    c->vxr_ci = span(sz.src, nzr, VXR_TI);
    if (c->vxr_cj * c->vxr_ci > 256) return set_err(c, AFID_E_ARG, "refinement ratio too small for the vxr tiles");
    const size_t es = c->mixed ? sizeof(float) : sizeof(double);      // element size of the transfer LDS
    c->lds_vxr = es * size_t(c->vxr_xc * c->vxr_cj * c->vxr_ci + c->vxr_xc * c->vxr_cj * VXR_TI +
                             c->vxr_xc * VXR_TJ * VXR_TI);
    c->vy_cs = span(sy.src, nyr - 1, VYZ_CHUNK); c->vy_co = span(sz.src, nzr, VYZ_LINES);
    c->vz_cs = span(sz.src, nzr - 1, VYZ_CHUNK); c->vz_co = span(sy.src, nyr, VYZ_LINES);
    if (std::max(std::max(c->vy_cs, c->vz_cs), std::max(c->vy_co, c->vz_co)) > 32)
        return set_err(c, AFID_E_ARG, "refinement ratio too small for the vyr/vzr tiles");
    auto lds_yz = [&](int cs, int co) {
        return es * size_t(std::max(c->vxr_xc * cs * co, cs * VYZ_LINES * 64) + c->vxr_xc * cs * VYZ_LINES);
    };
    c->lds_vy = lds_yz(c->vy_cs, c->vy_co);
    c->lds_vz = lds_yz(c->vz_cs, c->vz_co);
    // refined -> coarse tiles: 64 coarse x by SALC_TJ coarse y, two refined
    // planes of the tile in LDS (storage type of the refined field)
    for (int q = 0; q < 2; ++q) {
        const HostTab &h = q == 0 ? hs : hp;
        CoarseTab &t = q == 0 ? c->tab_sal : c->tab_phi;
        if ((q == 0 && !c->sal) || (q == 1 && !c->pf)) continue;
        t.xr = span(h.xs, nxm, 64);
        t.yr = span(h.ys, c->ny, SALC_TJ);
        if (t.yr > 64) return set_err(c, AFID_E_ARG, "refinement ratio too large for the refined -> coarse tiles");
        t.lds = es * size_t(2 * t.yr * t.xr);
        if (t.lds > 65536) return set_err(c, AFID_E_ARG, "refinement ratio too small for the refined -> coarse tiles");
    }
    if (c->lds_vxr > 65536 || c->lds_vy > 65536 || c->lds_vz > 65536)
        return set_err(c, AFID_E_ARG, "refinement ratio too small for the transfer tiles (LDS plan above 64 KiB)");
    return AFID_OK;
}

// ---------------------------------------------------------------------------
// ABI
// ---------------------------------------------------------------------------
extern "C" {

}  // extern "C"

static int scratch_ptr_for_device(afid_ctx *c, int dev);

// A context on the caller's stream (shared) or on a stream of its own.
static int ctx_create(const afid_config *cfg, hipStream_t shared, afid_ctx **out) {
    if (!cfg || !out) return AFID_E_ARG;
    *out = nullptr;
    afid_ctx *c = new afid_ctx();
    c->c = *cfg;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        fprintf(stderr, "afid_gpu_init: no HIP device\n");
        delete c;
        return AFID_E_HIP;
    }
    c->device = dev;
    if (const char *e = std::getenv("AFID_COLUMN_FOUR_PASS")) c->col_four_pass = std::atoi(e) != 0;
    if (const char *e = std::getenv("AFID_VYZ_TILE")) c->vyz_tile = std::atoi(e) != 0;
    if (const char *e = std::getenv("AFID_FFT_Z16")) c->fft_z16 = std::atoi(e) != 0;
    if (shared) {
        c->stream = shared;
        c->own_stream = false;
    } else if (hipStreamCreate(&c->stream) != hipSuccess) {
        delete c;
        return AFID_E_HIP;
    }
    int rc;
    try {
        rc = setup(c);
    } catch (const std::exception &e) {
        rc = set_err(c, AFID_E_ARG, e.what());
    }
    // the static scratch of afid_calc_mean_profiles_noalloc: its address on
    // this device, once per device
    if (rc == AFID_OK) rc = scratch_ptr_for_device(c, dev);
    if (rc == AFID_OK) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = set_err(c, AFID_E_HIP, std::string("set-up: ") + hipGetErrorString(e));
    }
    *out = c;   // returned even on failure so afid_last_error can report; caller finalizes
    return rc;
}

extern "C" {

int afid_gpu_init(const afid_config *cfg, afid_ctx **out) { return ctx_create(cfg, nullptr, out); }

int afid_gpu_finalize(afid_ctx *c) {
    if (!c) return AFID_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    for (auto &t : c->timers) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (auto &b : c->bufs) (void)hipFree(b.p);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return AFID_OK;
}

const char *afid_last_error(const afid_ctx *c) { return c ? c->err.c_str() : "null context"; }
int64_t afid_device_bytes(const afid_ctx *c) { return c ? c->bytes : 0; }

static int field_dims(const afid_ctx *c, int f, int &nxf, int &ny, int &nz, int &ld) {
    if (f < 0 || f >= AFID_NFIELDS || (!c->f[f] && !c->g[f])) return AFID_E_ARG;
    const bool r = is_refined(f);
    ny = r ? c->nyr : c->ny;
    nz = r ? c->nzr : c->nz;
    if (is_plane(f)) { nxf = 1; ld = 1; }
    else { nxf = r ? c->nxr : c->nx; ld = r ? c->ldxr : c->ldx; }
    return AFID_OK;
}

int afid_upload(afid_ctx *c, int f, const double *host, const int64_t lb[3], const int64_t ext[3]) {
    int nxf, ny, nz, ld;
    if (!c || !host || field_dims(c, f, nxf, ny, nz, ld)) return set_err(c, AFID_E_ARG, "afid_upload: bad field");
    if (lb[1] > 1 || lb[2] > 1 || lb[1] + ext[1] - 1 < ny || lb[2] + ext[2] - 1 < nz || lb[0] > 1)
        return set_err(c, AFID_E_ARG, "afid_upload: host array does not cover the interior");
    const int kmax = int(std::min<int64_t>(nxf, lb[0] + ext[0] - 1));
    std::vector<double> slab(size_t(ny) * ld, 0.0);
    std::vector<float> slabf(c->g[f] ? slab.size() : 0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 1; i <= nz; ++i) {
        for (int j = 1; j <= ny; ++j)
            for (int k = 1; k <= kmax; ++k)
                slab[size_t(j - 1) * ld + (k - 1)] = host[(k - lb[0]) + ext[0] * ((j - lb[1]) + ext[1] * (int64_t)(i - lb[2]))];
        if (f == AFID_RUX && c->g[f] && c->d_bref)        // the FP32 history holds H - bref
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < kmax; ++k) slab[size_t(j) * ld + k] -= c->h_bref[size_t(k)];
        if (c->g[f]) {       // FP32 field of the mixed form: rounded to nearest
            for (size_t q = 0; q < slab.size(); ++q) slabf[q] = float(slab[q]);
            HIPCHK(c, hipMemcpy(c->g[f] + size_t(i - 1) * ny * ld, slabf.data(), sizeof(float) * slabf.size(), hipMemcpyHostToDevice));
        } else {
            HIPCHK(c, hipMemcpy(c->f[f] + size_t(i - 1) * ny * ld, slab.data(), sizeof(double) * slab.size(), hipMemcpyHostToDevice));
        }
    }
    if (f == AFID_TEMPBP || f == AFID_TEMPTP) {
        double sum = 0.0;
        for (int i = 1; i <= nz; ++i)
            for (int j = 1; j <= ny; ++j) sum += host[(1 - lb[0]) + ext[0] * ((j - lb[1]) + ext[1] * (int64_t)(i - lb[2]))];
        c->wall_mean[f == AFID_TEMPBP ? 0 : 1] = sum / (double(ny) * nz);
        c->bref_dirty = true;
    }
    return AFID_OK;
}

int afid_download(afid_ctx *c, int f, double *host, const int64_t lb[3], const int64_t ext[3]) {
    int nxf, ny, nz, ld;
    if (!c || !host || field_dims(c, f, nxf, ny, nz, ld)) return set_err(c, AFID_E_ARG, "afid_download: bad field");
    if (lb[1] > 1 || lb[2] > 1 || lb[1] + ext[1] - 1 < ny || lb[2] + ext[2] - 1 < nz || lb[0] > 1)
        return set_err(c, AFID_E_ARG, "afid_download: host array does not cover the interior");
    const int kmax = int(std::min<int64_t>(nxf, lb[0] + ext[0] - 1));
    std::vector<double> slab(size_t(ny) * ld);
    std::vector<float> slabf(c->g[f] ? slab.size() : 0);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t ih = lb[2]; ih < lb[2] + ext[2]; ++ih) {
        const int i = int(((ih - 1) % nz + nz) % nz);    // periodic halo (halo_common.F90:66-71 at one rank)
        if (c->g[f]) {
            HIPCHK(c, hipMemcpy(slabf.data(), c->g[f] + size_t(i) * ny * ld, sizeof(float) * slabf.size(), hipMemcpyDeviceToHost));
            for (size_t q = 0; q < slab.size(); ++q) slab[q] = double(slabf[q]);
            if (f == AFID_RUX && c->d_bref)
                for (int j = 0; j < ny; ++j)
                    for (int k = 0; k < kmax; ++k) slab[size_t(j) * ld + k] += c->h_bref[size_t(k)];
        } else {
            HIPCHK(c, hipMemcpy(slab.data(), c->f[f] + size_t(i) * ny * ld, sizeof(double) * slab.size(), hipMemcpyDeviceToHost));
        }
        for (int64_t jh = lb[1]; jh < lb[1] + ext[1]; ++jh) {
            const int j = int(((jh - 1) % ny + ny) % ny);
            for (int k = 1; k <= kmax; ++k)
                host[(k - lb[0]) + ext[0] * ((jh - lb[1]) + ext[1] * (ih - lb[2]))] = slab[size_t(j) * ld + (k - 1)];
        }
    }
    return AFID_OK;
}

int afid_set_substep(afid_ctx *c, double dt, int ns) {
    if (!c || ns < 1 || ns > c->c.nsst || !(dt > 0.0)) return set_err(c, AFID_E_ARG, "afid_set_substep: bad dt or substep");
    c->dt = dt;
    c->beta = dt / c->ren * 0.5;
    c->al = c->alm[ns - 1]; c->ga = c->gam[ns - 1]; c->ro = c->rom[ns - 1];
    c->ns = ns;
    if (c->d_bref && c->bref_dirty && ns == 1) {
        // the reference buoyancy of the conductive profile T = Tb + (Tt - Tb) x / alx3
        // at the vx faces: byct active_T (T(xm_k) + T(xm_{k-1})) / 2, k = 1..nxm-1
        // (0-based); taken at the first substep, whose history weight is zero
        const std::vector<double> &xm = c->host1d["xm"];
        const double tb = c->wall_mean[0], tt = c->wall_mean[1];
        auto tl = [&](double x) { return tb + (tt - tb) * x / c->c.alx3; };
        c->h_bref.assign(size_t(c->ldx), 0.0);
        for (int k = 1; k < c->nxm; ++k)
            c->h_bref[size_t(k)] = c->byct * (c->c.active_t * 0.5 * (tl(xm[size_t(k)]) + tl(xm[size_t(k - 1)])));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipMemcpy(c->d_bref, c->h_bref.data(), sizeof(double) * c->h_bref.size(), hipMemcpyHostToDevice));
        c->bref_dirty = false;
    }
    return ensure_factors(c);
}

#define NEED_SUBSTEP(c) if (!(c) || (c)->ns == 0) return set_err((c), AFID_E_STATE, "no substep set (afid_set_substep)");
#define NEED_SAL(c) if (!(c)->sal) return set_err((c), AFID_E_STATE, "salinity is not enabled");
#define NEED_FP64(c) if ((c)->mixed) return set_err((c), AFID_E_STATE, "the per-routine calls belong to the FP64 form; the mixed form runs through afid_time_marcher");
#define NEED_MOIST(c) if (!(c) || !(c)->moist) return set_err((c), AFID_E_STATE, "moisture is not enabled (moist = 0)");

static int launch_point(afid_ctx *c, void (*kern)(KP), bool refined, const char *name) {
    KP p = make_kp(c);
    dim3 gr = refined ? point_grid(c->nxmr, c->nyr, c->nzr) : point_grid(c->nxm, c->ny, c->nz);
    hipLaunchKernelGGL(kern, gr, kPointBlock, 0, c->stream, p);
    return check_launch(c, name);
}

int afid_explicit_terms_vx(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) return launch_point(c, k_explicit_vx, false, "k_explicit_vx"); }
int afid_explicit_terms_vy(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) return launch_point(c, k_explicit_vy, false, "k_explicit_vy"); }
int afid_explicit_terms_vz(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) return launch_point(c, k_explicit_vz, false, "k_explicit_vz"); }
int afid_explicit_terms_temp(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) return launch_point(c, k_explicit_temp, false, "k_explicit_temp"); }
int afid_explicit_salinity(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) NEED_SAL(c) return launch_point(c, k_explicit_sal, true, "k_explicit_sal"); }

int afid_add_sal_buoyancy(afid_ctx *c) {
    NEED_SUBSTEP(c) NEED_FP64(c) NEED_SAL(c)
    double *rkv = c->c.gaxis == 1 ? c->f[AFID_QCAP] : c->c.gaxis == 2 ? c->f[AFID_DPH] : c->f[AFID_DQ];
    KP p = make_kp(c);
    hipLaunchKernelGGL(k_add_sal_buoyancy, point_grid(c->nxm, c->ny, c->nz), kPointBlock, 0, c->stream, p, rkv);
    return check_launch(c, "k_add_sal_buoyancy");
}

}  // extern "C"

// immersed-boundary solve (SolveImpEqnUpdate_*_ibm) of slot 0..4
template <int CH>
static void launch_tri_ibm(afid_ctx *c, const TriFactors &f, const IbmJob &jb, const double *rhs, double *q) {
    const unsigned nb = unsigned(((long long)jb.ny * jb.nz + 3) / 4);
    hipLaunchKernelGGL((k_tri_update_ibm<CH>), dim3(nb), dim3(256), 0, c->stream, f, jb, rhs, q);
}
static int run_tri_ibm(afid_ctx *c, int slot, int kind, const double *rhs, double *q) {
    TriFactors f = factors(c, kind);
    const bool ref = slot == 4;
    IbmJob jb{c->ibm_mask[slot], c->ibm_dist[slot], 1.0, slot, ref ? c->nxmr : c->nxm, ref ? c->ldxr : c->ldx,
              ref ? c->nyr : c->ny, ref ? c->nzr : c->nz, ref ? (long long)c->nyr * c->ldxr : (long long)c->ny * c->ldx};
    switch (f.chunk) {
#define AFID_TI_CASE(CH) case CH: launch_tri_ibm<CH>(c, f, jb, rhs, q); break;
        AFID_TI_CASE(1) AFID_TI_CASE(2) AFID_TI_CASE(3) AFID_TI_CASE(4) AFID_TI_CASE(5) AFID_TI_CASE(6)
        AFID_TI_CASE(7) AFID_TI_CASE(8) AFID_TI_CASE(9) AFID_TI_CASE(10) AFID_TI_CASE(11) AFID_TI_CASE(12)
        AFID_TI_CASE(13) AFID_TI_CASE(14) AFID_TI_CASE(15) AFID_TI_CASE(16)
#undef AFID_TI_CASE
    default: return set_err(c, AFID_E_ARG, "wall-normal size above 1024 points is not supported");
    }
    return check_launch(c, "k_tri_update_ibm");
}

extern "C" {

// ibm_param.F90: the mask and distances of one solve (see afid_gpu.h)
int afid_set_ibm(afid_ctx *c, int field, const int32_t *mask, const double *dist, int64_t ndist) {
    if (!c) return AFID_E_ARG;
    const int slot = field == AFID_VX ? 0 : field == AFID_VY ? 1 : field == AFID_VZ ? 2 : field == AFID_TEMP ? 3 : field == AFID_SAL ? 4 : -1;
    if (slot < 0) return set_err(c, AFID_E_ARG, "afid_set_ibm: field must be AFID_VX, AFID_VY, AFID_VZ, AFID_TEMP or AFID_SAL");
    if (slot == 4 && !c->sal) return set_err(c, AFID_E_STATE, "salinity is not enabled");
    if (c->mixed) return set_err(c, AFID_E_STATE, "the immersed-boundary solves belong to the FP64 call tree");
    int rc;
    if (!mask) {                       // switch this solve back to the plain form
        c->ibm_mask[slot] = nullptr;
        c->ibm_dist[slot] = nullptr;
        return AFID_OK;
    }
    const bool ref = slot == 4;
    const int nx = ref ? c->nxr : c->nx, ny = ref ? c->nyr : c->ny, nz = ref ? c->nzr : c->nz, ld = ref ? c->ldxr : c->ldx;
    const int nrow = ref ? c->nxmr : c->nxm;
    std::vector<signed char> m(size_t(nz) * ny * ld, 2);
    std::vector<double> d(m.size(), 0.0);
    int64_t q = 0;
    // the reference consumes the distances in its loop order: ic, jc, kc = 1..nxm
    for (int i = 0; i < nz; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nx; ++k) {
                const int32_t v = mask[(size_t(i) * ny + j) * nx + k];
                if (v != 2 && v != 0 && v != 1 && v != -1) return set_err(c, AFID_E_ARG, "afid_set_ibm: mask values are 2, 0, 1 or -1");
                const size_t o = (size_t(i) * ny + j) * ld + k;
                m[o] = (signed char)v;
                if (k < nrow && (v == 1 || v == -1)) {
                    if (!ref) {
                        if (!dist || q >= ndist) return set_err(c, AFID_E_ARG, "afid_set_ibm: fewer distances than boundary points");
                        d[o] = dist[q];
                    }
                    ++q;
                }
            }
    if (!c->ibm_mask[slot]) {
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->ibm_mask[slot]), m.size()))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->ibm_dist[slot]), sizeof(double) * d.size()))) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->ibm_mask[slot], m.data(), m.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->ibm_dist[slot], d.data(), sizeof(double) * d.size(), hipMemcpyHostToDevice));
    return AFID_OK;
}

}  // extern "C"

// masked velocity solve of the phase field (SolveImpEqnUpdate_X_pf / _YZ_pf)
template <int CH, int AX>
static void launch_tri_pf(afid_ctx *c, const TriFactors &f, const double *rhs, double *q) {
    KP p = make_kp(c);
    const unsigned nb = unsigned(((long long)c->ny * c->nz + 3) / 4);
    hipLaunchKernelGGL((k_tri_update_pf<CH, AX>), dim3(nb), dim3(256), 0, c->stream, p, f, rhs, q);
}
template <int AX>
static int run_tri_pf(afid_ctx *c, int kind, const double *rhs, double *q) {
    TriFactors f = factors(c, kind);
    switch (f.chunk) {
#define AFID_TP_CASE(CH) case CH: launch_tri_pf<CH, AX>(c, f, rhs, q); break;
        AFID_TP_CASE(1) AFID_TP_CASE(2) AFID_TP_CASE(3) AFID_TP_CASE(4) AFID_TP_CASE(5) AFID_TP_CASE(6)
        AFID_TP_CASE(7) AFID_TP_CASE(8) AFID_TP_CASE(9) AFID_TP_CASE(10) AFID_TP_CASE(11) AFID_TP_CASE(12)
        AFID_TP_CASE(13) AFID_TP_CASE(14) AFID_TP_CASE(15) AFID_TP_CASE(16)
#undef AFID_TP_CASE
    default: return set_err(c, AFID_E_ARG, "wall-normal size above 1024 points is not supported");
    }
    return check_launch(c, "k_tri_update_pf");
}

extern "C" {

int afid_implicit_and_update_vx(afid_ctx *c) {
    NEED_SUBSTEP(c) NEED_FP64(c)
    int rc = launch_point(c, k_rhs_vx, false, "k_rhs_vx");
    if (!rc) rc = c->ibm_mask[0] ? run_tri_ibm(c, 0, MAT_X, c->rhs4[0], c->f[AFID_VX])
                  : c->pf ? run_tri_pf<0>(c, MAT_X, c->rhs4[0], c->f[AFID_VX]) : run_tri(c, MAT_X, c->rhs4[0], c->f[AFID_VX]);
    c->f[AFID_RHS] = c->rhs4[0];
    return rc;
}
int afid_implicit_and_update_vy(afid_ctx *c) {
    NEED_SUBSTEP(c) NEED_FP64(c)
    int rc = launch_point(c, k_rhs_vy, false, "k_rhs_vy");
    if (!rc) rc = c->ibm_mask[1] ? run_tri_ibm(c, 1, MAT_YZ, c->rhs4[1], c->f[AFID_VY])
                  : c->pf ? run_tri_pf<1>(c, MAT_YZ, c->rhs4[1], c->f[AFID_VY]) : run_tri(c, MAT_YZ, c->rhs4[1], c->f[AFID_VY]);
    c->f[AFID_RHS] = c->rhs4[1];
    return rc;
}
int afid_implicit_and_update_vz(afid_ctx *c) {
    NEED_SUBSTEP(c) NEED_FP64(c)
    int rc = launch_point(c, k_rhs_vz, false, "k_rhs_vz");
    if (!rc) rc = c->ibm_mask[2] ? run_tri_ibm(c, 2, MAT_YZ, c->rhs4[2], c->f[AFID_VZ])
                  : c->pf ? run_tri_pf<2>(c, MAT_YZ, c->rhs4[2], c->f[AFID_VZ]) : run_tri(c, MAT_YZ, c->rhs4[2], c->f[AFID_VZ]);
    c->f[AFID_RHS] = c->rhs4[2];
    return rc;
}
int afid_implicit_and_update_temp(afid_ctx *c) {
    NEED_SUBSTEP(c) NEED_FP64(c)
    int rc = launch_point(c, k_rhs_temp, false, "k_rhs_temp");
    // ImplicitAndUpdateTemp.F90:69-70: the immersed-boundary solve unless the phase field is on
    if (!rc) rc = c->ibm_mask[3] && !c->pf ? run_tri_ibm(c, 3, MAT_T, c->rhs4[3], c->f[AFID_TEMP])
                                           : run_tri(c, MAT_T, c->rhs4[3], c->f[AFID_TEMP]);
    c->f[AFID_RHS] = c->rhs4[3];
    return rc;
}
int afid_implicit_salinity(afid_ctx *c) {
    NEED_SUBSTEP(c) NEED_FP64(c) NEED_SAL(c)
    int rc = launch_point(c, k_rhs_sal, true, "k_rhs_sal");
    if (!rc) rc = c->ibm_mask[4] ? run_tri_ibm(c, 4, MAT_S, c->f[AFID_RHSR], c->f[AFID_SAL])     // salinity.F90:403-404
                                 : run_tri(c, MAT_S, c->f[AFID_RHSR], c->f[AFID_SAL]);
    return rc;
}

int afid_calc_local_divergence(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) return launch_point(c, k_local_divergence, false, "k_local_divergence"); }

// moisture.F90 (the per-routine call tree; afid_time_marcher fuses them)
int afid_set_humidity_bcs(afid_ctx *c) {
    NEED_MOIST(c)
    KP p = make_kp(c);
    const long long n = (long long)c->ny * c->nz;
    hipLaunchKernelGGL(k_humidity_bcs, dim3(unsigned((n + 255) / 256)), dim3(256), 0, c->stream, p, c->f[AFID_HUMBP],
                       c->f[AFID_HUMTP], c->c.q_fixn == 2 ? 1 : 0);
    return check_launch(c, "k_humidity_bcs");
}
int afid_explicit_humidity(afid_ctx *c) {
    NEED_MOIST(c) NEED_SUBSTEP(c) NEED_FP64(c)
    return launch_point(c, k_explicit_humid, false, "k_explicit_humid");
}
int afid_add_condensation(afid_ctx *c) {
    NEED_MOIST(c) NEED_SUBSTEP(c) NEED_FP64(c)
    return launch_point(c, k_add_condensation, false, "k_add_condensation");
}
int afid_implicit_humidity(afid_ctx *c) {
    NEED_MOIST(c) NEED_SUBSTEP(c) NEED_FP64(c)
    KP p = make_kp(c);
    hipLaunchKernelGGL(k_rhs_humid, point_grid(c->nxm, c->ny, c->nz), kPointBlock, 0, c->stream, p, c->rhs4[3]);
    int rc = check_launch(c, "k_rhs_humid");
    if (!rc) rc = run_tri(c, MAT_Q, c->rhs4[3], c->f[AFID_HUMID]);
    c->f[AFID_RHS] = c->rhs4[3];            // the reference's rhs holds the last solve's increment
    return rc;
}
int afid_update_saturation(afid_ctx *c) {
    NEED_MOIST(c)
    return launch_point(c, k_update_saturation, false, "k_update_saturation");
}

}  // extern "C"

// One Poisson sweep: forward transforms of the source, per-mode tridiagonals,
// inverse transforms into dph.  fwd_gate gates the forward y transform, gate
// the rest; slot_out receives max|div u| of a velocity source.
// z transform of the spectrum (k_fft_z; the FP32 form may take 16 values per
// thread and stage, see setup)
template <typename R>
static void launch_fft_z(afid_ctx *c, const KP &p, const FftPlan<R> &pz, cpx<R> *spec, int sign, Gate gate) {
    const dim3 gr(unsigned((c->nxm + c->Lz - 1) / c->Lz), unsigned(c->nymh));
    if constexpr (sizeof(R) == sizeof(float)) {
        if (c->fftz_maxv == 16) {
            hipLaunchKernelGGL((k_fft_z<R, 16>), gr, dim3(FFT_THREADS), c->lds_z, c->stream, p, pz, spec, c->Lz, sign, gate);
            return;
        }
    }
    hipLaunchKernelGGL((k_fft_z<R>), gr, dim3(FFT_THREADS), c->lds_z, c->stream, p, pz, spec, c->Lz, sign, gate);
}
template <typename R>
static int poisson_sweep(afid_ctx *c, const FftPlan<R> &py, const FftPlan<R> &pz, const PoissonFactors<R> &pf, int src,
                         Gate fwd_gate, Gate gate, unsigned long long *slot_out, unsigned long long *clear) {
    KP p = make_kp(c);
    cpx<R> *spec = static_cast<cpx<R> *>(c->spec);
    const int TXy = 2 * c->Ly;
    tstart(c, "poisson_fwd_y");
    hipLaunchKernelGGL(k_div_fft_y<R>, dim3(unsigned((c->nxm + TXy - 1) / TXy), unsigned(c->nz)), dim3(FFT_THREADS), c->lds_y,
                       c->stream, p, py, spec, c->Ly, src, fwd_gate, slot_out, clear);
    tstop(c);
    int rc = check_launch(c, "k_div_fft_y");
    if (rc) return rc;
    tstart(c, "poisson_fft_z");
    launch_fft_z<R>(c, p, pz, spec, -1, gate);
    tstop(c);
    if ((rc = check_launch(c, "k_fft_z"))) return rc;
    tstart(c, "poisson_tridiag");
    switch (c->pchunk) {
#define AFID_P_CASE(CH) case CH: hipLaunchKernelGGL((k_poisson_tri<CH, R>), dim3(unsigned(c->nymh), unsigned(c->nz / 2 + 1)), dim3(128), 0, c->stream, p, pf, spec, gate); break;
        AFID_P_CASE(1) AFID_P_CASE(2) AFID_P_CASE(3) AFID_P_CASE(4) AFID_P_CASE(5) AFID_P_CASE(6) AFID_P_CASE(7) AFID_P_CASE(8)
        AFID_P_CASE(9) AFID_P_CASE(10) AFID_P_CASE(11) AFID_P_CASE(12) AFID_P_CASE(13) AFID_P_CASE(14) AFID_P_CASE(15) AFID_P_CASE(16)
#undef AFID_P_CASE
    default: return set_err(c, AFID_E_ARG, "nxm too large");
    }
    tstop(c);
    if ((rc = check_launch(c, "k_poisson_tri"))) return rc;
    tstart(c, "poisson_fft_z");
    launch_fft_z<R>(c, p, pz, spec, +1, gate);
    tstop(c);
    if ((rc = check_launch(c, "k_fft_z"))) return rc;
    tstart(c, "poisson_inv_y");
    hipLaunchKernelGGL(k_ifft_y<R>, dim3(unsigned((c->nxm + TXy - 1) / TXy), unsigned(c->nz)), dim3(FFT_THREADS), c->lds_y,
                       c->stream, p, py, spec, c->Ly, gate);
    tstop(c);
    return check_launch(c, "k_ifft_y");
}

// FP64 form: one sweep, from dph (from_dph = 1) or fused with the divergence (0)
static int poisson(afid_ctx *c, int from_dph) {
    const Gate open{nullptr, 0.0};
    return poisson_sweep<double>(c, c->plan_y, c->plan_z, c->pois, from_dph ? SRC_DPH : SRC_VELOCITY, open, open, nullptr, nullptr);
}

static int project(afid_ctx *c, int vel, int pr, Gate gate = Gate{nullptr, 0.0}) {
    KP p = make_kp(c);
    dim3 gr = point_grid(c->nxm, c->ny, c->nz);
    if (gate.slot) gr.z = unsigned(std::min(c->nz, 32));
    if (c->mixed) hipLaunchKernelGGL(k_project<float>, gr, kPointBlock, 0, c->stream, p, vel, pr, gate);
    else hipLaunchKernelGGL(k_project<double>, gr, kPointBlock, 0, c->stream, p, vel, pr, gate);
    return check_launch(c, "k_project");
}

// Mixed form: pressure solve and projection of one substep as FP32 sweeps
// inside FP64-residual iterative refinement.  Sweep 1 solves for the
// divergence of the velocity and is applied by k_project_res, which also
// forms the FP64 divergence of the corrected field and max-reduces it into
// slots[1].  Sweep s > 1 transforms the divergence of the current velocity
// (the FP64 residual left by sweep s-1), solves, and projects in place,
// adding its correction to the pressure; it runs only while slots[s-1] >
// div_tol.  The gates are read on the device, so the step never waits for
// the host.
static int poisson_mixed(afid_ctx *c) {
    const Gate open{nullptr, 0.0};
    int rc = poisson_sweep<float>(c, c->plan_yf, c->plan_zf, c->pff, SRC_VELOCITY, open, open, nullptr, c->slots);
    if (rc) return rc;
    {
        KP p = make_kp(c);
        tstart(c, "project");
        const dim3 gr = point_grid(c->nxm, c->ny, (c->nz + PROJ_ZC - 1) / PROJ_ZC);
        hipLaunchKernelGGL(k_project_res, gr, kPointBlock, 0, c->stream, p, c->alt[0], c->alt[1], c->alt[2], c->redsum);
        hipLaunchKernelGGL(k_max_partials, dim3(256), dim3(256), 0, c->stream, c->redsum, (long long)gr.x * gr.y * gr.z,
                           c->slots + 1);
        tstop(c);
        if ((rc = check_launch(c, "k_project_res"))) return rc;
        std::swap(c->f[AFID_VX], c->alt[0]);
        std::swap(c->f[AFID_VY], c->alt[1]);
        std::swap(c->f[AFID_VZ], c->alt[2]);
    }
    for (int s = 2; s <= c->max_sweeps; ++s) {
        // sweep s runs while the divergence left by sweep s-1 (slots[s-1]) is
        // above div_tol; for s > 2 its forward transform is what measures
        // that divergence, so it is gated on the sweep before (slots[s-2])
        const Gate fwd{c->slots + (s == 2 ? 1 : s - 2), c->div_tol};
        const Gate rest{c->slots + (s - 1), c->div_tol};
        rc = poisson_sweep<float>(c, c->plan_yf, c->plan_zf, c->pff, SRC_VELOCITY, fwd, rest,
                                  s > 2 ? c->slots + (s - 1) : nullptr, nullptr);
        if (rc) return rc;
        tstart(c, "project");
        rc = project(c, 1, 1, rest);
        tstop(c);
        if (rc) return rc;
    }
    return AFID_OK;
}

// vyr / vzr, one wavefront per refined line (k_vyz_wave)
template <int AX, bool TWO>
static void launch_vyz_wave(afid_ctx *c, const KP &p, dim3 grid, int nxc) {
    if (c->mixed) hipLaunchKernelGGL((k_vyz_wave<AX, float, TWO>), grid, dim3(256), 0, c->stream, p, nxc);
    else hipLaunchKernelGGL((k_vyz_wave<AX, double, TWO>), grid, dim3(256), 0, c->stream, p, nxc);
}

extern "C" {

int afid_solve_pressure_correction(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) return poisson(c, 1); }

int afid_correct_velocity(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) return project(c, 1, 0); }
int afid_correct_pressure(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) return project(c, 0, 1); }

static int interp_vel(afid_ctx *c) {
    KP p = make_kp(c);
    int rc;
    const dim3 gx(tile_grid((long long)((c->nyr + VXR_TJ - 1) / VXR_TJ) * ((c->nzr + VXR_TI - 1) / VXR_TI)));
    const dim3 gy(unsigned((c->nxmr + 63) / 64), unsigned((c->nzr + VYZ_LINES - 1) / VYZ_LINES));
    const dim3 gz(unsigned((c->nxmr + 63) / 64), unsigned((c->nyr + VYZ_LINES - 1) / VYZ_LINES));
    tstart(c, "mgrd_vxr");
    if (c->mixed) hipLaunchKernelGGL(k_vxr_tile<float>, gx, dim3(256), c->lds_vxr, c->stream, p, c->vxr_xc, c->vxr_cj, c->vxr_ci);
    else hipLaunchKernelGGL(k_vxr_tile<double>, gx, dim3(256), c->lds_vxr, c->stream, p, c->vxr_xc, c->vxr_cj, c->vxr_ci);
    tstop(c);
    if ((rc = check_launch(c, "k_vxr_tile"))) return rc;
    if (c->vyz_tile) {
        tstart(c, "mgrd_vyr");
        if (c->mixed) hipLaunchKernelGGL((k_vyz_tile<1, float>), gy, dim3(64 * VYZ_LINES), c->lds_vy, c->stream, p, c->vxr_xc, c->vy_cs, c->vy_co);
        else hipLaunchKernelGGL((k_vyz_tile<1, double>), gy, dim3(64 * VYZ_LINES), c->lds_vy, c->stream, p, c->vxr_xc, c->vy_cs, c->vy_co);
        tstop(c);
        if ((rc = check_launch(c, "k_vyz_tile<1>"))) return rc;
        tstart(c, "mgrd_vzr");
        if (c->mixed) hipLaunchKernelGGL((k_vyz_tile<2, float>), gz, dim3(64 * VYZ_LINES), c->lds_vz, c->stream, p, c->vxr_xc, c->vz_cs, c->vz_co);
        else hipLaunchKernelGGL((k_vyz_tile<2, double>), gz, dim3(64 * VYZ_LINES), c->lds_vz, c->stream, p, c->vxr_xc, c->vz_cs, c->vz_co);
        tstop(c);
        return check_launch(c, "k_vyz_tile<2>");
    }
    // one wavefront per refined line: (x chunks) x (groups of 4 lines)
    // workgroups, padded to a multiple of 8 for the XCD map
    // (a chunk of 64 refined x spanning more than 64 coarse points, refinement
    // ratio below 1.05 in x, takes the second register per lane)
    const int nxc = (c->nxmr + 63) / 64;
    const dim3 wy(unsigned(((long long)nxc * ((c->nzr + 3) / 4) + 7) / 8 * 8));
    const dim3 wz(unsigned(((long long)nxc * ((c->nyr + 3) / 4) + 7) / 8 * 8));
    tstart(c, "mgrd_vyr");
    if (c->vxr_xc > 64) launch_vyz_wave<1, true>(c, p, wy, nxc);
    else launch_vyz_wave<1, false>(c, p, wy, nxc);
    tstop(c);
    if ((rc = check_launch(c, "k_vyz_wave<1>"))) return rc;
    tstart(c, "mgrd_vzr");
    if (c->vxr_xc > 64) launch_vyz_wave<2, true>(c, p, wz, nxc);
    else launch_vyz_wave<2, false>(c, p, wz, nxc);
    tstop(c);
    return check_launch(c, "k_vyz_wave<2>");
}
int afid_interp_vel_mgrd(afid_ctx *c) { NEED_SAL(c) return interp_vel(c); }

// refined -> coarse cell centres (k_interp_coarse_tile)
// (FP32 weights of the table for the mixed form's FP32 refined fields)
static int to_coarse(afid_ctx *c, CoarseJob jb, const CoarseTab &t, const char *timer) {
    KP p = make_kp(c);
    jb.x_src = t.x_src; jb.y_src = t.y_src; jb.z_src = t.z_src;
    if (c->mixed) { jb.x_w = t.x_wf; jb.y_w = t.y_wf; jb.z_w = t.z_wf; }
    else { jb.x_w = t.x_w; jb.y_w = t.y_w; jb.z_w = t.z_w; }
    const dim3 gr(tile_grid((long long)((c->nxm + 63) / 64) * ((c->ny + SALC_TJ - 1) / SALC_TJ) * ((c->nz + SALC_TZ - 1) / SALC_TZ)));
    tstart(c, timer);
    if (c->mixed) hipLaunchKernelGGL(k_interp_coarse_tile<float>, gr, dim3(256), t.lds, c->stream, p, jb, t.xr, t.yr);
    else hipLaunchKernelGGL(k_interp_coarse_tile<double>, gr, dim3(256), t.lds, c->stream, p, jb, t.xr, t.yr);
    tstop(c);
    return check_launch(c, "k_interp_coarse_tile");
}

int afid_interp_sal_multigrid(afid_ctx *c) {
    NEED_SAL(c)
    CoarseJob jb{};
    jb.mode = CM_SAL;
    if (c->mixed) { jb.src = c->g[AFID_SAL]; jb.out = c->g[AFID_SALC]; }
    else { jb.src = c->f[AFID_SAL]; jb.out = c->f[AFID_SALC]; }
    jb.scale = 1.0;
    return to_coarse(c, jb, c->tab_sal, "mgrd_interp_sal");
}

#define NEED_PF(c) if (!(c)->pf) return set_err((c), AFID_E_STATE, "the phase field is not enabled");

int afid_explicit_phase(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) NEED_PF(c) return launch_point(c, k_explicit_phase, true, "k_explicit_phase"); }
int afid_add_volume_penalty(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) NEED_PF(c) return launch_point(c, k_volume_penalty, false, "k_volume_penalty"); }
int afid_implicit_phase(afid_ctx *c) {
    NEED_SUBSTEP(c) NEED_FP64(c) NEED_PF(c)
    int rc = launch_point(c, k_rhs_phase, true, "k_rhs_phase");
    if (!rc) rc = run_tri(c, MAT_P, c->f[AFID_RHSR], c->f[AFID_PHI]);   // rhsr keeps the increment
    return rc;
}
static int latent_heat(afid_ctx *c, double *out, int accumulate) {
    CoarseJob jb{};
    jb.mode = CM_LAT; jb.out = out;
    jb.src = c->mixed ? static_cast<const void *>(c->g[AFID_RHSR]) : static_cast<const void *>(c->f[AFID_RHSR]);
    jb.scale = c->c.pf_s / (c->al * c->dt);          // pf_S d(phi)/(al dt), phasefield.F90:666,701
    jb.accumulate = accumulate;
    return to_coarse(c, jb, c->tab_phi, "pf_latent_heat");
}
int afid_add_latent_heat(afid_ctx *c) { NEED_SUBSTEP(c) NEED_FP64(c) NEED_PF(c) return latent_heat(c, c->f[AFID_HRO], 1); }
int afid_interp_phi_multigrid(afid_ctx *c) {
    NEED_PF(c)
    CoarseJob jb{};
    jb.mode = CM_PHI; jb.out = c->f[AFID_PHIC];
    jb.src = c->mixed ? static_cast<const void *>(c->g[AFID_PHI]) : static_cast<const void *>(c->f[AFID_PHI]);
    jb.scale = 1.0;
    return to_coarse(c, jb, c->tab_phi, "pf_interp_phi");
}
int afid_interp_temp_multigrid(afid_ctx *c) {
    NEED_PF(c)
    KP p = make_kp(c);
    tstart(c, "pf_interp_temp");
    const dim3 gr(unsigned((c->nyr + VXR_TJ - 1) / VXR_TJ), unsigned((c->nzr + VXR_TI - 1) / VXR_TI));
    if (c->mixed) hipLaunchKernelGGL(k_interp_temp_tile<float>, gr, dim3(256), c->lds_vxr, c->stream, p, c->vxr_xc, c->vxr_cj, c->vxr_ci);
    else hipLaunchKernelGGL(k_interp_temp_tile<double>, gr, dim3(256), c->lds_vxr, c->stream, p, c->vxr_xc, c->vxr_cj, c->vxr_ci);
    tstop(c);
    return check_launch(c, "k_interp_temp_tile");
}

}  // extern "C"

// fused column kernels of the step (one sweep per grid, ping-pong buffers)
template <int CH, typename HT, bool PF>
static void launch_column_base_form(afid_ctx *c, const KP &p) {
    BaseOut o{c->alt[0], c->alt[1], c->alt[2], c->alt[3]};
    const long long nblk = strip_blocks(c->ny, c->nz);
    hipLaunchKernelGGL((k_column_base<CH, HT, PF>), dim3(unsigned((nblk + 7) / 8 * 8)), dim3(256), 0, c->stream, p,
                       factors(c, MAT_X), factors(c, MAT_YZ), factors(c, MAT_T), o);
}
template <int CH>
static void launch_column_base(afid_ctx *c, const KP &p) {
    if (c->mixed && c->pf) launch_column_base_form<CH, float, true>(c, p);
    else if (c->mixed) launch_column_base_form<CH, float, false>(c, p);
    else if (c->pf) launch_column_base_form<CH, double, true>(c, p);
    else launch_column_base_form<CH, double, false>(c, p);
}
// single-pass kernel; CH here is the chunk of the nxm-row systems
template <int CH, typename HT, bool PF>
static void launch_column_base1_form(afid_ctx *c, const KP &p) {
    BaseOut o{c->alt[0], c->alt[1], c->alt[2], c->alt[3]};
    const dim3 gr(unsigned((strip_blocks(c->ny, c->nz) + 7) / 8 * 8));
    hipLaunchKernelGGL((k_column_base1<CH, HT, PF>), gr, dim3(256), 0, c->stream, p, factors(c, MAT_X), factors(c, MAT_YZ),
                       factors(c, MAT_T), o);
}
template <int CH>
static void launch_column_base1(afid_ctx *c, const KP &p) {
    if (c->mixed && c->pf) launch_column_base1_form<CH, float, true>(c, p);
    else if (c->mixed) launch_column_base1_form<CH, float, false>(c, p);
    else if (c->pf) launch_column_base1_form<CH, double, true>(c, p);
    else launch_column_base1_form<CH, double, false>(c, p);
}
// ExplicitHumidity + ImplicitHumidity + SolveImpEqnUpdate_Humidity of every
// column into humid_alt; the caller swaps the buffers once the base column
// kernel (which reads the old humidity for the condensation heat) has run
template <int CH>
static void launch_column_humid(afid_ctx *c, const KP &p) {
    const long long nblk = strip_blocks(c->ny, c->nz);
    hipLaunchKernelGGL((k_column_humid<CH>), dim3(unsigned((nblk + 7) / 8 * 8)), dim3(256), 0, c->stream, p, factors(c, MAT_Q),
                       c->humid_alt);
}
static int column_humid(afid_ctx *c) {
    KP p = make_kp(c);
    switch (c->chunk[MAT_Q]) {
#define AFID_CQ_CASE(CH) case CH: launch_column_humid<CH>(c, p); break;
        AFID_CQ_CASE(1) AFID_CQ_CASE(2) AFID_CQ_CASE(3) AFID_CQ_CASE(4) AFID_CQ_CASE(5) AFID_CQ_CASE(6)
        AFID_CQ_CASE(7) AFID_CQ_CASE(8) AFID_CQ_CASE(9) AFID_CQ_CASE(10) AFID_CQ_CASE(11) AFID_CQ_CASE(12)
        AFID_CQ_CASE(13) AFID_CQ_CASE(14) AFID_CQ_CASE(15) AFID_CQ_CASE(16)
#undef AFID_CQ_CASE
    default: return set_err(c, AFID_E_ARG, "nxm too large");
    }
    return check_launch(c, "k_column_humid");
}
static int column_base(afid_ctx *c) {
    KP p = make_kp(c);
    // single pass up to 512 wall-normal cells (its registers hold two
    // right-hand sides of CH rows), four passes above
    if (!c->col_four_pass && c->chunk[MAT_YZ] <= 8) {
        switch (c->chunk[MAT_YZ]) {
#define AFID_CB1_CASE(CH) case CH: launch_column_base1<CH>(c, p); break;
            AFID_CB1_CASE(1) AFID_CB1_CASE(2) AFID_CB1_CASE(3) AFID_CB1_CASE(4) AFID_CB1_CASE(5) AFID_CB1_CASE(6)
            AFID_CB1_CASE(7) AFID_CB1_CASE(8)
#undef AFID_CB1_CASE
        }
    } else {
        switch ((c->nx + 63) / 64) {
#define AFID_CB_CASE(CH) case CH: launch_column_base<CH>(c, p); break;
            AFID_CB_CASE(1) AFID_CB_CASE(2) AFID_CB_CASE(3) AFID_CB_CASE(4) AFID_CB_CASE(5) AFID_CB_CASE(6)
            AFID_CB_CASE(7) AFID_CB_CASE(8) AFID_CB_CASE(9) AFID_CB_CASE(10) AFID_CB_CASE(11) AFID_CB_CASE(12)
            AFID_CB_CASE(13) AFID_CB_CASE(14) AFID_CB_CASE(15) AFID_CB_CASE(16)
#undef AFID_CB_CASE
        default: return set_err(c, AFID_E_ARG, "nx too large");
        }
    }
    int rc = check_launch(c, "k_column_base");
    if (rc) return rc;
    std::swap(c->f[AFID_VX], c->alt[0]);
    std::swap(c->f[AFID_VY], c->alt[1]);
    std::swap(c->f[AFID_VZ], c->alt[2]);
    std::swap(c->f[AFID_TEMP], c->alt[3]);
    return AFID_OK;
}
template <int CH>
static void launch_column_sal(afid_ctx *c, const KP &p, long long nblk) {
    const dim3 gr(unsigned((nblk + 7) / 8 * 8));
    if (c->mixed) hipLaunchKernelGGL((k_column_sal<CH, float>), gr, dim3(256), 0, c->stream, p, factors(c, MAT_S), c->sal_alt_f);
    else hipLaunchKernelGGL((k_column_sal<CH, double>), gr, dim3(256), 0, c->stream, p, factors(c, MAT_S), c->sal_alt);
}
static int column_sal(afid_ctx *c) {
    KP p = make_kp(c);
    const long long nblk = strip_blocks(c->nyr, c->nzr);
    const int ch = (c->nxmr + 63) / 64;
    switch (ch) {
#define AFID_CS_CASE(CH) case CH: launch_column_sal<CH>(c, p, nblk); break;
        AFID_CS_CASE(1) AFID_CS_CASE(2) AFID_CS_CASE(3) AFID_CS_CASE(4) AFID_CS_CASE(5) AFID_CS_CASE(6)
        AFID_CS_CASE(7) AFID_CS_CASE(8) AFID_CS_CASE(9) AFID_CS_CASE(10) AFID_CS_CASE(11) AFID_CS_CASE(12)
        AFID_CS_CASE(13) AFID_CS_CASE(14) AFID_CS_CASE(15) AFID_CS_CASE(16)
#undef AFID_CS_CASE
    default: return set_err(c, AFID_E_ARG, "nxmr too large");
    }
    int rc = check_launch(c, "k_column_sal");
    if (rc) return rc;
    if (c->mixed) std::swap(c->g[AFID_SAL], c->sal_alt_f);
    else std::swap(c->f[AFID_SAL], c->sal_alt);
    return AFID_OK;
}

template <int CH>
static void launch_column_phi(afid_ctx *c, const KP &p, long long nblk) {
    const dim3 gr(unsigned((nblk + 7) / 8 * 8));
    if (c->mixed) hipLaunchKernelGGL((k_column_phi<CH, float>), gr, dim3(256), 0, c->stream, p, factors(c, MAT_P), c->phi_alt_f);
    else hipLaunchKernelGGL((k_column_phi<CH, double>), gr, dim3(256), 0, c->stream, p, factors(c, MAT_P), c->phi_alt);
}
static int column_phi(afid_ctx *c) {
    KP p = make_kp(c);
    const long long nblk = strip_blocks(c->nyr, c->nzr);
    const int ch = (c->nxmr + 63) / 64;
    switch (ch) {
#define AFID_CP_CASE(CH) case CH: launch_column_phi<CH>(c, p, nblk); break;
        AFID_CP_CASE(1) AFID_CP_CASE(2) AFID_CP_CASE(3) AFID_CP_CASE(4) AFID_CP_CASE(5) AFID_CP_CASE(6)
        AFID_CP_CASE(7) AFID_CP_CASE(8) AFID_CP_CASE(9) AFID_CP_CASE(10) AFID_CP_CASE(11) AFID_CP_CASE(12)
        AFID_CP_CASE(13) AFID_CP_CASE(14) AFID_CP_CASE(15) AFID_CP_CASE(16)
#undef AFID_CP_CASE
    default: return set_err(c, AFID_E_ARG, "nxmr too large");
    }
    int rc = check_launch(c, "k_column_phi");
    if (rc) return rc;
    if (c->mixed) std::swap(c->g[AFID_PHI], c->phi_alt_f);
    else std::swap(c->f[AFID_PHI], c->phi_alt);
    return AFID_OK;
}

extern "C" {

// TimeMarcher.F90:13-165
int afid_time_marcher(afid_ctx *c, double dt) {
    if (!c) return AFID_E_ARG;
    for (int q = 0; q < 5; ++q)
        if (c->ibm_mask[q]) return set_err(c, AFID_E_STATE, "the immersed-boundary solves run through the per-routine call tree");
    for (auto &t : c->timers) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    c->timers.clear();
    int rc;
    for (int ns = 1; ns <= c->c.nsst; ++ns) {
        if ((rc = afid_set_substep(c, dt, ns))) return rc;
        if (c->pf) {
            // ExplicitPhase + ImplicitPhase (phi, ruphi, rhsr), then the
            // latent heat into qlat, read by the base column as hro
            tstart(c, "pf_column_phi");
            if ((rc = column_phi(c))) return rc;
            tstop(c);
            if ((rc = latent_heat(c, c->f[AFID_HRO], 0))) return rc;
        }
        if (c->moist) {
            // TimeMarcher.F90:79-81, :97: the humidity column reads the old
            // velocity and humidity, the base column reads the old humidity
            // for the condensation heat; both write second buffers
            tstart(c, "moist_column_humid");
            if ((rc = column_humid(c))) return rc;
            tstop(c);
        }
        tstart(c, "column_base");
        if ((rc = column_base(c))) return rc;
        tstop(c);
        if (c->moist) std::swap(c->f[AFID_HUMID], c->humid_alt);
        if (c->sal) {
            tstart(c, "column_sal");
            if ((rc = column_sal(c))) return rc;
            tstop(c);
        }
        if (c->mixed) {
            if ((rc = poisson_mixed(c))) return rc;
        } else {
            if ((rc = poisson(c, 0))) return rc;
            tstart(c, "project");
            if ((rc = project(c, 1, 1))) return rc;
            tstop(c);
        }
        if (c->sal) {
            if ((rc = interp_vel(c))) return rc;
            if ((rc = afid_interp_sal_multigrid(c))) return rc;
        }
        if (c->pf) {
            if ((rc = afid_interp_temp_multigrid(c))) return rc;
            if ((rc = afid_interp_phi_multigrid(c))) return rc;
        }
        if (c->moist) {                           // TimeMarcher.F90:154
            tstart(c, "moist_saturation");
            if ((rc = afid_update_saturation(c))) return rc;
            tstop(c);
        }
    }
    return AFID_OK;
}

// afid_time_marcher replayed from a captured graph.  Two steps are captured
// (kernel nodes only: the step holds no memset or copy; the Poisson gates and
// the refinement slots live on the device) and replayed nsteps / 2 times; an
// odd last step runs eagerly.  A new dt (the factors and the substep
// scalars are baked into the nodes) or moved buffers recapture.
int afid_time_marcher_graph(afid_ctx *c, double dt, int nsteps) {
    if (!c || nsteps < 0) return AFID_E_ARG;
    int rc;
    bool same = c->graph && c->graph_dt == dt;
    for (int f = 0; same && f < AFID_NFIELDS; ++f) same = c->graph_ptrs[f] == c->f[f];
    if (!same && nsteps >= 2) {
        if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
        // the factors of dt are computed and copied before the capture
        if ((rc = afid_set_substep(c, dt, 1))) return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const bool timing = c->timing;
        c->timing = false;
        for (auto &t : c->timers) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
        c->timers.clear();
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = afid_time_marcher(c, dt);
        if (!rc) rc = afid_time_marcher(c, dt);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->stream, &g);
        c->timing = timing;
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (ec != hipSuccess) return set_err(c, AFID_E_HIP, std::string("graph capture: ") + hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) { c->graph = nullptr; return set_err(c, AFID_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ei)); }
        c->graph_dt = dt;
        for (int f = 0; f < AFID_NFIELDS; ++f) c->graph_ptrs[f] = c->f[f];
    }
    int done = 0;
    if (c->graph)
        for (; done + 2 <= nsteps; done += 2) HIPCHK(c, hipGraphLaunch(c->graph, c->stream));
    for (; done < nsteps; ++done)
        if ((rc = afid_time_marcher(c, dt))) return rc;
    return AFID_OK;
}

// CreateInitialConditions.F90:150-201 (RBC branch) with u = 0 and seeded
// noise; also the scalar of CreateInitialSalinity on the refined grid.
int afid_create_initial_conditions(afid_ctx *c, uint64_t seed, double eps) {
    if (!c) return AFID_E_ARG;
    const size_t nb = size_t(c->nz) * c->ny * c->ldx;
    for (int f : {AFID_VX, AFID_VY, AFID_VZ, AFID_PR, AFID_RUX, AFID_RUY, AFID_RUZ, AFID_RUTEMP}) {
        if (c->g[f]) HIPCHK(c, hipMemsetAsync(c->g[f], 0, sizeof(float) * nb, c->stream));
        else HIPCHK(c, hipMemsetAsync(c->f[f], 0, sizeof(double) * nb, c->stream));
    }
    hipLaunchKernelGGL(k_init_scalar<double>, point_grid(c->nx, c->ny, c->nz), kPointBlock, 0, c->stream, c->f[AFID_TEMP],
                       c->f[AFID_TEMPBP], c->f[AFID_TEMPTP], c->d_xm, c->nxm, c->nx, c->ldx, c->ny, c->nz, c->c.alx3, eps,
                       (unsigned long long)seed);
    int rc = check_launch(c, "k_init_scalar");
    if (rc) return rc;
    if (c->pf) {
        // set_ice_sphere (phasefield.F90:311-341) of radius pf_r0, then
        // tempr and phic as at start-up (main.F90:210-212)
        KP p = make_kp(c);
        const size_t nr = size_t(c->nzr) * c->nyr * c->ldxr;
        for (int f : {AFID_RUPHI, AFID_HPHI, AFID_RHSR}) {
            if (c->g[f]) HIPCHK(c, hipMemsetAsync(c->g[f], 0, sizeof(float) * nr, c->stream));
            else if (c->f[f]) HIPCHK(c, hipMemsetAsync(c->f[f], 0, sizeof(double) * nr, c->stream));
        }
        hipLaunchKernelGGL(k_init_temp_sphere, point_grid(c->nxm, c->ny, c->nz), kPointBlock, 0, c->stream, p, c->d_xm, c->pf_r0,
                           c->c.alx3, c->c.ylen, c->c.zlen);
        if (c->mixed)
            hipLaunchKernelGGL(k_init_phi_sphere<float>, point_grid(c->nxr, c->nyr, c->nzr), kPointBlock, 0, c->stream, p, c->d_xmr,
                               c->pf_r0, c->pf_eps, c->c.alx3, c->c.ylen, c->c.zlen);
        else
            hipLaunchKernelGGL(k_init_phi_sphere<double>, point_grid(c->nxr, c->nyr, c->nzr), kPointBlock, 0, c->stream, p, c->d_xmr,
                               c->pf_r0, c->pf_eps, c->c.alx3, c->c.ylen, c->c.zlen);
        if ((rc = check_launch(c, "k_init_phi_sphere"))) return rc;
        if ((rc = afid_interp_temp_multigrid(c))) return rc;
        return afid_interp_phi_multigrid(c);
    }
    if (!c->sal) return rc;
    const size_t nr = size_t(c->nzr) * c->nyr * c->ldxr;
    for (int f : {AFID_RUSAL, AFID_VXR, AFID_VYR, AFID_VZR}) {
        if (c->g[f]) HIPCHK(c, hipMemsetAsync(c->g[f], 0, sizeof(float) * nr, c->stream));
        else HIPCHK(c, hipMemsetAsync(c->f[f], 0, sizeof(double) * nr, c->stream));
    }
    if (c->mixed)
        hipLaunchKernelGGL(k_init_scalar<float>, point_grid(c->nxr, c->nyr, c->nzr), kPointBlock, 0, c->stream, c->g[AFID_SAL],
                           c->f[AFID_SALBP], c->f[AFID_SALTP], c->d_xmr, c->nxmr, c->nxr, c->ldxr, c->nyr, c->nzr, c->c.alx3, eps,
                           (unsigned long long)(seed + 1));
    else
        hipLaunchKernelGGL(k_init_scalar<double>, point_grid(c->nxr, c->nyr, c->nzr), kPointBlock, 0, c->stream, c->f[AFID_SAL],
                           c->f[AFID_SALBP], c->f[AFID_SALTP], c->d_xmr, c->nxmr, c->nxr, c->ldxr, c->nyr, c->nzr, c->c.alx3, eps,
                           (unsigned long long)(seed + 1));
    if ((rc = check_launch(c, "k_init_scalar"))) return rc;
    // salc from the initial S (the reference interpolates at start-up, main.F90:206-213)
    return afid_interp_sal_multigrid(c);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// restart onto a new grid (ReadFlowInterp.F90:90-141, IC_interpolation/)
// ---------------------------------------------------------------------------
// device buffers of one restart call, freed after the stream has drained
struct RestartScratch {
    afid_ctx *c;
    std::vector<void *> p;
    explicit RestartScratch(afid_ctx *ctx) : c(ctx) {}
    ~RestartScratch() {
        (void)hipStreamSynchronize(c->stream);
        for (void *q : p) (void)hipFree(q);
    }
    template <typename T>
    int alloc(size_t n, T **out) {
        void *q = nullptr;
        const hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
        if (e != hipSuccess) return set_err(c, AFID_E_MEM, std::string("restart scratch: ") + hipGetErrorString(e));
        p.push_back(q);
        *out = static_cast<T *>(q);
        return AFID_OK;
    }
    template <typename T>
    int upload(const std::vector<T> &v, const T **out) {
        T *q;
        int rc = alloc(v.size(), &q);
        if (rc) return rc;
        HIPCHK(c, hipMemcpyAsync(q, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, c->stream));
        *out = q;
        return AFID_OK;
    }
};

// CreateOldGrid.F90:11-115 (one grid of the old run) and the old -> new
// cell-centre tables of CreateNewInputStencil.F90 / CreateNewSalStencil.F90.
// Old cell counts nxmo x nymo x nzmo; returns the old faces xco (0-based).
static int restart_tables(afid_ctx *c, RestartScratch &sc, int nxmo, int nymo, int nzmo, int istr3o, double str3o,
                          bool refined, RestartTabs &t, const double **xco_dev) {
    const afid_config &g = c->c;
    if (nxmo < 2 || nymo < 1 || nzmo < 1) return set_err(c, AFID_E_ARG, "restart: the old grid is too small");
    int opt = istr3o;
    if (!refined && g.istr3 == 5) opt = 5;     // CreateOldGrid.F90:77 tests the current istr3
    else if (!refined && istr3o == 5) return set_err(c, AFID_E_ARG, "restart: old grid option 5 needs istr3 = 5 (the reference leaves it unset)");
    FArr xco, xmo;
    try {
        // option 3 takes the parameter 1.0 on both old grids (CreateOldGrid.F90:62-67)
        make_wall_grid(opt, nxmo, g.alx3, str3o, opt == 5 ? g.dpdy : 1.0, xco, xmo);
    } catch (const std::exception &e) {
        return set_err(c, AFID_E_ARG, std::string("restart: ") + e.what());
    }
    auto uni = [](int n, double L) {
        FArr m(1, n);
        for (int i = 1; i <= n; ++i) m(i) = 0.5 * (L * double(i - 1) / n + L * double(i) / n);
        return m;
    };
    const int nxn = refined ? c->nxmr : c->nxm, nyn = refined ? c->nyr : c->ny, nzn = refined ? c->nzr : c->nz;
    const std::vector<double> &xmv = c->host1d[refined ? "xmr" : "xm"];
    FArr xn(1, nxn), xo(1, nxmo);
    for (int k = 1; k <= nxn; ++k) xn(k) = xmv[size_t(k - 1)];
    for (int k = 1; k <= nxmo; ++k) xo(k) = xmo(k);
    FArr yo = uni(nymo, g.ylen), zo = uni(nzmo, g.zlen), yn = uni(nyn, g.ylen), zn = uni(nzn, g.zlen);
    Stencil sx, sy, sz;
    try {
        sx = construct_stencil(xo, nxmo, xn, nxn, g.alx3, interpolation_indices(xo, nxmo, xn, nxn, g.alx3), 'x');
        sy = construct_stencil(yo, nymo, yn, nyn, g.ylen, interpolation_indices(yo, nymo, yn, nyn, g.ylen), 'y');
        sz = construct_stencil(zo, nzmo, zn, nzn, g.zlen, interpolation_indices(zo, nzmo, zn, nzn, g.zlen), 'z');
    } catch (const std::exception &e) {
        return set_err(c, AFID_E_ARG, std::string("restart: ") + e.what());
    }
    int rc;
    if ((rc = sc.upload(sx.src, &t.xs)) || (rc = sc.upload(sy.src, &t.ys)) || (rc = sc.upload(sz.src, &t.zs))) return rc;
    if ((rc = sc.upload(sx.w, &t.xw)) || (rc = sc.upload(sy.w, &t.yw)) || (rc = sc.upload(sz.w, &t.zw))) return rc;
    return sc.upload(farr_range(xco, 1, nxmo + 1), xco_dev);
}

static int check_old_grid(afid_ctx *c, const afid_old_grid *og, bool refined) {
    const int nx = refined ? og->nxro : og->nxo, ny = refined ? og->nyro : og->nyo, nz = refined ? og->nzro : og->nzo;
    if (nx < 3 || ny < 2 || nz < 2 || nx > 4097 || ny > 65536 || nz > 65536)
        return set_err(c, AFID_E_ARG, "restart: old grid sizes out of range");
    if (std::fabs(og->ylen - c->c.ylen) > 1e-8 || std::fabs(og->zlen - c->c.zlen) > 1e-8)    // ReadFlowInterp.F90:76-82
        return set_err(c, AFID_E_ARG, "restart: continua domain size does not match the configuration");
    return AFID_OK;
}

extern "C" {

// InterpInputVel.F90:1-350 (with CreateOldGrid and CreateNewInputStencil)
int afid_interp_input_vel(afid_ctx *c, const afid_old_grid *og, const double *vxo, const double *vyo, const double *vzo,
                          const double *tempo) {
    if (!c || !og || !vxo || !vyo || !vzo || !tempo) return AFID_E_ARG;
    int rc;
    if ((rc = check_old_grid(c, og, false))) return rc;
    const int nxo = og->nxo, nyo = og->nyo - 1, nzo = og->nzo - 1, nxmo = nxo - 1;
    RestartScratch sc(c);
    RestartTabs t{};
    const double *xco;
    if ((rc = restart_tables(c, sc, nxmo, nyo, nzo, og->istr3o, og->str3o, false, t, &xco))) return rc;
    const size_t nold = size_t(nxo) * nyo * nzo, next = size_t(nxo + 3) * nyo * nzo;
    const size_t nb = size_t(c->nz) * c->ny * c->ldx;
    double *F, *E, *G;
    const double *lxa;
    if ((rc = sc.alloc(nold, &F)) || (rc = sc.alloc(next, &E)) || (rc = sc.alloc(nb, &G))) return rc;
    {
        const std::vector<double> &xc = c->host1d["xc"];
        std::vector<double> l;
        for (int k = 0; k < c->nxm; ++k) l.push_back(xc[size_t(k + 1)] - xc[size_t(k)]);   // lxa, InterpInputVel.F90:75
        if ((rc = sc.upload(l, &lxa))) return rc;
    }
    // the new state starts from zero (pr and the RK history are not in the checkpoint)
    for (int f : {AFID_VX, AFID_VY, AFID_VZ, AFID_TEMP, AFID_PR, AFID_RUX, AFID_RUY, AFID_RUZ, AFID_RUTEMP}) {
        if (c->g[f]) HIPCHK(c, hipMemsetAsync(c->g[f], 0, sizeof(float) * nb, c->stream));
        else if (c->f[f]) HIPCHK(c, hipMemsetAsync(c->f[f], 0, sizeof(double) * nb, c->stream));
    }
    const double dyo = double(nyo) / c->c.ylen, dzo = double(nzo) / c->c.zlen;
    RestartGhost vel{0, double(1 - 2 * c->c.inslws), double(1 - 2 * c->c.inslwn), 0, 0, 0.0, 0.0};
    const dim3 sgr(unsigned((nxo + 3 + 255) / 256), unsigned(nyo), unsigned(nzo));
    const dim3 tgr = point_grid(c->nxm, c->ny, c->nz);
    const double *old[3] = {vxo, vyo, vzo};
    double *vel_out[3] = {c->f[AFID_VX], c->f[AFID_VY], c->f[AFID_VZ]};
    for (int comp = 0; comp < 3; ++comp) {
        HIPCHK(c, hipMemcpyAsync(F, old[comp], sizeof(double) * nold, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_restart_stage, sgr, dim3(256), 0, c->stream, F, comp, nxo, nyo, nzo, xco, dyo, dzo, vel, E);
        hipLaunchKernelGGL(k_restart_transfer, tgr, kPointBlock, 0, c->stream, E, nxo, nyo, nzo, t, G, c->nxm, c->ny, c->nz,
                           c->ldx, (long long)c->ny * c->ldx);
        if ((rc = check_launch(c, "k_restart_transfer"))) return rc;
        if (comp == 0) {
            hipLaunchKernelGGL(k_restart_integrate, dim3(unsigned((c->ny * c->nz + 255) / 256)), dim3(256), 0, c->stream,
                               vel_out[0], G, 0, c->nxm, c->ny, c->nz, c->ldx, (long long)c->ny * c->ldx, lxa, 0.0);
        } else {
            const int nm = comp == 1 ? c->nz : c->ny;
            hipLaunchKernelGGL(k_restart_plane, dim3(unsigned((c->nxm + 63) / 64), unsigned((nm + 3) / 4)), kPointBlock, 0,
                               c->stream, F, comp, nxo, nyo, nzo, t, vel, vel_out[comp], c->nxm, nm, c->ldx, (long long)c->ny * c->ldx);
            const int lines = comp == 1 ? c->nxm * c->nz : c->nxm * c->ny;
            hipLaunchKernelGGL(k_restart_integrate, dim3(unsigned((lines + 255) / 256)), dim3(256), 0, c->stream, vel_out[comp],
                               G, comp, c->nxm, c->ny, c->nz, c->ldx, (long long)c->ny * c->ldx, lxa,
                               comp == 1 ? c->c.ylen / double(c->ny) : c->c.zlen / double(c->nz));
        }
        if ((rc = check_launch(c, "k_restart_integrate"))) return rc;
    }
    // T with its wall ghosts (InterpInputVel.F90:280-330)
    double tup, tlo;
    if (c->c.rayt > 0) {
        if (c->c.inslwn == 0) { tup = 0.0; tlo = 1.0; } else { tup = -0.5; tlo = 0.5; }
    } else { tup = 0.5; tlo = -0.5; }
    if (c->pf) {
        if (c->c.rayt > 0) { tup = 0.0; tlo = 1.0; } else { tup = 1.0; tlo = 0.0; }
    }
    RestartGhost sg{1, 0.0, 0.0, c->c.tfixs, c->c.tfixn, tlo, tup};
    HIPCHK(c, hipMemcpyAsync(F, tempo, sizeof(double) * nold, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_restart_stage, sgr, dim3(256), 0, c->stream, F, 3, nxo, nyo, nzo, xco, dyo, dzo, sg, E);
    hipLaunchKernelGGL(k_restart_transfer, tgr, kPointBlock, 0, c->stream, E, nxo, nyo, nzo, t, c->f[AFID_TEMP], c->nxm, c->ny,
                       c->nz, c->ldx, (long long)c->ny * c->ldx);
    if ((rc = check_launch(c, "k_restart_transfer"))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AFID_OK;
}

}  // extern "C"

// InterpInputSal.F90 (phi = false) and InterpInputPhi.F90 (phi = true)
static int interp_input_refined(afid_ctx *c, const afid_old_grid *og, const double *fo, bool phi) {
    if (!c || !og || !fo) return AFID_E_ARG;
    if (phi ? !c->pf : !c->sal) return set_err(c, AFID_E_STATE, phi ? "the phase field is not enabled" : "salinity is not enabled");
    int rc;
    if ((rc = check_old_grid(c, og, true))) return rc;
    const int nxo = og->nxro, nyo = og->nyro - 1, nzo = og->nzro - 1, nxmo = nxo - 1;
    RestartScratch sc(c);
    RestartTabs t{};
    const double *xco;
    if ((rc = restart_tables(c, sc, nxmo, nyo, nzo, og->istr3ro, og->str3o, true, t, &xco))) return rc;
    const size_t nold = size_t(nxo) * nyo * nzo, next = size_t(nxo + 3) * nyo * nzo;
    const size_t nr = size_t(c->nzr) * c->nyr * c->ldxr;
    const int fid = phi ? AFID_PHI : AFID_SAL, hid = phi ? AFID_RUPHI : AFID_RUSAL;
    double *F, *E, *out = c->f[fid];
    if ((rc = sc.alloc(nold, &F)) || (rc = sc.alloc(next, &E))) return rc;
    if (!out && (rc = sc.alloc(nr, &out))) return rc;         // FP32 storage: staged in FP64
    HIPCHK(c, hipMemsetAsync(out, 0, sizeof(double) * nr, c->stream));
    if (c->g[hid]) HIPCHK(c, hipMemsetAsync(c->g[hid], 0, sizeof(float) * nr, c->stream));
    else if (c->f[hid]) HIPCHK(c, hipMemsetAsync(c->f[hid], 0, sizeof(double) * nr, c->stream));
    // S: Dirichlet ghosts from the sign of Ra_S, 1 / 0 with the phase field
    // (InterpInputSal.F90:33-56); phi: zero gradient (InterpInputPhi.F90:28-33)
    double sup = c->c.rays > 0 ? 0.5 : -0.5, slo = c->c.rays > 0 ? -0.5 : 0.5;
    if (c->pf) { sup = 0.0; slo = 1.0; }
    RestartGhost gh{1, 0.0, 0.0, phi ? 0 : c->c.sfixs, phi ? 0 : c->c.sfixn, slo, sup};
    HIPCHK(c, hipMemcpyAsync(F, fo, sizeof(double) * nold, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_restart_stage, dim3(unsigned((nxo + 3 + 255) / 256), unsigned(nyo), unsigned(nzo)), dim3(256), 0, c->stream,
                       F, 3, nxo, nyo, nzo, xco, 0.0, 0.0, gh, E);
    hipLaunchKernelGGL(k_restart_transfer, point_grid(c->nxmr, c->nyr, c->nzr), kPointBlock, 0, c->stream, E, nxo, nyo, nzo, t, out,
                       c->nxmr, c->nyr, c->nzr, c->ldxr, (long long)c->nyr * c->ldxr);
    if ((rc = check_launch(c, "k_restart_transfer"))) return rc;
    if (!c->f[fid]) {
        hipLaunchKernelGGL(k_restart_to_float, dim3(unsigned((nr + 255) / 256)), dim3(256), 0, c->stream, out, c->g[fid], (long long)nr);
        if ((rc = check_launch(c, "k_restart_to_float"))) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AFID_OK;
}

extern "C" {

int afid_interp_input_sal(afid_ctx *c, const afid_old_grid *og, const double *salo) { return interp_input_refined(c, og, salo, false); }
int afid_interp_input_phi(afid_ctx *c, const afid_old_grid *og, const double *phio) { return interp_input_refined(c, og, phio, true); }

// InterpInputHum, moisture.F90:539-582: the old humidity onto the base grid
// with Dirichlet (qfixS / qfixN) or zero-gradient ghosts, the wall values
// being those of the first column of humbp / humtp (the reference assumes
// them uniform, :557-559; SetHumidityBCs runs first); the RK history restarts
// from zero
int afid_interp_input_humid(afid_ctx *c, const afid_old_grid *og, const double *qo) {
    if (!c || !og || !qo) return AFID_E_ARG;
    NEED_MOIST(c)
    int rc;
    if ((rc = check_old_grid(c, og, false))) return rc;
    const int nxo = og->nxo, nyo = og->nyo - 1, nzo = og->nzo - 1, nxmo = nxo - 1;
    RestartScratch sc(c);
    RestartTabs t{};
    const double *xco;
    if ((rc = restart_tables(c, sc, nxmo, nyo, nzo, og->istr3o, og->str3o, false, t, &xco))) return rc;
    const size_t nold = size_t(nxo) * nyo * nzo, next = size_t(nxo + 3) * nyo * nzo;
    const size_t nb = size_t(c->nz) * c->ny * c->ldx;
    double *F, *E;
    if ((rc = sc.alloc(nold, &F)) || (rc = sc.alloc(next, &E))) return rc;
    double walls[2] = {0.0, 0.0};
    HIPCHK(c, hipMemcpyAsync(&walls[0], c->f[AFID_HUMBP], sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&walls[1], c->f[AFID_HUMTP], sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemsetAsync(c->f[AFID_HUMID], 0, sizeof(double) * nb, c->stream));
    HIPCHK(c, hipMemsetAsync(c->f[AFID_RUHUMID], 0, sizeof(double) * nb, c->stream));
    RestartGhost gh{1, 0.0, 0.0, c->q_fixs, c->q_fixn, walls[0], walls[1]};
    HIPCHK(c, hipMemcpyAsync(F, qo, sizeof(double) * nold, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_restart_stage, dim3(unsigned((nxo + 3 + 255) / 256), unsigned(nyo), unsigned(nzo)), dim3(256), 0, c->stream,
                       F, 3, nxo, nyo, nzo, xco, 0.0, 0.0, gh, E);
    hipLaunchKernelGGL(k_restart_transfer, point_grid(c->nxm, c->ny, c->nz), kPointBlock, 0, c->stream, E, nxo, nyo, nzo, t,
                       c->f[AFID_HUMID], c->nxm, c->ny, c->nz, c->ldx, (long long)c->ny * c->ldx);
    if ((rc = check_launch(c, "k_restart_transfer"))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AFID_OK;
}

int afid_synchronize(afid_ctx *c) {
    if (!c) return AFID_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AFID_OK;
}

static int read_max(afid_ctx *c, double *a, double *b) {
    unsigned long long h[4];
    HIPCHK(c, hipMemcpyAsync(h, c->redmax, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double v[4];
    std::memcpy(v, h, sizeof(v));
    *a = v[0]; *b = v[1];
    return AFID_OK;
}

int afid_calc_max_cfl(afid_ctx *c, double *cflm, double *cflmr) {
    if (!c || !cflm || !cflmr) return AFID_E_ARG;
    HIPCHK(c, hipMemsetAsync(c->redmax, 0, sizeof(unsigned long long) * 4, c->stream));
    KP p = make_kp(c);
    hipLaunchKernelGGL(k_cfl_div, red_grid(c->nxm, c->ny, c->nz), kPointBlock, 0, c->stream, p, 0);
    int rc = check_launch(c, "k_cfl_div");
    if (rc) return rc;
    if (c->sal) {
        if (c->mixed) hipLaunchKernelGGL(k_cfl_div_r<float>, red_grid(c->nxmr, c->nyr, c->nzr), kPointBlock, 0, c->stream, p, 0);
        else hipLaunchKernelGGL(k_cfl_div_r<double>, red_grid(c->nxmr, c->nyr, c->nzr), kPointBlock, 0, c->stream, p, 0);
        if ((rc = check_launch(c, "k_cfl_div_r"))) return rc;
    }
    if ((rc = read_max(c, cflm, cflmr))) return rc;
    *cflm = std::max(*cflm, 0.00000001);          // CalcMaxCFL.F90:24
    *cflmr = c->sal ? std::max(*cflmr, 1.0e-8) : *cflm;
    return AFID_OK;
}

int afid_check_divergence(afid_ctx *c, double *qmax, double *qmaxr) {
    if (!c || !qmax || !qmaxr) return AFID_E_ARG;
    HIPCHK(c, hipMemsetAsync(c->redmax, 0, sizeof(unsigned long long) * 4, c->stream));
    KP p = make_kp(c);
    hipLaunchKernelGGL(k_cfl_div, red_grid(c->nxm, c->ny, c->nz), kPointBlock, 0, c->stream, p, 1);
    int rc = check_launch(c, "k_cfl_div");
    if (rc) return rc;
    if (c->sal) {
        if (c->mixed) hipLaunchKernelGGL(k_cfl_div_r<float>, red_grid(c->nxmr, c->nyr, c->nzr), kPointBlock, 0, c->stream, p, 1);
        else hipLaunchKernelGGL(k_cfl_div_r<double>, red_grid(c->nxmr, c->nyr, c->nzr), kPointBlock, 0, c->stream, p, 1);
        if ((rc = check_launch(c, "k_cfl_div_r"))) return rc;
    }
    unsigned long long h[4];
    HIPCHK(c, hipMemcpyAsync(h, c->redmax, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double v[4];
    std::memcpy(v, h, sizeof(v));
    *qmax = v[2];
    *qmaxr = c->sal ? v[3] : v[2];
    return AFID_OK;
}

int afid_global_quantities(afid_ctx *c, double out[4]) {
    if (!c || !out) return AFID_E_ARG;
    KP p = make_kp(c);
    hipLaunchKernelGGL(k_global_sums, point_grid(c->nxm, c->ny, c->nz), kPointBlock, 0, c->stream, p);
    int rc = check_launch(c, "k_global_sums");
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, c->stream, c->redsum, c->nred, c->redout);
    if ((rc = check_launch(c, "k_sum_partials"))) return rc;
    double s[4];
    HIPCHK(c, hipMemcpyAsync(s, c->redout, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double vol = 1.0 / (c->c.alx3 * c->dx * double(c->nz) * double(c->ny));   // GlobalQuantities.F90:35
    out[0] = 1.0 + std::sqrt(c->c.prat * std::fabs(c->c.rayt)) * s[0] * vol;       // :79
    out[1] = 0.5 * s[2] * vol;
    out[2] = std::sqrt(s[2] * vol) * std::sqrt(std::fabs(c->c.rayt) / c->c.prat);  // :91
    out[3] = s[1] * vol;                                                            // :89
    return AFID_OK;
}

int afid_calc_mean_profiles(afid_ctx *c, double *out) {
    if (!c || !out) return AFID_E_ARG;
    if (c->nxm < 2) return set_err(c, AFID_E_ARG, "CalcMeanProfiles needs nxm >= 2");
    int rc;
    const int nxm = c->nxm;
    if (!c->d_mean_part) {
        c->mean_groups = mean_ygroups(c->ny);
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->d_mean_part), sizeof(double) * size_t(c->mean_groups) * MEAN_NQ * nxm))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->d_mean_out), sizeof(double) * size_t(MEAN_NQ) * nxm))) return rc;
    }
    KP p = make_kp(c);
    hipLaunchKernelGGL(k_mean_partials, dim3(unsigned((nxm + 63) / 64), unsigned(c->mean_groups)), dim3(256), 0, c->stream, p,
                       c->d_sdx, c->d_mean_part);
    if ((rc = check_launch(c, "k_mean_partials"))) return rc;
    hipLaunchKernelGGL(k_mean_finish, dim3(unsigned((MEAN_NQ * nxm + 255) / 256)), dim3(256), 0, c->stream, p, c->d_mean_part,
                       c->mean_groups, c->d_mean_out);
    if ((rc = check_launch(c, "k_mean_finish"))) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->d_mean_out, sizeof(double) * size_t(MEAN_NQ) * nxm, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AFID_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// CalcMeanProfiles without allocation: the partial sums live in the
// file-scope device array `scratch`, whose address on each device is taken
// once at context set-up (scratch_ptr_for_device) and used from then on.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[AFID_NOALLOC_SCRATCH_DOUBLES];

namespace {
std::mutex g_scratch_mu;
std::map<int, double *> g_scratch_ptr;     // device ordinal -> address of `scratch` there
}  // namespace

// Address of `scratch` on the current device (`dev`), looked up once per
// device.  Called from context set-up with that device current.
static int scratch_ptr_for_device(afid_ctx *c, int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr.count(dev)) return AFID_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return set_err(c, AFID_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
    return AFID_OK;
}

static double *scratch_ptr(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    return it == g_scratch_ptr.end() ? nullptr : it->second;
}

extern "C" {

int afid_calc_mean_profiles_noalloc(afid_ctx *c, double *out_dev) {
    if (!c || !out_dev) return AFID_E_ARG;
    if (c->nxm < 2) return set_err(c, AFID_E_ARG, "CalcMeanProfiles needs nxm >= 2");
    const int nxm = c->nxm, groups = mean_ygroups(c->ny);
    const long long need = (long long)groups * MEAN_NQ * nxm;
    if (need > AFID_NOALLOC_SCRATCH_DOUBLES)
        return set_err(c, AFID_E_ARG, "afid_calc_mean_profiles_noalloc: the partial sums need " + std::to_string(need) +
                                          " doubles (15 x nxm x ceil(ny/4)); the static scratch holds " +
                                          std::to_string((long long)AFID_NOALLOC_SCRATCH_DOUBLES));
    double *part = scratch_ptr(c->device);
    if (!part) return set_err(c, AFID_E_STATE, "afid_calc_mean_profiles_noalloc: no scratch address for this device");
    int rc;
    KP p = make_kp(c);
    // every (group, profile, x) partial is written before k_mean_finish reads
    // it, so the scratch needs no clearing
    hipLaunchKernelGGL(k_mean_partials, dim3(unsigned((nxm + 63) / 64), unsigned(groups)), dim3(256), 0, c->stream, p,
                       c->d_sdx, part);
    if ((rc = check_launch(c, "k_mean_partials"))) return rc;
    hipLaunchKernelGGL(k_mean_finish, dim3(unsigned((MEAN_NQ * nxm + 255) / 256)), dim3(256), 0, c->stream, p, part, groups,
                       out_dev);
    return check_launch(c, "k_mean_finish");
}

int afid_calc_moist_stats(afid_ctx *c, double *out) {
    if (!c || !out) return AFID_E_ARG;
    NEED_MOIST(c)
    int rc;
    const int nxm = c->nxm, groups = mean_ygroups(c->ny);
    if (!c->d_moist_part) {
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->d_moist_part), sizeof(double) * size_t(groups) * MOIST_NQ * nxm))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->d_moist_out), sizeof(double) * size_t(MOIST_NQ) * nxm))) return rc;
    }
    KP p = make_kp(c);
    hipLaunchKernelGGL(k_moist_partials, dim3(unsigned((nxm + 63) / 64), unsigned(groups)), dim3(256), 0, c->stream, p, c->d_moist_part);
    if ((rc = check_launch(c, "k_moist_partials"))) return rc;
    hipLaunchKernelGGL(k_moist_finish, dim3(unsigned((MOIST_NQ * nxm + 255) / 256)), dim3(256), 0, c->stream, p, c->d_moist_part,
                       groups, c->d_moist_out);
    if ((rc = check_launch(c, "k_moist_finish"))) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->d_moist_out, sizeof(double) * size_t(MOIST_NQ) * nxm, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AFID_OK;
}

int afid_calc_sal_stats(afid_ctx *c, double *out) {
    if (!c || !out) return AFID_E_ARG;
    if (!c->sal) return set_err(c, AFID_E_ARG, "CalcSalStats needs the salinity field");
    if (c->nxmr < 2) return set_err(c, AFID_E_ARG, "CalcSalStats needs nxmr >= 2");
    int rc;
    const int n = c->nxmr;
    if (!c->d_sal_part) {
        c->sal_groups = mean_ygroups(c->nyr);
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->d_sal_part), sizeof(double) * size_t(c->sal_groups) * SAL_NQ * n))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->d_sal_out), sizeof(double) * size_t(SAL_NQ) * n))) return rc;
    }
    KP p = make_kp(c);
    const dim3 gr(unsigned((n + 63) / 64), unsigned(mean_ygroups(c->nyr)));
    if (c->mixed) hipLaunchKernelGGL(k_sal_stats_partials<float>, gr, dim3(256), 0, c->stream, p, c->d_sdxr, c->d_sal_part);
    else hipLaunchKernelGGL(k_sal_stats_partials<double>, gr, dim3(256), 0, c->stream, p, c->d_sdxr, c->d_sal_part);
    if ((rc = check_launch(c, "k_sal_stats_partials"))) return rc;
    hipLaunchKernelGGL(k_sal_stats_finish, dim3(unsigned((SAL_NQ * n + 255) / 256)), dim3(256), 0, c->stream, p, c->d_sal_part,
                       c->sal_groups, c->d_sal_out);
    if ((rc = check_launch(c, "k_sal_stats_finish"))) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->d_sal_out, sizeof(double) * size_t(SAL_NQ) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AFID_OK;
}

int afid_calc_phi_stats(afid_ctx *c, double *out) {
    if (!c || !out) return AFID_E_ARG;
    if (!c->pf) return set_err(c, AFID_E_ARG, "CalcPhiStats needs the phase field");
    int rc;
    const int n = c->nxmr;
    if (!c->d_phi_part) {
        c->phi_groups = mean_ygroups(c->nyr);
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->d_phi_part), sizeof(double) * size_t(c->phi_groups) * 2 * n))) return rc;
        if ((rc = dalloc(c, reinterpret_cast<void **>(&c->d_phi_out), sizeof(double) * 2 * size_t(n)))) return rc;
    }
    KP p = make_kp(c);
    const dim3 gr(unsigned((n + 63) / 64), unsigned(mean_ygroups(c->nyr)));
    if (c->mixed) hipLaunchKernelGGL(k_phi_stats_partials<float>, gr, dim3(256), 0, c->stream, p, c->d_phi_part);
    else hipLaunchKernelGGL(k_phi_stats_partials<double>, gr, dim3(256), 0, c->stream, p, c->d_phi_part);
    if ((rc = check_launch(c, "k_phi_stats_partials"))) return rc;
    hipLaunchKernelGGL(k_phi_stats_finish, dim3(unsigned((2 * n + 255) / 256)), dim3(256), 0, c->stream, p, c->d_phi_part,
                       c->phi_groups, c->d_phi_out);
    if ((rc = check_launch(c, "k_phi_stats_finish"))) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->d_phi_out, sizeof(double) * 2 * size_t(n), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return AFID_OK;
}

int afid_refinement_residuals(afid_ctx *c, double *out, int max, int *count) {
    if (!c || !count) return AFID_E_ARG;
    unsigned long long h[4] = {0, 0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(h, c->slots, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *count = c->mixed ? c->max_sweeps - 1 : 0;
    for (int q = 0; q < *count && q < max && out; ++q) std::memcpy(&out[q], &h[q + 1], sizeof(double));
    return AFID_OK;
}

int afid_set_timing(afid_ctx *c, int on) {
    if (!c) return AFID_E_ARG;
    c->timing = on != 0;
    return AFID_OK;
}

int afid_get_timing(afid_ctx *c, const char **names, double *ms, int max, int *count) {
    if (!c || !count) return AFID_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<std::pair<std::string, double>> acc;
    for (auto &t : c->timers) {
        float v = 0.f;
        if (hipEventElapsedTime(&v, t.a, t.b) != hipSuccess) continue;
        bool found = false;
        for (auto &a : acc) if (a.first == t.name) { a.second += v; found = true; }
        if (!found) acc.push_back({t.name, double(v)});
    }
    c->last_timing = acc;
    *count = int(acc.size());
    for (int q = 0; q < int(acc.size()) && q < max; ++q) {
        if (names) names[q] = c->last_timing[size_t(q)].first.c_str();
        if (ms) ms[q] = acc[size_t(q)].second;
    }
    return AFID_OK;
}

int afid_grid_array(afid_ctx *c, const char *name, double *out, int64_t n, int64_t *count) {
    if (!c || !name || !count) return AFID_E_ARG;
    auto it = c->host1d.find(name);
    if (it == c->host1d.end()) return set_err(c, AFID_E_ARG, std::string("unknown grid array ") + name);
    *count = int64_t(it->second.size());
    if (out) std::memcpy(out, it->second.data(), sizeof(double) * size_t(std::min<int64_t>(n, *count)));
    return AFID_OK;
}

}  // extern "C"

// the batched 2-D ensemble (include/afid_ensemble.h)
#include "afid_ensemble_api.hpp"
