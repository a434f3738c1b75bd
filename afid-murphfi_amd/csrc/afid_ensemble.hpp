// afid_ensemble.hpp -- an ensemble of independent 2-D runs (ny = 1, x-z
// plane) advanced together: one launch per RK3 substep for every run.
//
// The reference advances one simulation per process (main.F90:281-422 around
// TimeMarcher.F90:13-165).  Here a workgroup owns whole runs: it walks its
// share of a static largest-first schedule and, for each run, does the full
// substep inside the workgroup, with workgroup barriers between the phases
// (runs never share data, so no grid-wide synchronisation exists):
//
//   factors   LU factors of the three wall-normal matrices for this run's dt
//             (three wavefronts, Moebius prefix scan), into LDS;
//   columns   explicit terms + implicit RHS + solve + update of vx, vz, T,
//             one wavefront per column (the base column kernel's stencils;
//             with ny = 1 every y difference is an exact zero and v stays 0);
//   forward   divergence of two x-adjacent rows packed into one complex
//             line, FFT along z in LDS, split into the two real spectra;
//   modes     per-mode x-tridiagonal with the run's precomputed factors;
//   inverse   the two spectra packed back, inverse FFT, into dph;
//   project   RemoveDivergence + CorrectPressure;
//   driver    (last substep) CFL and max|div| of the new state, time, the
//             next dt and the run status (main.F90:285-302, 355-385).
//
// Per-run dt, time and status live in device memory, so a captured graph of
// substep launches replays with no host round trip.
#pragma once
#include <hip/hip_runtime.h>
#include "afid_params.hpp"
#include "afid_stencils.hpp"
#include "afid_kernels.hpp"
#include "afid_poisson.hpp"

namespace afid {

// Read-only description of one run: its arrays and metrics (the KP of its
// 2-D context, buffer set 0 in vx/vz/temp), both buffers of the fields the
// column phase ping-pongs, the Poisson factors, spectrum and z transform,
// the RK coefficients and the driver settings of its bou.in.
struct EnsRun {
    KP p;
    double *vxb[2], *vzb[2], *tb[2];
    PoissonFactors<double> pf;
    cpx<double> *spec;
    FftPlan<double> plan;
    // mixed form: FP32 spectrum, factors and transform; the residual of each
    // Poisson sweep (max|div u| bit patterns, slots 1..3) gates the next
    PoissonFactors<float> pff;
    cpx<float> *specf;
    FftPlan<float> planf;
    unsigned long long *rslot;
    double div_tol;
    int max_sweeps;
    double gam[3], rom[3], alm[3];
    double dtmin, dtmax, limit_cfl, tmax, resid;
    int variable_dt, ntst, nsst, pad;
};

// Mutable state of one run (device), updated by the last substep of a step.
// status: 0 running; otherwise the reference's exit code (165 CFL too high,
// 166 dt below DTMIN, 169 divergence above RESID, 333 time above TMAX, 555
// NTST steps done); a run with nonzero status is left untouched.
struct EnsState {
    double dt, time, cfl, dmax;
    int ntime, status, cur, steps;
};

// order the LDS accesses of one wavefront (cross-lane exchange through LDS)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LU factors of one normalised constant tridiagonal matrix in the form
// wave_tri_solve consumes, built by one wavefront into l/e/g/s (64*CH rows).
// Rows [r0, r1) carry a_r = -am_r bd sb, c_r = -ap_r bd sb, s_r = sb =
// 1/(1 - ac_r bd) (SolveImpEqnUpdate_X.F90:25-42, _YZ.F90:25-44,
// _Temp.F90:25-46); the others are identity rows with zero right-hand side.
// u_r = 1 - a_r c'_{r-1} comes from the Moebius prefix of c'.
template <int CH>
__device__ __forceinline__ void wave_build_factors(const double *am, const double *ac, const double *ap, double bd,
                                                   int r0, int r1, double *l, double *e, double *g, double *s) {
    const int lane = threadIdx.x & 63;
    double a[CH], c[CH], gg[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;
        if (r >= r0 && r < r1) {
            const double sb = 1.0 / (1.0 - ac[r] * bd);
            a[m] = -am[r] * bd * sb;
            c[m] = -ap[r] * bd * sb;
            s[r] = sb;
        } else {
            a[m] = 0.0; c[m] = 0.0; s[r] = 0.0;
        }
    }
    double t = wave_mobius_prefix<CH>(a, c);
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;
        gg[m] = 1.0 / (1.0 - a[m] * t);
        t = c[m] * gg[m];
        e[r] = t;
        g[r] = gg[m];
    }
    double gprev = __shfl_up(gg[CH - 1], 1, 64);
    if (lane == 0) gprev = 0.0;
#pragma unroll
    for (int m = 0; m < CH; ++m) l[lane * CH + m] = a[m] * (m == 0 ? gprev : gg[m - 1]);
}

// The reference's time-step control at the top of step s.ntime
// (main.F90:285-302) from the CFL number of the state the step starts from.
__device__ __forceinline__ void ens_top_of_step(const EnsRun &R, EnsState &s, double cfl) {
    if (s.status) return;
    if (R.variable_dt) {
        if (s.ntime > 1) {
            s.dt = cfl < 1.0e-8 ? R.dtmax : R.limit_cfl / cfl;
            if (s.dt > R.dtmax) s.dt = R.dtmax;
        }
        if (s.dt < R.dtmin) s.status = 166;
    } else if (cfl * s.dt > R.limit_cfl) {
        s.status = 165;
    }
}

// Workgroup max of two values (LDS scratch of 2*NT/64 doubles)
template <int NT>
__device__ __forceinline__ void block_max2(double &a, double &b, double *red) {
    for (int off = 32; off > 0; off >>= 1) {
        a = fmax(a, __shfl_down(a, off, 64));
        b = fmax(b, __shfl_down(b, off, 64));
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) { red[wave] = a; red[NT / 64 + wave] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / 64; ++w) { a = fmax(a, red[w]); b = fmax(b, red[NT / 64 + w]); }
    }
}

// CalcMaxCFL (CalcMaxCFL.F90:23-44) and CheckDivergence (CheckDivergence.F90:
// 29-41) of one 2-D run, by the whole workgroup; the results are valid in
// thread 0.  NaN propagates into dmax (fmax keeps it only if both are NaN, so
// a NaN cell is turned into +inf first).
template <int NT>
__device__ __forceinline__ void ens_cfl_div(const KP &p, double &cfl, double &dmax, double *red) {
    cfl = 1.0e-8;
    dmax = 0.0;
    const int nxm = p.nxm, nz = p.nz, lane = threadIdx.x & 63;
    // one z column per wavefront, x across the lanes (a max is exact in any order)
    for (int i = threadIdx.x >> 6; i < nz; i += NT / 64) {
        const int ip = wrapp(i, nz);
        for (int k = lane; k < nxm; k += 64) {
            const long long c = (long long)i * p.plane + k, cp = (long long)ip * p.plane + k;
            const double q = fabs((p.vz[c] + p.vz[cp]) * 0.5 * p.dz) + fabs((p.vx[c] + p.vx[c + 1]) * 0.5 * p.udx3m[k]);
            const double d = fabs((p.vz[cp] - p.vz[c]) * p.dz + (p.vx[c + 1] - p.vx[c]) * p.udx3m[k]);
            cfl = fmax(cfl, q != q ? HUGE_VAL : q);
            dmax = fmax(dmax, d != d ? HUGE_VAL : d);
        }
    }
    block_max2<NT>(cfl, dmax, red);
}

// LDS of the substep kernel (bytes): the column phase (three factor sets and
// two column buffers per wavefront) or the transform lines, whichever is
// larger, plus the reduction scratch.
template <int NT, int CH>
__host__ __device__ constexpr int ens_col_lds() {
    return int(sizeof(double)) * (12 * 64 * CH + (NT / 64) * 2 * lds_col_size<CH>());
}
// complex values per thread in a transform stage of the ensemble kernel
constexpr int ENS_FFT_MAXV = 4;
__host__ __device__ inline int ens_fft_lines(int n, int cap, int nt) {
    int L = (ENS_FFT_MAXV * nt) / n;
    if (L > cap) L = cap;
    return L < 1 ? 1 : L;
}

// ---------------------------------------------------------------------------
// phases of one run's substep, shared by the fused kernel (one workgroup per
// run, all phases) and the phased kernels (one launch per phase over every
// run's columns, lines, modes or cells)
// ---------------------------------------------------------------------------
// the run's KP for substep ns with dt, reading buffer set cur
__device__ __forceinline__ KP ens_kp(const EnsRun &R, int ns, double dt, int cur) {
    KP p = R.p;
    p.dt = dt;
    p.al = R.alm[ns - 1]; p.ga = R.gam[ns - 1]; p.ro = R.rom[ns - 1];
    p.beta = dt / p.ren * 0.5;
    p.vx = R.vxb[cur]; p.vz = R.vzb[cur]; p.temp = R.tb[cur];
    return p;
}

// LU factors of matrix which (0 vx, 1 vz, 2 T) for p's dt and substep, by
// one wavefront, into f (4 arrays of 64*CH rows)
template <int CH>
__device__ __forceinline__ void ens_factors(const KP &p, int which, double *f) {
    constexpr int NP = 64 * CH;
    if (which == 0)        // identity rows 1 and nx (SolveImpEqnUpdate_X.F90:22-54)
        wave_build_factors<CH>(p.am3ck, p.ac3ck, p.ap3ck, p.beta * p.al, 1, p.nxm, f, f + NP, f + 2 * NP, f + 3 * NP);
    else if (which == 1)
        wave_build_factors<CH>(p.am3sk, p.ac3sk, p.ap3sk, p.beta * p.al, 0, p.nxm, f, f + NP, f + 2 * NP, f + 3 * NP);
    else
        wave_build_factors<CH>(p.am3ssk, p.ac3ssk, p.ap3ssk, 0.5 * p.al * p.dt / p.pect, 0, p.nxm, f, f + NP,
                               f + 2 * NP, f + 3 * NP);
}

// Explicit terms of a 2-D run (ny = 1, v = 0): the stencils of
// ExplicitTermsVX.F90:42-127, ExplicitTermsVZ.F90:45-114 and
// ExplicitTermsTemp.F90:39-77 with every y term dropped.  With one periodic
// y cell each y difference of the 3-D stencils is an exact zero (and v is
// identically zero), so the arithmetic that remains is the same.
// Column (i) pointers: c = i, m = i - 1, p = i + 1 (periodic).
struct Nb2 {
    const double *vx, *vx_m, *vx_p, *vz, *vz_m, *vz_p, *t, *t_m, *t_p, *pr, *pr_m;
    double tbp, ttp;
};
__device__ __forceinline__ Nb2 make_nb2(const KP &p, int i) {
    const long long c = (long long)i * p.plane, m = (long long)wrapm(i, p.nz) * p.plane, q = (long long)wrapp(i, p.nz) * p.plane;
    Nb2 n;
    n.vx = p.vx + c; n.vx_m = p.vx + m; n.vx_p = p.vx + q;
    n.vz = p.vz + c; n.vz_m = p.vz + m; n.vz_p = p.vz + q;
    n.t = p.temp + c; n.t_m = p.temp + m; n.t_p = p.temp + q;
    n.pr = p.pr + c; n.pr_m = p.pr + m;
    n.tbp = p.tempbp[i];
    n.ttp = p.temptp[i];
    return n;
}
// (arithmetic in T: double in the FP64 form, float in the mixed form, the
// FP64 state rounded on load)
template <typename T>
__device__ __forceinline__ T ex_vx_2d(const KP &p, const Nb2 &n, unsigned k) {
    const T udz = T(p.dz) * T(0.25), udzq = T(p.dzq) / T(p.ren);
    const T rkm = T(p.rudx3m[k - 1]), rkc = T(p.rudx3m[k]), u3 = T(p.udx3c[k]);
    const T vxc = T(n.vx[k]), vxzp = T(n.vx_p[k]), vxzm = T(n.vx_m[k]);
    const T hxz = u3 * udz * ((rkm * T(n.vz_p[k - 1]) + rkc * T(n.vz_p[k])) * (vxzp + vxc)
                              - (rkm * T(n.vz[k - 1]) + rkc * T(n.vz[k])) * (vxc + vxzm));
    const T a = T(n.vx[k + 1]) + vxc, b = vxc + T(n.vx[k - 1]);
    const T hxx = (a * a - b * b) * u3 * T(0.25);
    const T dzzvx = (vxzm - T(2.0) * vxc + vxzp) * udzq;
    T q = -(hxx + hxz) + dzzvx;
    if (p.gaxis == 1) q += T(p.byct) * (T(p.active_t) * T(0.5) * (T(n.t[k]) + T(n.t[k - 1])));
    return q;
}
template <typename T>
__device__ __forceinline__ T ex_vz_2d(const KP &p, const Nb2 &n, unsigned k) {
    const T udzq = T(p.dzq) / T(p.ren), udz = T(p.dz) * T(0.25);
    const unsigned kmm = k == 0 ? 0 : k - 1, kpp = k == unsigned(p.nxm - 1) ? k : k + 1;
    const T vzc = T(n.vz[k]), vzzp = T(n.vz_p[k]), vzzm = T(n.vz_m[k]);
    const T a = vzzp + vzc, b = vzzm + vzc;
    const T hzz = (a * a - b * b) * udz;
    const T hzx = ((T(n.vx[k + 1]) + T(n.vx_m[k + 1])) * (T(n.vz[kpp]) + vzc)
                   - (T(n.vx[k]) + T(n.vx_m[k])) * (vzc + T(n.vz[kmm]))) * T(p.udx3m[k]) * T(0.25);
    const T dzzvz = (vzzp - T(2.0) * vzc + vzzm) * udzq;
    T q = -(hzx + hzz) + dzzvz;
    if (p.gaxis == 3) q += T(p.byct) * (T(p.active_t) * T(0.5) * (T(n.t[k]) + T(n.t_m[k])));
    return q;
}
template <typename T>
__device__ __forceinline__ T ex_temp_2d(const KP &p, const Nb2 &n, unsigned k) {
    const T udz = T(p.dz) * T(0.5), udzq = T(p.dzq) / T(p.pect);
    const T tc = T(n.t[k]), tzp = T(n.t_p[k]), tzm = T(n.t_m[k]);
    const T vxp = T(n.vx[k + 1]), vxc = T(n.vx[k]);
    T htx;
    if (k == 0) htx = (vxp * (T(n.t[k + 1]) + tc) - vxc * T(2.0) * T(n.tbp)) * T(0.5) * T(p.udx3m[k]);
    else if (k == unsigned(p.nxm - 1)) htx = (vxp * T(2.0) * T(n.ttp) - vxc * (tc + T(n.t[k - 1]))) * T(0.5) * T(p.udx3m[k]);
    else htx = (vxp * (T(n.t[k + 1]) + tc) - vxc * (tc + T(n.t[k - 1]))) * T(0.5) * T(p.udx3m[k]);
    const T htz = (T(n.vz_p[k]) * (tzp + tc) - T(n.vz[k]) * (tc + tzm)) * udz;
    const T dzzt = (tzp - T(2.0) * tc + tzm) * udzq;
    return -(htx + htz) + dzzt;
}

// Right-hand sides of vx, vz and T at row r of a 2-D column in one pass
// (ImplicitAndUpdateVX / VZ / Temp with the 2-D explicit terms above; rows
// without an equation give 0); the new explicit terms go to the histories.
template <typename HT>
__device__ __forceinline__ void ens_rhs3(const KP &p, const Nb2 &n, long long c0, unsigned r, double &rx, double &rz,
                                         double &rt) {
    const unsigned nxm = p.nxm;
    rx = 0.0; rz = 0.0; rt = 0.0;
    if (r >= nxm) return;
    HT *rux = sel<HT>(p.rux, p.ruxf) + c0, *ruz = sel<HT>(p.ruz, p.ruzf) + c0, *rut = sel<HT>(p.rut, p.rutf) + c0;
    if (r >= 1) {
        const HT h = ex_vx_2d<HT>(p, n, r);
        const double dxx = n.vx[r + 1] * p.ap3ck[r] + n.vx[r] * p.ac3ck[r] + n.vx[r - 1] * p.am3ck[r];
        const double dxp = (n.pr[r] - n.pr[r - 1]) * (p.al * p.udx3c[r]);
        rx = (p.ga * double(h) + p.ro * double(rux[r]) + (p.al / p.ren) * dxx - dxp) * p.dt;
        rux[r] = h;
    }
    {
        const HT h = ex_vz_2d<HT>(p, n, r);
        const unsigned kmm = r == 0 ? 0 : r - 1, kpp = r == nxm - 1 ? r : r + 1;
        const double dxx = n.vz[kpp] * p.ap3sk[r] + n.vz[r] * p.ac3sk[r] + n.vz[kmm] * p.am3sk[r];
        const double dzp = (n.pr[r] - n.pr_m[r]) * p.dz * p.al;
        rz = (p.ga * double(h) + p.ro * double(ruz[r]) + (p.al / p.ren) * dxx - dzp) * p.dt;
        ruz[r] = h;
    }
    {
        const HT h = ex_temp_2d<HT>(p, n, r);
        double dxx;
        if (r == 0)
            dxx = n.t[r + 1] * p.ap3ssk[r] + n.t[r] * p.ac3ssk[r] - (p.ap3ssk[r] + p.ac3ssk[r]) * n.tbp * p.tfixs;
        else if (r == nxm - 1)
            dxx = n.t[r] * p.ac3ssk[r] + n.t[r - 1] * p.am3ssk[r] - (p.am3ssk[r] + p.ac3ssk[r]) * n.ttp * p.tfixn;
        else
            dxx = n.t[r + 1] * p.ap3ssk[r] + n.t[r] * p.ac3ssk[r] + n.t[r - 1] * p.am3ssk[r];
        rt = (p.ga * double(h) + p.ro * double(rut[r]) + (p.al / p.pect) * dxx) * p.dt;
        rut[r] = h;
    }
}

// column i of the run by one wavefront, new values into buffer set cur ^ 1.
// One sweep over the rows forms the three right-hand sides (each neighbour
// value loaded once): vx's and vz's into this wavefront's two LDS columns
// b0 / b1, T's in registers; the vx and vz solves and updates follow, then
// T's in b0.  The updates read the old values again (from cache).  fac holds
// the three factor sets.  Each field's arithmetic is that of a pass over one
// field at a time, so the two orders give the same results bit for bit.
template <int CH, typename HT = double>
__device__ __forceinline__ void ens_column(const EnsRun &R, const KP &p, int cur, int i, const double *fac, double *b0,
                                           double *b1, unsigned lane) {
    constexpr int NP = 64 * CH;
    const int nx = p.nx, nxm = p.nxm;
    const Nb2 n = make_nb2(p, i);
    const long long c0 = (long long)i * p.plane;
    double keep[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const unsigned r = lane + 64u * m;
        double rx, rz, rt;
        ens_rhs3<HT>(p, n, c0, r, rx, rz, rt);
        b0[lds_row<CH>(r)] = rx;
        b1[lds_row<CH>(r)] = rz;
        keep[m] = rt;
    }
    wave_sync();
    auto factors_of = [&](int fi) {
        const double *f = fac + fi * 4 * NP;
        TriFactors tf;
        tf.l = f; tf.e = f + NP; tf.g = f + 2 * NP; tf.s = f + 3 * NP; tf.a = nullptr; tf.c = nullptr; tf.ln = nullptr; tf.sn = nullptr;
        tf.n = fi == 0 ? nx : nxm; tf.chunk = CH;
        return tf;
    };
    column_solve_lds<CH>(b0, factors_of(0));
    column_solve_lds<CH>(b1, factors_of(1));
    wave_sync();
    {
        double *ox = R.vxb[cur ^ 1] + c0, *oz = R.vzb[cur ^ 1] + c0;
        for (unsigned rr = lane; rr < unsigned(nx); rr += 64) {
            const unsigned qq = lds_row<CH>(rr);
            ox[rr] = (rr >= 1 && rr < unsigned(nxm)) ? n.vx[rr] + b0[qq] : 0.0;   // walls, ImplicitAndUpdateVX.F90:86-87
            oz[rr] = rr < unsigned(nxm) ? n.vz[rr] + b1[qq] : n.vz[rr];
        }
    }
    wave_sync();
#pragma unroll
    for (int m = 0; m < CH; ++m) b0[lds_row<CH>(lane + 64u * m)] = keep[m];
    wave_sync();
    column_solve_lds<CH>(b0, factors_of(2));
    wave_sync();
    {
        double *ot = R.tb[cur ^ 1] + c0;
        for (unsigned rr = lane; rr < unsigned(nx); rr += 64)
            ot[rr] = rr < unsigned(nxm) ? n.t[rr] + b0[lds_row<CH>(rr)] : n.t[rr];
    }
    wave_sync();
}

// the run's spectrum, transform and mode factors in arithmetic R
template <typename R> struct EnsPoisson;
template <> struct EnsPoisson<double> {
    static __device__ __forceinline__ cpx<double> *spec(const EnsRun &E) { return E.spec; }
    static __device__ __forceinline__ const FftPlan<double> &plan(const EnsRun &E) { return E.plan; }
    static __device__ __forceinline__ const PoissonFactors<double> &pf(const EnsRun &E) { return E.pf; }
};
template <> struct EnsPoisson<float> {
    static __device__ __forceinline__ cpx<float> *spec(const EnsRun &E) { return E.specf; }
    static __device__ __forceinline__ const FftPlan<float> &plan(const EnsRun &E) { return E.planf; }
    static __device__ __forceinline__ const PoissonFactors<float> &pf(const EnsRun &E) { return E.pff; }
};

// divergence (pressure.F90:147-161, formed in FP64) of the line pairs
// [l0, l0 + nl) and their forward transform along z (pressure.F90:171-179,
// normalised by 1/nz since ny = 1) in arithmetic R into the half spectrum;
// p reads the current velocity.  Whole workgroup of NT threads; lb holds nl
// lines.
template <int NT, typename R = double>
__device__ __forceinline__ void ens_forward(const EnsRun &E, const KP &p, int l0, int nl, cpx<R> *lb) {
    const int nxm = p.nxm, N = p.nz, S = line_stride<R>(N), nh = N / 2 + 1, TX = 2 * nl, k0 = 2 * l0;
    const double usdtal = 1.0 / (p.dt * p.al);
    const R scale = R(1.0 / double(N));
    cpx<R> *spec = EnsPoisson<R>::spec(E);
    const FastDiv dtx(TX);                  // t < 2 nl N: t * TX < 2^32
    for (int t = threadIdx.x; t < TX * N; t += NT) {
        const int i = dtx.div(t), xo = t - i * TX, k = k0 + xo;
        double d = 0.0;
        if (k < nxm) {
            const long long c = (long long)i * p.plane + k, cp = (long long)wrapp(i, N) * p.plane + k;
            d = ((p.vz[cp] - p.vz[c]) * p.dz + (p.vx[c + 1] - p.vx[c]) * p.udx3m[k]) * usdtal;
        }
        R *slot = reinterpret_cast<R *>(&lb[(xo >> 1) * S + lpad<R>(i)]);
        slot[xo & 1] = R(d);
    }
    __syncthreads();
    lds_fft<R, NT, ENS_FFT_MAXV>(lb, nl, S, EnsPoisson<R>::plan(E), -1);
    for (int t = threadIdx.x; t < TX * nh; t += NT) {
        const int m = dtx.div(t), xo = t - m * TX, k = k0 + xo;
        if (k >= nxm) continue;
        const int line = xo >> 1;
        const cpx<R> zm = lb[line * S + lpad<R>(m)], zn = lb[line * S + lpad<R>(m == 0 ? 0 : N - m)];
        cpx<R> o;
        if ((xo & 1) == 0) o = mk<R>(R(0.5) * (zm.x + zn.x), R(0.5) * (zm.y - zn.y));
        else o = mk<R>(R(0.5) * (zm.y + zn.y), R(-0.5) * (zm.x - zn.x));
        spec[(long long)m * nxm + k] = mk<R>(o.x * scale, o.y * scale);
    }
    __syncthreads();
}

// per-mode tridiagonal (pressure.F90:201-236) of mode m by one wavefront;
// modes m and N-m share the matrix and the real data make them conjugate
template <int CH, typename R = double>
__device__ __forceinline__ void ens_mode(const EnsRun &E, int nxm, int m) {
    const PoissonFactors<R> &pf = EnsPoisson<R>::pf(E);
    mode_tri_solve<CH, R>(pf.m, pf.g, pf.e, (long long)m * pf.npad, pf.npad, EnsPoisson<R>::spec(E) + (long long)m * nxm, nxm);
}

// inverse transform (pressure.F90:189-195) of the line pairs [l0, l0 + nl)
// into dph (FP64).  Whole workgroup.
template <int NT, typename R = double>
__device__ __forceinline__ void ens_inverse(const EnsRun &E, const KP &p, int l0, int nl, cpx<R> *lb) {
    const int nxm = p.nxm, N = p.nz, S = line_stride<R>(N), nh = N / 2 + 1, TX = 2 * nl, k0 = 2 * l0;
    const cpx<R> *spec = EnsPoisson<R>::spec(E);
    const FastDiv dnl(nl), dtx(TX);
    for (int t = threadIdx.x; t < nl * nh; t += NT) {
        const int m = dnl.div(t), line = t - m * nl;
        const int ka = k0 + 2 * line, kb = ka + 1;
        const cpx<R> *row = spec + (long long)m * nxm;
        cpx<R> A = ka < nxm ? row[ka] : mk<R>(R(0), R(0));
        cpx<R> B = kb < nxm ? row[kb] : mk<R>(R(0), R(0));
        if (m == 0 || 2 * m == N) { A.y = R(0); B.y = R(0); }
        lb[line * S + lpad<R>(m)] = mk<R>(A.x - B.y, A.y + B.x);
        if (m > 0 && N - m > m) lb[line * S + lpad<R>(N - m)] = mk<R>(A.x + B.y, B.x - A.y);
    }
    __syncthreads();
    lds_fft<R, NT, ENS_FFT_MAXV>(lb, nl, S, EnsPoisson<R>::plan(E), +1);
    for (int t = threadIdx.x; t < TX * N; t += NT) {
        const int i = dtx.div(t), xo = t - i * TX, k = k0 + xo;
        if (k >= nxm) continue;
        const cpx<R> z = lb[(xo >> 1) * S + lpad<R>(i)];
        sel<R>(p.dph, p.dphf)[(long long)i * p.plane + k] = (xo & 1) ? z.y : z.x;     // dph in arithmetic R
    }
    __syncthreads();
}

// RemoveDivergence + CorrectPressure (pressure.F90:245-300) of the cells
// t = k + nxm i in [t0, t1), t0 + threadIdx.x stepping by NT
template <int NT, typename DT = double>
__device__ __forceinline__ void ens_project(const KP &p, int t0, int t1) {
    const int nxm = p.nxm, nz = p.nz;
    const double adt = p.al * p.dt, be = p.al * p.beta;
    const DT *d = sel<DT>(p.dph, p.dphf);         // FP32 in the mixed form
    const FastDiv dn(nxm);                  // t < nxm nz: t * nxm < 2^32 for every supported run
    for (int t = t0 + threadIdx.x; t < t1; t += NT) {
        const int i = dn.div(t), k = t - i * nxm, im = wrapm(i, nz), ip = wrapp(i, nz);
        const int km = k == 0 ? 0 : k - 1, kp = k == nxm - 1 ? k : k + 1;
        const long long c = (long long)i * p.plane + k, row = c - k;
        const double dc = d[c], dkm = d[row + km], dkp = d[row + kp];
        const double dim = d[(long long)im * p.plane + k], dip = d[(long long)ip * p.plane + k];
        const double ux = p.vx[c], uz = p.vz[c], prc = p.pr[c];
        const double nux = ux - (dc - dkm) * (adt * p.udx3c[k]);
        const double nuz = uz - (dc - dim) * (adt * p.dz);
        const double npr = prc + (dc - be * ((dip - 2.0 * dc + dim) * p.dzq +
                                            (dkp * p.apphk[k] + dc * p.acphk[k] + dkm * p.amphk[k])));
        p.vx[c] = nux; p.vz[c] = nuz; p.pr[c] = npr;
    }
}

// the driver's update after step s.ntime (thread 0): time, exit codes
// 169 / 333 / 555 (main.F90:355-385) and the next step's dt
__device__ __forceinline__ void ens_end_of_step(const EnsRun &R, EnsState &s, double cfl, double dmax) {
    s.time += s.dt;
    s.cfl = cfl; s.dmax = dmax;
    if (!(dmax <= R.resid)) s.status = 169;
    else if (s.time > R.tmax) s.status = 333;
    else if (s.ntime == R.ntst) s.status = 555;
    s.ntime += 1;
    s.steps += 1;
    ens_top_of_step(R, s, cfl);
}

// ---------------------------------------------------------------------------
// fused form: one RK3 substep ns (1-based) of every scheduled run in one
// launch; workgroup w owns runs sched_run[sched_off[w] .. sched_off[w+1]).
// Lines: transform lines per LDS batch (the tuned tile size).
// ---------------------------------------------------------------------------
template <int NT, int CH>
__global__ void __launch_bounds__(NT) k_ens_substep(const EnsRun *__restrict__ runs, EnsState *st,
                                                    const int *__restrict__ sched_off, const int *__restrict__ sched_run,
                                                    int ns, int lines) {
    extern __shared__ __align__(16) unsigned char smraw[];
    constexpr int NW = NT / 64, NP = 64 * CH;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned lane = threadIdx.x & 63;
    double *fac = reinterpret_cast<double *>(smraw);
    double *cols = fac + 12 * NP;
    cpx<double> *lb = reinterpret_cast<cpx<double> *>(smraw);
    double *red = reinterpret_cast<double *>(smraw);       // reused after the last phase
    const int q0 = sched_off[blockIdx.x], q1 = sched_off[blockIdx.x + 1];
    for (int q = q0; q < q1; ++q) {
        const int r = sched_run[q];
        const EnsRun &R = runs[r];
        if (__builtin_amdgcn_readfirstlane(st[r].status)) continue;     // uniform over the workgroup
        const int cur = __builtin_amdgcn_readfirstlane(st[r].cur);
        KP p = ens_kp(R, ns, st[r].dt, cur);
        if (wave < 3) ens_factors<CH>(p, wave, fac + wave * 4 * NP);
        __syncthreads();
        {
            double *buf = cols + wave * 2 * lds_col_size<CH>(), *own = buf + lds_col_size<CH>();
            for (int i = wave; i < p.nz; i += NW) ens_column<CH>(R, p, cur, i, fac, buf, own, lane);
        }
        __syncthreads();
        p.vx = R.vxb[cur ^ 1]; p.vz = R.vzb[cur ^ 1]; p.temp = R.tb[cur ^ 1];
        const int nlines = (p.nxm + 1) / 2, L = ens_fft_lines(p.nz, lines, NT), nh = p.nz / 2 + 1;
        for (int l0 = 0; l0 < nlines; l0 += L) ens_forward<NT>(R, p, l0, min(L, nlines - l0), lb);
        for (int m = wave; m < nh; m += NW) ens_mode<CH>(R, p.nxm, m);
        __syncthreads();
        for (int l0 = 0; l0 < nlines; l0 += L) ens_inverse<NT>(R, p, l0, min(L, nlines - l0), lb);
        ens_project<NT>(p, 0, p.nxm * p.nz);
        __syncthreads();
        if (ns == R.nsst) {
            double cfl, dmax;
            ens_cfl_div<NT>(p, cfl, dmax, red);
            if (threadIdx.x == 0) {
                EnsState s = st[r];
                ens_end_of_step(R, s, cfl, dmax);
                s.cur = cur ^ 1;
                st[r] = s;
            }
        } else if (threadIdx.x == 0) {
            st[r].cur = cur ^ 1;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// phased form: the same substep as a sequence of launches, each over the
// work items of every running run (a run's items share nothing with other
// runs; the launch boundaries order the phases):
//   factors (one workgroup per run) -> columns (4 per workgroup) ->
//   forward (line batches) -> modes (4 per workgroup) -> inverse ->
//   project (cell chunks) -> finish (one workgroup per run: buffer parity,
//   and after the last substep the driver's step update).
// An item is (run, first index); frozen runs' items return at once.
// ---------------------------------------------------------------------------
struct EnsItem { int run, start; };
constexpr int ENSP_NT = 256;                      // threads of the phased kernels
constexpr int ENSP_CELLS = 4 * ENSP_NT;           // cells per projection item

// Mixed form: Poisson sweep s (1-based) of a run runs when s = 1, or while
// the FP64 residual max|div u| left by sweep s-1 (rslot[s-1]) is above
// div_tol, up to max_sweeps; the FP64 form has one sweep.
// any[s] != 0: some run's residual after sweep s exceeded its div_tol (set by
// k_ensp_resid, cleared by k_ensp_factors); a launch of sweep s >= 2 whose
// flag is clear exits before reading its work item
__device__ __forceinline__ bool ens_any_open(const unsigned long long *any, int s) {
    return s <= 1 || __builtin_amdgcn_readfirstlane(int(any[s - 1] != 0ull));
}
__device__ __forceinline__ bool ens_sweep_open(const EnsRun &E, int s) {
    if (s <= 1) return true;
    if (s > E.max_sweeps) return false;
    return __longlong_as_double((long long)E.rslot[s - 1]) > E.div_tol;
}

template <int CH>
__global__ void __launch_bounds__(192) k_ensp_factors(const EnsRun *__restrict__ runs, const EnsState *st, int ns,
                                                      double *fac, unsigned long long *any) {
    const int r = blockIdx.x;
    if (r == 0 && threadIdx.x < 4) any[threadIdx.x] = 0ull;         // sweep flags of this substep
    if (__builtin_amdgcn_readfirstlane(st[r].status)) return;
    const EnsRun &R = runs[r];
    const KP p = ens_kp(R, ns, st[r].dt, 0);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (R.rslot && threadIdx.x < 4) R.rslot[threadIdx.x] = 0ull;      // residual slots of this substep
    ens_factors<CH>(p, wave, fac + ((long long)r * 3 + wave) * 4 * 64 * CH);
}

template <int CH, typename HT>
__global__ void __launch_bounds__(ENSP_NT) k_ensp_columns(const EnsRun *__restrict__ runs, const EnsState *st,
                                                          const EnsItem *__restrict__ items, int ns, const double *fac) {
    __shared__ double lds[ENSP_NT / 64][2][lds_col_size<CH>()];
    const EnsItem it = items[blockIdx.x];
    const int r = it.run;
    if (__builtin_amdgcn_readfirstlane(st[r].status)) return;
    const EnsRun &R = runs[r];
    const int cur = __builtin_amdgcn_readfirstlane(st[r].cur);
    const KP p = ens_kp(R, ns, st[r].dt, cur);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = it.start + wave;
    if (i >= p.nz) return;
    ens_column<CH, HT>(R, p, cur, i, fac + (long long)r * 12 * 64 * CH, lds[wave][0], lds[wave][1], threadIdx.x & 63);
}

// forward (inverse = 0) or inverse (inverse = 1) transform of one line batch
// of Poisson sweep `sweep`, in arithmetic R
template <typename R>
__global__ void __launch_bounds__(ENSP_NT) k_ensp_transform(const EnsRun *__restrict__ runs, const EnsState *st,
                                                            const EnsItem *__restrict__ items, int ns, int lines,
                                                            int inverse, int sweep, const unsigned long long *any) {
    if (!ens_any_open(any, sweep)) return;
    extern __shared__ __align__(16) unsigned char smraw[];
    cpx<R> *lb = reinterpret_cast<cpx<R> *>(smraw);
    const EnsItem it = items[blockIdx.x];
    const int r = it.run;
    if (__builtin_amdgcn_readfirstlane(st[r].status)) return;
    const EnsRun &R_ = runs[r];
    if (!ens_sweep_open(R_, sweep)) return;
    const int cur = __builtin_amdgcn_readfirstlane(st[r].cur);
    const KP p = ens_kp(R_, ns, st[r].dt, cur ^ 1);          // the columns' new velocity
    const int nlines = (p.nxm + 1) / 2, L = ens_fft_lines(p.nz, lines, ENSP_NT);
    const int nl = min(L, nlines - it.start);
    if (inverse) ens_inverse<ENSP_NT, R>(R_, p, it.start, nl, lb);
    else ens_forward<ENSP_NT, R>(R_, p, it.start, nl, lb);
}

template <int CH, typename R>
__global__ void __launch_bounds__(ENSP_NT) k_ensp_modes(const EnsRun *__restrict__ runs, const EnsState *st,
                                                        const EnsItem *__restrict__ items, int sweep,
                                                        const unsigned long long *any) {
    if (!ens_any_open(any, sweep)) return;
    const EnsItem it = items[blockIdx.x];
    const int r = it.run;
    if (__builtin_amdgcn_readfirstlane(st[r].status)) return;
    const EnsRun &E = runs[r];
    if (!ens_sweep_open(E, sweep)) return;
    const int m = it.start + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (m <= E.p.nz / 2) ens_mode<CH, R>(E, E.p.nxm, m);
}

template <typename RT>
__global__ void __launch_bounds__(ENSP_NT) k_ensp_project(const EnsRun *__restrict__ runs, const EnsState *st,
                                                          const EnsItem *__restrict__ items, int ns, int sweep,
                                                          const unsigned long long *any) {
    if (!ens_any_open(any, sweep)) return;
    const EnsItem it = items[blockIdx.x];
    const int r = it.run;
    if (__builtin_amdgcn_readfirstlane(st[r].status)) return;
    const EnsRun &R = runs[r];
    if (!ens_sweep_open(R, sweep)) return;
    const int cur = __builtin_amdgcn_readfirstlane(st[r].cur);
    const KP p = ens_kp(R, ns, st[r].dt, cur ^ 1);
    ens_project<ENSP_NT, RT>(p, it.start, min(it.start + ENSP_CELLS, p.nxm * p.nz));
}

// Mixed form: FP64 residual max|div u| (CheckDivergence units) after the
// projection of sweep `sweep`, max-reduced per run into rslot[sweep]
__global__ void __launch_bounds__(ENSP_NT) k_ensp_resid(const EnsRun *__restrict__ runs, const EnsState *st,
                                                        const EnsItem *__restrict__ items, int ns, int sweep,
                                                        unsigned long long *any) {
    if (!ens_any_open(any, sweep)) return;
    const EnsItem it = items[blockIdx.x];
    const int r = it.run;
    if (__builtin_amdgcn_readfirstlane(st[r].status)) return;
    const EnsRun &R = runs[r];
    if (!ens_sweep_open(R, sweep)) return;
    const int cur = __builtin_amdgcn_readfirstlane(st[r].cur);
    const KP p = ens_kp(R, ns, st[r].dt, cur ^ 1);
    const int nxm = p.nxm, nz = p.nz, t1 = min(it.start + ENSP_CELLS, nxm * nz);
    double dmax = 0.0;
    const FastDiv dn(nxm);
    for (int t = it.start + threadIdx.x; t < t1; t += ENSP_NT) {
        const int i = dn.div(t), k = t - i * nxm;
        const long long c = (long long)i * p.plane + k, cp = (long long)wrapp(i, nz) * p.plane + k;
        const double d = fabs((p.vz[cp] - p.vz[c]) * p.dz + (p.vx[c + 1] - p.vx[c]) * p.udx3m[k]);
        dmax = fmax(dmax, d != d ? HUGE_VAL : d);
    }
    dmax = block_max(dmax);
    if (threadIdx.x == 0) {
        atomicMax(R.rslot + sweep, (unsigned long long)__double_as_longlong(dmax));
        // the next sweep opens for this run; one atomic on the shared flag
        // suffices, so it is skipped once the flag reads set
        if (!(dmax <= R.div_tol) && __hip_atomic_load(any + sweep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull)
            atomicOr(any + sweep, 1ull);
    }
}

// one workgroup of ENSP_FIN_NT threads per run (the CFL / divergence sweep
// of the last substep reads the whole run)
constexpr int ENSP_FIN_NT = 1024;
__global__ void __launch_bounds__(ENSP_FIN_NT) k_ensp_finish(const EnsRun *__restrict__ runs, EnsState *st, int ns) {
    __shared__ double red[2 * ENSP_FIN_NT / 64];
    const int r = blockIdx.x;
    if (__builtin_amdgcn_readfirstlane(st[r].status)) return;
    const EnsRun &R = runs[r];
    const int cur = __builtin_amdgcn_readfirstlane(st[r].cur);
    if (ns == R.nsst) {
        const KP p = ens_kp(R, ns, st[r].dt, cur ^ 1);
        double cfl, dmax;
        ens_cfl_div<ENSP_FIN_NT>(p, cfl, dmax, red);
        if (threadIdx.x == 0) {
            EnsState s = st[r];
            ens_end_of_step(R, s, cfl, dmax);
            s.cur = cur ^ 1;
            st[r] = s;
        }
    } else if (threadIdx.x == 0) {
        st[r].cur = cur ^ 1;
    }
}

// CFL and divergence of the current state of every scheduled run; with
// top != 0 the time-step control of the step about to start is applied
// (after new initial conditions or uploads).
template <int NT>
__global__ void __launch_bounds__(NT) k_ens_prepare(const EnsRun *__restrict__ runs, EnsState *st,
                                                    const int *__restrict__ sched_off, const int *__restrict__ sched_run,
                                                    int top) {
    __shared__ double red[2 * NT / 64];
    const int q0 = sched_off[blockIdx.x], q1 = sched_off[blockIdx.x + 1];
    for (int q = q0; q < q1; ++q) {
        const int r = sched_run[q];
        const EnsRun &R = runs[r];
        EnsState s = st[r];
        KP p = R.p;
        p.vx = R.vxb[s.cur]; p.vz = R.vzb[s.cur];
        double cfl, dmax;
        ens_cfl_div<NT>(p, cfl, dmax, red);
        if (threadIdx.x == 0) {
            s.cfl = cfl; s.dmax = dmax;
            if (top) ens_top_of_step(R, s, cfl);
            st[r] = s;
        }
        __syncthreads();
    }
}

}  // namespace afid
