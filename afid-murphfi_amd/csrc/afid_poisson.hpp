// afid_poisson.hpp -- pressure Poisson solve of SolvePressureCorrection
// (pressure.F90:165-236): real FFT along y, complex FFT along z, one
// wall-normal tridiagonal per Fourier mode, inverse transforms.
//
// Transforms run along the strided y and z directions with the contiguous x
// direction mapped to lanes, so every global access is a run of x-adjacent
// values; a workgroup holds L lines of length N in LDS and runs a
// mixed-radix (8,4,3,2) Stockham transform on them.  The real y transform
// packs two x-adjacent real lines into one complex line.
//
// Spectral layout: S[i][m][k] (complex, k = x fastest), m = 0..ny/2,
// i = 0..nz-1.  The forward transforms fold in the 1/(ny nz) normalisation
// (pressure.F90:179) and the 1/(al dt) of CalcLocalDivergence.
#pragma once
#include <hip/hip_runtime.h>
#include "afid_params.hpp"
#include "afid_stencils.hpp"
#include "afid_kernels.hpp"

namespace afid {

// complex value of real type R (double in the FP64 form, float in the mixed form)
template <typename R>
struct __align__(2 * sizeof(R)) cpx { R x, y; };
template <typename R>
__device__ __forceinline__ cpx<R> mk(R a, R b) { cpx<R> z; z.x = a; z.y = b; return z; }

template <typename R>
struct FftPlan {
    int n = 1;
    int nst = 0;
    int rad[12] = {0};
    const cpx<R> *tw = nullptr;   // exp(-2 pi i t/n), t = 0..n-1
};

template <typename R>
__device__ __forceinline__ cpx<R> cmul(cpx<R> a, cpx<R> b) { return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <typename R>
__device__ __forceinline__ cpx<R> cadd(cpx<R> a, cpx<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <typename R>
__device__ __forceinline__ cpx<R> csub(cpx<R> a, cpx<R> b) { return mk<R>(a.x - b.x, a.y - b.y); }
// multiply by i*sign
template <typename R>
__device__ __forceinline__ cpx<R> cmulis(cpx<R> a, int sign) { return sign < 0 ? mk<R>(a.y, -a.x) : mk<R>(-a.y, a.x); }

template <int N, typename R>
struct DftSmall;
template <typename R>
struct DftSmall<2, R> {
    static __device__ __forceinline__ void run(cpx<R> (&v)[2], int) {
        cpx<R> a = v[0], b = v[1];
        v[0] = cadd(a, b); v[1] = csub(a, b);
    }
};
template <typename R>
struct DftSmall<3, R> {
    static __device__ __forceinline__ void run(cpx<R> (&v)[3], int sign) {
        const R s3 = R(0.86602540378443864676);   // sin(2 pi/3)
        cpx<R> t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        cpx<R> m = mk<R>(v[0].x - R(0.5) * t1.x, v[0].y - R(0.5) * t1.y);
        cpx<R> r = cmulis(mk<R>(s3 * t2.x, s3 * t2.y), sign);
        v[0] = cadd(v[0], t1);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    }
};
template <typename R>
struct DftSmall<4, R> {
    static __device__ __forceinline__ void run(cpx<R> (&v)[4], int sign) {
        cpx<R> a = cadd(v[0], v[2]), b = csub(v[0], v[2]), c = cadd(v[1], v[3]), d = cmulis(csub(v[1], v[3]), sign);
        v[0] = cadd(a, c); v[2] = csub(a, c);
        v[1] = cadd(b, d); v[3] = csub(b, d);
    }
};
template <typename R>
struct DftSmall<8, R> {
    static __device__ __forceinline__ void run(cpx<R> (&v)[8], int sign) {
        const R h = R(0.70710678118654752440);    // 1/sqrt(2)
        cpx<R> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        DftSmall<4, R>::run(e, sign);
        DftSmall<4, R>::run(o, sign);
        // w8^k = exp(sign 2 pi i k/8)
        cpx<R> w1 = mk<R>(h, R(sign) * h), w3 = mk<R>(-h, R(sign) * h);
        cpx<R> t0 = o[0], t1 = cmul(o[1], w1), t2 = cmulis(o[2], sign), t3 = cmul(o[3], w3);
        v[0] = cadd(e[0], t0); v[4] = csub(e[0], t0);
        v[1] = cadd(e[1], t1); v[5] = csub(e[1], t1);
        v[2] = cadd(e[2], t2); v[6] = csub(e[2], t2);
        v[3] = cadd(e[3], t3); v[7] = csub(e[3], t3);
    }
};

// Workgroup size of the transform kernels and the most complex values a
// thread holds in a stage (L*N <= FFT_MAXV*FFT_THREADS): 16 wavefronts whose
// butterflies of a stage stay in registers between its read and write phase.
constexpr int FFT_THREADS = 1024, FFT_MAXV = 8;

// Division by a workgroup-uniform d through one multiply: whenever x * d <=
// 2^32, floor(x / d) = umulhi(x, ceil(2^32 / d)) (the error term, below
// x / 2^32 <= 1/d, never reaches the next integer, since frac(x/d) <= 1 - 1/d).
// d = 1 is passed through.
struct FastDiv {
    unsigned d, m;
    __device__ __forceinline__ explicit FastDiv(unsigned d_) : d(d_), m(d_ > 1 ? 0xFFFFFFFFu / d_ + 1u : 0u) {}
    __device__ __forceinline__ unsigned div(unsigned x) const { return d == 1 ? x : __umulhi(x, m); }
};

// LDS layout of L lines of N complex values: element j of line l sits at
// l*S + lpad(j), with one pad element per 32 banks of data (16 FP32 or 8
// FP64 complex values) so the stride-R writes of a Stockham stage spread
// over the banks, and a line stride S = lpad(N)+1 that is not a multiple of
// the 32 banks, so x-adjacent lines (the global-load order) do too.
template <typename R>
__host__ __device__ __forceinline__ int lpad(int i) { return i + (i >> (sizeof(R) == 4 ? 4 : 3)); }
template <typename R>
__host__ __device__ __forceinline__ int line_stride(int n) { return lpad<R>(n) + 1; }

// One Stockham stage of radix RD over L lines of length N held in LDS, by
// the NT threads of the workgroup (L*N <= FFT_MAXV*NT).  Each thread keeps
// at most MAXB butterflies of the stage in registers between the read and
// the write phase, so the transform runs in place.
template <int RD, int MAXB, typename R, int NT>
__device__ __forceinline__ void fft_stage(cpx<R> *buf, int L, int N, int S, int Ns, const cpx<R> *tw, int sign) {
    const int nbl = N / RD, nb = L * nbl;
    const FastDiv dnbl(nbl), dns(Ns);
    cpx<R> v[MAXB][RD];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * NT;
        if (b < nb) {
            const int line = dnbl.div(b), jj = b - line * nbl;
#pragma unroll
            for (int r = 0; r < RD; ++r) v[q][r] = buf[line * S + lpad<R>(jj + r * nbl)];
        }
    }
    __syncthreads();
    const int tstride = N / (Ns * RD);
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * NT;
        if (b < nb) {
            const int line = dnbl.div(b), jj = b - line * nbl, jq = dns.div(jj), kk = jj - jq * Ns;
#pragma unroll
            for (int r = 1; r < RD; ++r) {
                cpx<R> w = tw[kk * r * tstride];
                if (sign > 0) w.y = -w.y;
                v[q][r] = cmul(v[q][r], w);
            }
            DftSmall<RD, R>::run(v[q], sign);
            const int dst = jq * Ns * RD + kk;
#pragma unroll
            for (int r = 0; r < RD; ++r) buf[line * S + lpad<R>(dst + r * Ns)] = v[q][r];
        }
    }
    __syncthreads();
}

// Full transform of L lines (L*N <= FFT_MAXV*NT), line stride S, by a
// workgroup of NT threads.
// MAXV: the most complex values a thread holds in a stage (L*N <= MAXV*NT).
template <typename R, int NT = FFT_THREADS, int MAXV = FFT_MAXV>
__device__ __forceinline__ void lds_fft(cpx<R> *buf, int L, int S, const FftPlan<R> &pl, int sign) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int RD = pl.rad[s];
        switch (RD) {
        case 8: fft_stage<8, (MAXV + 7) / 8, R, NT>(buf, L, pl.n, S, Ns, pl.tw, sign); break;
        case 4: fft_stage<4, (MAXV + 3) / 4, R, NT>(buf, L, pl.n, S, Ns, pl.tw, sign); break;
        case 3: fft_stage<3, (MAXV + 2) / 3, R, NT>(buf, L, pl.n, S, Ns, pl.tw, sign); break;
        default: fft_stage<2, (MAXV + 1) / 2, R, NT>(buf, L, pl.n, S, Ns, pl.tw, sign); break;
        }
        Ns *= RD;
    }
}

// Source of the forward transform
enum PoissonSrc { SRC_VELOCITY = 0, SRC_DPH = 1 };

// Divergence (CalcLocalDivergence, pressure.F90:147-161) fused with the real
// forward y transform (pressure.F90:171-174) and the 1/(ny nz) scaling.
// With slot_out, the maximum of |div u| (CheckDivergence units) is also
// reduced into *slot_out; with slots_clear, workgroup 0 zeroes 4 slots.
// Block: 2L x-adjacent real lines at one z.  grid (ceil(nxm/2L), nz).
template <typename R>
__global__ void __launch_bounds__(FFT_THREADS) k_div_fft_y(KP p, FftPlan<R> pl, cpx<R> *spec, int L, int src, Gate gate,
                                                           unsigned long long *slot_out, unsigned long long *slots_clear) {
    if (gate_closed(gate)) return;
    extern __shared__ __align__(16) unsigned char smraw[];
    cpx<R> *lb = reinterpret_cast<cpx<R> *>(smraw);
    const int N = p.ny, S = line_stride<R>(N), TX = 2 * L, i = blockIdx.y, k0 = blockIdx.x * TX;
    const int ip = wrapp(i, p.nz);
    const double usdtal = 1.0 / (p.dt * p.al);
    if (slots_clear && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x < 4) slots_clear[threadIdx.x] = 0ull;
    double dmax = 0.0;
    const FastDiv dtx(TX);
    for (int t = threadIdx.x; t < TX * N; t += FFT_THREADS) {
        const int j = dtx.div(t), xo = t - j * TX, k = k0 + xo;
        double d = 0.0;
        if (k < p.nxm) {
            if (src == SRC_DPH) d = p.dph[bidx(p, k, j, i)];
            else {
                const int jp = wrapp(j, N);
                const double dv = (p.vz[bidx(p, k, j, ip)] - p.vz[bidx(p, k, j, i)]) * p.dz
                                + (p.vy[bidx(p, k, jp, i)] - p.vy[bidx(p, k, j, i)]) * p.dy
                                + (p.vx[bidx(p, k + 1, j, i)] - p.vx[bidx(p, k, j, i)]) * p.udx3m[k];
                dmax = fmax(dmax, fabs(dv));
                d = dv * usdtal;
            }
        }
        R *slot = reinterpret_cast<R *>(&lb[(xo >> 1) * S + lpad<R>(j)]);
        slot[xo & 1] = R(d);
    }
    if (slot_out) {
        dmax = block_max(dmax);
        max_to_slot(slot_out, dmax);
    }
    __syncthreads();
    lds_fft<R>(lb, L, S, pl, -1);
    const R scale = R(1.0 / (double(p.ny) * double(p.nz)));
    for (int t = threadIdx.x; t < TX * p.nymh; t += FFT_THREADS) {
        const int m = dtx.div(t), xo = t - m * TX, k = k0 + xo;
        if (k >= p.nxm) continue;
        const int line = xo >> 1;
        const cpx<R> zm = lb[line * S + lpad<R>(m)], zn = lb[line * S + lpad<R>(m == 0 ? 0 : N - m)];
        cpx<R> out;
        if ((xo & 1) == 0) out = mk<R>(R(0.5) * (zm.x + zn.x), R(0.5) * (zm.y - zn.y));     // (Z_m + conj Z_{N-m})/2
        else out = mk<R>(R(0.5) * (zm.y + zn.y), R(-0.5) * (zm.x - zn.x));                  // (Z_m - conj Z_{N-m})/(2i)
        spec[((long long)i * p.nymh + m) * p.nxm + k] = mk<R>(out.x * scale, out.y * scale);
    }
}

// Complex transform along z (pressure.F90:176 forward, :189 backward), in
// place on the spectral array.  Block: L x-adjacent lines at one y-mode m.
// MAXV: complex values per thread and stage (L*N <= MAXV*FFT_THREADS); 16
// lets the FP32 form take 16 x-adjacent lines (128-byte rows) at N = 768.
template <typename R, int MAXV = FFT_MAXV>
__global__ void __launch_bounds__(FFT_THREADS) k_fft_z(KP p, FftPlan<R> pl, cpx<R> *spec, int L, int sign, Gate gate) {
    if (gate_closed(gate)) return;
    extern __shared__ __align__(16) unsigned char smraw[];
    cpx<R> *lb = reinterpret_cast<cpx<R> *>(smraw);
    const int N = p.nz, S = line_stride<R>(N), m = blockIdx.y, k0 = blockIdx.x * L;
    const long long sz = (long long)p.nymh * p.nxm;
    cpx<R> *g = spec + (long long)m * p.nxm;
    const FastDiv dl(L);
    for (int t = threadIdx.x; t < L * N; t += FFT_THREADS) {
        const int i = dl.div(t), xo = t - i * L, k = k0 + xo;
        lb[xo * S + lpad<R>(i)] = k < p.nxm ? g[i * sz + k] : mk<R>(R(0), R(0));
    }
    __syncthreads();
    lds_fft<R, FFT_THREADS, MAXV>(lb, L, S, pl, sign);
    for (int t = threadIdx.x; t < L * N; t += FFT_THREADS) {
        const int i = dl.div(t), xo = t - i * L, k = k0 + xo;
        if (k < p.nxm) g[i * sz + k] = lb[xo * S + lpad<R>(i)];
    }
}

// Inverse real y transform (pressure.F90:193-195): two x-adjacent spectra
// A, B (Hermitian, imaginary parts of the m=0 and m=N/2 terms ignored as a
// c2r transform does) become Z = A + iB, whose inverse is a + ib, written
// to dph.
template <typename R>
__global__ void __launch_bounds__(FFT_THREADS) k_ifft_y(KP p, FftPlan<R> pl, const cpx<R> *spec, int L, Gate gate) {
    if (gate_closed(gate)) return;
    extern __shared__ __align__(16) unsigned char smraw[];
    cpx<R> *lb = reinterpret_cast<cpx<R> *>(smraw);
    const int N = p.ny, S = line_stride<R>(N), TX = 2 * L, i = blockIdx.y, k0 = blockIdx.x * TX;
    const FastDiv dl(L), dtx(TX);
    for (int t = threadIdx.x; t < L * p.nymh; t += FFT_THREADS) {
        const int m = dl.div(t), line = t - m * L;
        const int ka = k0 + 2 * line, kb = ka + 1;
        const cpx<R> *row = spec + ((long long)i * p.nymh + m) * p.nxm;
        cpx<R> A = ka < p.nxm ? row[ka] : mk<R>(R(0), R(0));
        cpx<R> B = kb < p.nxm ? row[kb] : mk<R>(R(0), R(0));
        if (m == 0 || 2 * m == N) { A.y = R(0); B.y = R(0); }
        lb[line * S + lpad<R>(m)] = mk<R>(A.x - B.y, A.y + B.x);
        if (m > 0 && N - m > m) lb[line * S + lpad<R>(N - m)] = mk<R>(A.x + B.y, B.x - A.y);
    }
    __syncthreads();
    lds_fft<R>(lb, L, S, pl, +1);
    for (int t = threadIdx.x; t < TX * N; t += FFT_THREADS) {
        const int j = dtx.div(t), xo = t - j * TX, k = k0 + xo;
        if (k >= p.nxm) continue;
        const cpx<R> z = lb[(xo >> 1) * S + lpad<R>(j)];
        sel<R>(p.dph, p.dphf)[bidx(p, k, j, i)] = (xo & 1) ? z.y : z.x;      // dph in the sweep's arithmetic R
    }
}

// Per-mode wall-normal tridiagonal (SolveTridiagonalPressure,
// pressure.F90:201-236) with factors precomputed at set-up.  Modes i and
// nz-i share a matrix (ak1 is even), so one block of two wavefronts solves
// both against one copy of the factors.  grid (nymh, nz/2+1).
template <typename R>
struct PoissonFactors {
    const R *m, *g, *e;   // [mode'][row padded to 64*CH], mode' = m*(nz/2+1) + i'
    int npad;
};
// Solve of one mode's column col (nxm complex values) by one wavefront,
// CH rows per lane, with the factors at fm/fg/fe + fo (rows >= npad read as
// zero, so any CH with 64*CH >= nxm serves).
template <int CH, typename R>
__device__ __forceinline__ void mode_tri_solve(const R *fmp, const R *fgp, const R *fep, long long fo, int npad,
                                               cpx<R> *col, int nxm) {
    const int lane = threadIdx.x & 63;
    R fm[CH], fg[CH], fe[CH];
    cpx<R> v[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
        const int r = lane * CH + q;
        const bool in = r < npad;
        fm[q] = in ? fmp[fo + r] : R(0); fg[q] = in ? fgp[fo + r] : R(0); fe[q] = in ? fep[fo + r] : R(0);
        v[q] = r < nxm ? col[r] : mk<R>(R(0), R(0));
    }
    // forward y_r = d_r - m_r y_{r-1} (complex B, real A)
    R A = R(1);
    cpx<R> Bc = mk<R>(R(0), R(0));
#pragma unroll
    for (int q = 0; q < CH; ++q) { Bc = mk<R>(-fm[q] * Bc.x + v[q].x, -fm[q] * Bc.y + v[q].y); A = -fm[q] * A; }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        R Ap = __shfl_up(A, off, 64), Bx = __shfl_up(Bc.x, off, 64), By = __shfl_up(Bc.y, off, 64);
        if (lane >= off) { Bc = mk<R>(A * Bx + Bc.x, A * By + Bc.y); A = A * Ap; }
    }
    cpx<R> y = mk<R>(__shfl_up(Bc.x, 1, 64), __shfl_up(Bc.y, 1, 64));
    if (lane == 0) y = mk<R>(R(0), R(0));
#pragma unroll
    for (int q = 0; q < CH; ++q) { y = mk<R>(-fm[q] * y.x + v[q].x, -fm[q] * y.y + v[q].y); v[q] = y; }
    // backward x_r = g_r y_r - e_r x_{r+1}
    A = R(1); Bc = mk<R>(R(0), R(0));
#pragma unroll
    for (int q = CH - 1; q >= 0; --q) {
        Bc = mk<R>(-fe[q] * Bc.x + fg[q] * v[q].x, -fe[q] * Bc.y + fg[q] * v[q].y); A = -fe[q] * A;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        R Ap = __shfl_down(A, off, 64), Bx = __shfl_down(Bc.x, off, 64), By = __shfl_down(Bc.y, off, 64);
        if (lane + off < 64) { Bc = mk<R>(A * Bx + Bc.x, A * By + Bc.y); A = A * Ap; }
    }
    cpx<R> x = mk<R>(__shfl_down(Bc.x, 1, 64), __shfl_down(Bc.y, 1, 64));
    if (lane == 63) x = mk<R>(R(0), R(0));
#pragma unroll
    for (int q = CH - 1; q >= 0; --q) {
        x = mk<R>(-fe[q] * x.x + fg[q] * v[q].x, -fe[q] * x.y + fg[q] * v[q].y); v[q] = x;
    }
#pragma unroll
    for (int q = 0; q < CH; ++q) {
        const int r = lane * CH + q;
        if (r < nxm) col[r] = v[q];
    }
}

template <int CH, typename R>
__global__ void __launch_bounds__(128) k_poisson_tri(KP p, PoissonFactors<R> pf, cpx<R> *spec, Gate gate) {
    if (gate_closed(gate)) return;
    const int wave = threadIdx.x >> 6;
    const int m = blockIdx.x, ih = blockIdx.y, nzh = p.nz / 2 + 1;
    const int i = wave == 0 ? ih : p.nz - ih;
    if (wave == 1 && (ih == 0 || i == ih || i >= p.nz)) return;
    const long long fo = ((long long)m * nzh + ih) * pf.npad;
    mode_tri_solve<CH, R>(pf.m, pf.g, pf.e, fo, pf.npad, spec + ((long long)i * p.nymh + m) * p.nxm, p.nxm);
}

// Set-up of the per-mode factors: one thread per (m, i'), sequential over x.
// Normalised rows (pressure.F90:211-220): a_k = amphk s_k, b = 1+1e-15,
// c_k = apphk s_k, rhs s_k with s_k = 1/(acphk_k - ak2_m - ak1_i).  The
// singular (0,0) mode pins its first unknown to zero instead.
// Stored: m_k = l_k s_{k-1}/s_k, g_k = s_k/u_k, e_k = c_k/u_k so that
// y'_k = d_k - m_k y'_{k-1}, x_k = g_k y'_k - e_k x_{k+1}.
__global__ void k_poisson_factors(KP p, const double *ak1, const double *ak2, double *fm, double *fg, double *fe, int npad) {
    const int nzh = p.nz / 2 + 1;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)p.nymh * nzh) return;
    const int m = int(t / nzh), ih = int(t % nzh);
    const long long fo = t * npad;
    const double K = ak2[m] + ak1[ih];
    const bool gauge = (m == 0 && ih == 0);
    double u_prev = 1.0, c_prev = 0.0, s_prev = 0.0;
    for (int k = 0; k < npad; ++k) {
        if (k >= p.nxm) { fm[fo + k] = 0.0; fg[fo + k] = 0.0; fe[fo + k] = 0.0; continue; }
        double s = 1.0 / (p.acphk[k] - K);
        double a = k > 0 ? p.amphk[k] * s : 0.0;
        double c = k < p.nxm - 1 ? p.apphk[k] * s : 0.0;
        double b = 1.0 + 1.0e-15;
        if (gauge && k == 0) { s = 0.0; a = 0.0; c = 0.0; b = 1.0; }
        double l = k > 0 ? a / u_prev : 0.0;
        double u = b - l * c_prev;
        fm[fo + k] = (k > 0 && s != 0.0) ? l * s_prev / s : 0.0;
        fg[fo + k] = s / u;
        fe[fo + k] = c / u;
        u_prev = u; c_prev = c; s_prev = s;
    }
}

// FP64 -> FP32 copy (factor tables and weights of the mixed form)
__global__ void k_to_float(const double *src, float *dst, long long n) {
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
        dst[q] = float(src[q]);
}

}  // namespace afid
