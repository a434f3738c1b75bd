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

struct FftPlan {
    int n = 1;
    int nst = 0;
    int rad[12] = {0};
    const double2 *tw = nullptr;   // exp(-2 pi i t/n), t = 0..n-1
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by i*sign
__device__ __forceinline__ double2 cmulis(double2 a, int sign) {
    return sign < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

template <int R>
__device__ __forceinline__ void dft_small(double2 (&v)[R], int sign);

template <>
__device__ __forceinline__ void dft_small<2>(double2 (&v)[2], int) {
    double2 a = v[0], b = v[1];
    v[0] = cadd(a, b); v[1] = csub(a, b);
}
template <>
__device__ __forceinline__ void dft_small<3>(double2 (&v)[3], int sign) {
    const double s3 = 0.86602540378443864676;   // sin(2 pi/3)
    double2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    double2 m = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
    double2 r = cmulis(make_double2(s3 * t2.x, s3 * t2.y), sign);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}
template <>
__device__ __forceinline__ void dft_small<4>(double2 (&v)[4], int sign) {
    double2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]), c = cadd(v[1], v[3]), d = cmulis(csub(v[1], v[3]), sign);
    v[0] = cadd(a, c); v[2] = csub(a, c);
    v[1] = cadd(b, d); v[3] = csub(b, d);
}
template <>
__device__ __forceinline__ void dft_small<8>(double2 (&v)[8], int sign) {
    const double h = 0.70710678118654752440;    // 1/sqrt(2)
    double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft_small<4>(e, sign);
    dft_small<4>(o, sign);
    // w8^k = exp(sign 2 pi i k/8)
    double2 w1 = make_double2(h, sign * h), w3 = make_double2(-h, sign * h);
    double2 t0 = o[0], t1 = cmul(o[1], w1), t2 = cmulis(o[2], sign), t3 = cmul(o[3], w3);
    v[0] = cadd(e[0], t0); v[4] = csub(e[0], t0);
    v[1] = cadd(e[1], t1); v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2); v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3); v[7] = csub(e[3], t3);
}

// One Stockham stage of radix R over L lines of length N held in LDS.
// Each thread keeps at most MAXB butterflies of the stage in registers
// between the read and the write phase, so the transform runs in place.
template <int R, int MAXB>
__device__ __forceinline__ void fft_stage(double2 *buf, int L, int N, int Ns, const double2 *tw, int sign) {
    const int nbl = N / R, nb = L * nbl, T = blockDim.x;
    double2 v[MAXB][R];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * T;
        if (b < nb) {
            const int line = b / nbl, jj = b - line * nbl;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = buf[line * N + jj + r * nbl];
        }
    }
    __syncthreads();
    const int tstride = N / (Ns * R);
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * T;
        if (b < nb) {
            const int line = b / nbl, jj = b - line * nbl, kk = jj % Ns;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                double2 w = tw[kk * r * tstride];
                if (sign > 0) w.y = -w.y;
                v[q][r] = cmul(v[q][r], w);
            }
            dft_small<R>(v[q], sign);
            const int dst = (jj / Ns) * Ns * R + kk;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[line * N + dst + r * Ns] = v[q][r];
        }
    }
    __syncthreads();
}

// Full transform of L lines; at most 32 complex values per thread (L*N <= 32*blockDim).
__device__ __forceinline__ void lds_fft(double2 *buf, int L, const FftPlan &pl, int sign) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; ++s) {
        const int R = pl.rad[s];
        switch (R) {
        case 8: fft_stage<8, 4>(buf, L, pl.n, Ns, pl.tw, sign); break;
        case 4: fft_stage<4, 8>(buf, L, pl.n, Ns, pl.tw, sign); break;
        case 3: fft_stage<3, 11>(buf, L, pl.n, Ns, pl.tw, sign); break;
        default: fft_stage<2, 16>(buf, L, pl.n, Ns, pl.tw, sign); break;
        }
        Ns *= R;
    }
}

// Divergence (CalcLocalDivergence, pressure.F90:147-161) fused with the real
// forward y transform (pressure.F90:171-174) and the 1/(ny nz) scaling.
// Block: 2L x-adjacent real lines at one z.  grid (ceil(nxm/2L), nz).
__global__ void __launch_bounds__(256) k_div_fft_y(KP p, FftPlan pl, double2 *spec, int L, int from_dph) {
    extern __shared__ double2 lb[];
    const int N = p.ny, TX = 2 * L, i = blockIdx.y, k0 = blockIdx.x * TX;
    const int ip = wrapp(i, p.nz);
    const double usdtal = 1.0 / (p.dt * p.al);
    double *lr = reinterpret_cast<double *>(lb);
    for (int t = threadIdx.x; t < TX * N; t += blockDim.x) {
        const int xo = t % TX, j = t / TX, k = k0 + xo;
        double d = 0.0;
        if (k < p.nxm) {
            if (from_dph) d = p.dph[bidx(p, k, j, i)];
            else {
                const int jp = wrapp(j, N);
                d = ((p.vz[bidx(p, k, j, ip)] - p.vz[bidx(p, k, j, i)]) * p.dz
                   + (p.vy[bidx(p, k, jp, i)] - p.vy[bidx(p, k, j, i)]) * p.dy
                   + (p.vx[bidx(p, k + 1, j, i)] - p.vx[bidx(p, k, j, i)]) * p.udx3m[k]) * usdtal;
            }
        }
        lr[2 * ((xo >> 1) * N + j) + (xo & 1)] = d;
    }
    __syncthreads();
    lds_fft(lb, L, pl, -1);
    const double scale = 1.0 / (double(p.ny) * double(p.nz));
    for (int t = threadIdx.x; t < TX * p.nymh; t += blockDim.x) {
        const int xo = t % TX, m = t / TX, k = k0 + xo;
        if (k >= p.nxm) continue;
        const int line = xo >> 1;
        const double2 zm = lb[line * N + m], zn = lb[line * N + (m == 0 ? 0 : N - m)];
        double2 out;
        if ((xo & 1) == 0) out = make_double2(0.5 * (zm.x + zn.x), 0.5 * (zm.y - zn.y));     // (Z_m + conj Z_{N-m})/2
        else out = make_double2(0.5 * (zm.y + zn.y), -0.5 * (zm.x - zn.x));                   // (Z_m - conj Z_{N-m})/(2i)
        spec[((long long)i * p.nymh + m) * p.nxm + k] = make_double2(out.x * scale, out.y * scale);
    }
}

// Complex transform along z (pressure.F90:176 forward, :189 backward), in
// place on the spectral array.  Block: L x-adjacent lines at one y-mode m.
__global__ void __launch_bounds__(256) k_fft_z(KP p, FftPlan pl, double2 *spec, int L, int sign) {
    extern __shared__ double2 lb[];
    const int N = p.nz, m = blockIdx.y, k0 = blockIdx.x * L;
    const long long sz = (long long)p.nymh * p.nxm;
    double2 *g = spec + (long long)m * p.nxm;
    for (int t = threadIdx.x; t < L * N; t += blockDim.x) {
        const int xo = t % L, i = t / L, k = k0 + xo;
        lb[xo * N + i] = k < p.nxm ? g[i * sz + k] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    lds_fft(lb, L, pl, sign);
    for (int t = threadIdx.x; t < L * N; t += blockDim.x) {
        const int xo = t % L, i = t / L, k = k0 + xo;
        if (k < p.nxm) g[i * sz + k] = lb[xo * N + i];
    }
}

// Inverse real y transform (pressure.F90:193-195): two x-adjacent spectra
// A, B (Hermitian, imaginary parts of the m=0 and m=N/2 terms ignored as a
// c2r transform does) become Z = A + iB, whose inverse is a + ib.
__global__ void __launch_bounds__(256) k_ifft_y(KP p, FftPlan pl, const double2 *spec, int L) {
    extern __shared__ double2 lb[];
    const int N = p.ny, TX = 2 * L, i = blockIdx.y, k0 = blockIdx.x * TX;
    for (int t = threadIdx.x; t < L * p.nymh; t += blockDim.x) {
        const int line = t % L, m = t / L;
        const int ka = k0 + 2 * line, kb = ka + 1;
        const double2 *row = spec + ((long long)i * p.nymh + m) * p.nxm;
        double2 A = ka < p.nxm ? row[ka] : make_double2(0.0, 0.0);
        double2 B = kb < p.nxm ? row[kb] : make_double2(0.0, 0.0);
        if (m == 0 || 2 * m == N) { A.y = 0.0; B.y = 0.0; }
        lb[line * N + m] = make_double2(A.x - B.y, A.y + B.x);
        if (m > 0 && N - m > m) lb[line * N + (N - m)] = make_double2(A.x + B.y, B.x - A.y);
    }
    __syncthreads();
    lds_fft(lb, L, pl, +1);
    for (int t = threadIdx.x; t < TX * N; t += blockDim.x) {
        const int xo = t % TX, j = t / TX, k = k0 + xo;
        if (k >= p.nxm) continue;
        const double2 z = lb[(xo >> 1) * N + j];
        p.dph[bidx(p, k, j, i)] = (xo & 1) ? z.y : z.x;
    }
}

// Per-mode wall-normal tridiagonal (SolveTridiagonalPressure,
// pressure.F90:201-236) with factors precomputed at set-up.  Modes i and
// nz-i share a matrix (ak1 is even), so one block of two wavefronts solves
// both against one copy of the factors.  grid (nymh, nz/2+1).
struct PoissonFactors {
    const double *m, *g, *e;   // [mode'][row padded to 64*CH], mode' = m*(nz/2+1) + i'
    int npad;
};
template <int CH>
__global__ void __launch_bounds__(128) k_poisson_tri(KP p, PoissonFactors pf, double2 *spec) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = blockIdx.x, ih = blockIdx.y, nzh = p.nz / 2 + 1;
    const int i = wave == 0 ? ih : p.nz - ih;
    if (wave == 1 && (ih == 0 || i == ih || i >= p.nz)) return;
    const long long fo = ((long long)m * nzh + ih) * pf.npad;
    double2 *col = spec + ((long long)i * p.nymh + m) * p.nxm;
    double fm[CH], fg[CH], fe[CH];
    double2 v[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
        const int r = lane * CH + q;
        fm[q] = pf.m[fo + r]; fg[q] = pf.g[fo + r]; fe[q] = pf.e[fo + r];
        v[q] = r < p.nxm ? col[r] : make_double2(0.0, 0.0);
    }
    // forward y_r = d_r - m_r y_{r-1} (complex B, real A)
    double A = 1.0;
    double2 Bc = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < CH; ++q) { Bc = make_double2(-fm[q] * Bc.x + v[q].x, -fm[q] * Bc.y + v[q].y); A = -fm[q] * A; }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double Ap = __shfl_up(A, off, 64), Bx = __shfl_up(Bc.x, off, 64), By = __shfl_up(Bc.y, off, 64);
        if (lane >= off) { Bc = make_double2(A * Bx + Bc.x, A * By + Bc.y); A = A * Ap; }
    }
    double2 y = make_double2(__shfl_up(Bc.x, 1, 64), __shfl_up(Bc.y, 1, 64));
    if (lane == 0) y = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < CH; ++q) { y = make_double2(-fm[q] * y.x + v[q].x, -fm[q] * y.y + v[q].y); v[q] = y; }
    // backward x_r = g_r y_r - e_r x_{r+1}
    A = 1.0; Bc = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = CH - 1; q >= 0; --q) {
        Bc = make_double2(-fe[q] * Bc.x + fg[q] * v[q].x, -fe[q] * Bc.y + fg[q] * v[q].y); A = -fe[q] * A;
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double Ap = __shfl_down(A, off, 64), Bx = __shfl_down(Bc.x, off, 64), By = __shfl_down(Bc.y, off, 64);
        if (lane + off < 64) { Bc = make_double2(A * Bx + Bc.x, A * By + Bc.y); A = A * Ap; }
    }
    double2 x = make_double2(__shfl_down(Bc.x, 1, 64), __shfl_down(Bc.y, 1, 64));
    if (lane == 63) x = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = CH - 1; q >= 0; --q) {
        x = make_double2(-fe[q] * x.x + fg[q] * v[q].x, -fe[q] * x.y + fg[q] * v[q].y); v[q] = x;
    }
#pragma unroll
    for (int q = 0; q < CH; ++q) {
        const int r = lane * CH + q;
        if (r < p.nxm) col[r] = v[q];
    }
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

}  // namespace afid
