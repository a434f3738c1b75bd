// afid_means.hpp -- wall-parallel mean profiles (CalcMeanProfiles.F90:11-229)
// on the base grid: the 15 profiles the reference writes to means.h5, each
// of nxm values, in its write order (CalcMeanProfiles.F90:230-283).
//
// Two kernels.  k_mean_partials: a block of 64 x points (one per lane) and 4
// wavefronts, one y index j each (group = j / 4); each wavefront marches
// along z, so the z neighbours it reads are its own previous and next
// columns and the y neighbours are its sibling wavefronts' (both in cache),
// and accumulates the 15 per-x sums of CalcMeanProfiles.F90:44-73, :104-127
// and :165-215 in FP64 registers; the four wavefronts' sums are combined in
// LDS and written to part[group][q][k].  k_mean_finish: one
// thread per (q, k) sums the groups in a fixed order and applies the
// normalisations of :87-101 and :219-223, so a call's result does not
// depend on scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include "afid_params.hpp"
#include "afid_stencils.hpp"

namespace afid {

constexpr int MEAN_NQ = 15;         // Tbar vybar vzbar vxT vyT vzT Trms vxrms vyrms vzrms vxvy vxvz vyvz chiT epsilon
// Grid of the partial-sum kernels: (x chunks of 64, ceil(ny / 4) y groups);
// every y group writes one partial per profile.
__host__ __device__ constexpr int mean_ygroups(int ny) { return (ny + 3) / 4; }

// sdx[k] = dx / g3rm(k+1) (CreateGrid.F90:120-123), k = 0..nxm-1
__global__ void __launch_bounds__(256) k_mean_partials(KP p, const double *__restrict__ sdx, double *__restrict__ part) {
    __shared__ double red[4][MEAN_NQ][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane, nxm = p.nxm;
    const bool kok = k < nxm;
    double a[MEAN_NQ];
#pragma unroll
    for (int q = 0; q < MEAN_NQ; ++q) a[q] = 0.0;
    const double dy = p.dy, dz = p.dz;
    const int j = blockIdx.y * 4 + wave;
    if (kok && j < p.ny) {
        const int kp = k + 1, km = k - 1;
        const double s0 = sdx[k];
        const int jp = wrapp(j, p.ny), jm = wrapm(j, p.ny);
        for (int i = 0; i < p.nz; ++i) {
            const int ip = wrapp(i, p.nz), im = wrapm(i, p.nz);
            auto at = [&](const double *f, int kk, int jj, int ii) { return f[bidx(p, kk, jj, ii)]; };
            const double *vx = p.vx, *vy = p.vy, *vz = p.vz, *tp = p.temp;
            // :44-73
            const double t = at(tp, k, j, i);
            const double vxk = at(vx, k, j, i), vxp = at(vx, kp, j, i);
            const double vyk = at(vy, k, j, i), vyp = at(vy, k, jp, i);
            const double vzk = at(vz, k, j, i), vzp = at(vz, k, j, ip);
            a[0] += t;
            a[1] += 0.5 * (vyk + vyp);
            a[2] += 0.5 * (vzk + vzp);
            a[3] += 0.5 * (vxk + vxp) * t;
            a[4] += 0.5 * (vyk + vyp) * t;
            a[5] += 0.5 * (vzk + vzp) * t;
            a[6] += t * t;
            a[7] += 0.5 * (vxk * vxk + vxp * vxp);
            a[8] += 0.5 * (vyk * vyk + vyp * vyp);
            a[9] += 0.5 * (vzk * vzk + vzp * vzp);
            a[10] += 0.25 * (vxk + vxp) * (vyk + vyp);
            a[11] += 0.25 * (vxk + vxp) * (vzk + vzp);
            a[12] += 0.25 * (vyk + vyp) * (vzk + vzp);
            // chiT, :104-127: y/z terms at every k, then the x term (wall
            // forms at the first and last cell, tdx = 0.5 dx / g3rm)
            {
                const double ga = (at(tp, k, j, ip) - at(tp, k, j, im)) * 0.5 * dz;
                const double gb = (at(tp, k, jp, i) - at(tp, k, jm, i)) * 0.5 * dy;
                double ch = ga * ga + gb * gb;
                const double tdx = 0.5 * s0;
                if (k == 0) {
                    const double w = (at(tp, 1, j, i) - t + 2.0 * p.tfixs * (t - p.tempbp[(long long)i * p.ny + j])) * tdx;
                    ch += w * w;
                }
                if (k >= 1 && k <= nxm - 2) {
                    const double d = (at(tp, kp, j, i) - at(tp, km, j, i)) * tdx;
                    ch += d * d;
                }
                if (k == nxm - 1) {
                    const double w = (t - at(tp, km, j, i) + 2.0 * p.tfixn * (p.temptp[(long long)i * p.ny + j] - t)) * tdx;
                    ch += w * w;
                }
                a[13] += ch;
            }
            // epsilon, :165-215
            {
                const double e1 = (vzp - vzk) * dz;
                const double b1 = at(vy, k, jp, ip) - at(vy, k, jp, im), b2 = at(vy, k, j, ip) - at(vy, k, j, im);
                const double c1 = at(vx, kp, j, ip) - at(vx, kp, j, im), c2 = at(vx, k, j, ip) - at(vx, k, j, im);
                const double d1 = at(vz, k, jp, ip) - at(vz, k, jm, ip), d2 = at(vz, k, jp, i) - at(vz, k, jm, i);
                const double e2 = (vyp - vyk) * dy;
                const double f1 = at(vx, kp, jp, i) - at(vx, kp, jm, i), f2 = at(vx, k, jp, i) - at(vx, k, jm, i);
                const double g = (vxp - vxk) * p.udx3m[k];
                double ep = e1 * e1 + 0.125 * dz * dz * (b1 * b1 + b2 * b2) + 0.125 * dz * dz * (c1 * c1 + c2 * c2)
                          + 0.125 * dy * dy * (d1 * d1 + d2 * d2) + e2 * e2 + 0.125 * dy * dy * (f1 * f1 + f2 * f2) + g * g;
                const double tdx = 0.125 * s0 * s0;
                if (k == 0) {
                    const double z1 = at(vz, 1, j, ip) - vzp + p.inslws * 2.0 * (vzp - 0.0);
                    const double z2 = at(vz, 1, j, i) - vzk + p.inslws * 2.0 * (vzk - 0.0);
                    const double y1 = at(vy, 1, jp, i) - vyp + p.inslws * 2.0 * (vyp - p.xminusu);
                    const double y2 = at(vy, 1, j, i) - vyk + p.inslws * 2.0 * (vyk - p.xminusu);
                    ep += tdx * (z1 * z1 + z2 * z2) + tdx * (y1 * y1 + y2 * y2);
                }
                if (k >= 1 && k <= nxm - 2) {
                    const double p1 = at(vz, kp, j, ip) - at(vz, km, j, ip), p2 = at(vz, kp, j, i) - at(vz, km, j, i);
                    const double q1 = at(vy, kp, jp, i) - at(vy, km, jp, i), q2 = at(vy, kp, j, i) - at(vy, km, j, i);
                    ep += tdx * (p1 * p1 + p2 * p2) + tdx * (q1 * q1 + q2 * q2);
                }
                if (k == nxm - 1) {
                    const double z1 = vzp - at(vz, km, j, ip) + p.inslwn * 2.0 * (0.0 - vzp);
                    const double z2 = vzk - at(vz, km, j, i) + p.inslwn * 2.0 * (0.0 - vzk);
                    const double y1 = vyp - at(vy, km, jp, i) + p.inslwn * 2.0 * (p.xplusu - vyp);
                    const double y2 = vyk - at(vy, km, j, i) + p.inslwn * 2.0 * (p.xplusu - vyk);
                    ep += tdx * (z1 * z1 + z2 * z2) + tdx * (y1 * y1 + y2 * y2);
                }
                a[14] += ep;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < MEAN_NQ; ++q) red[wave][q][lane] = a[q];
    __syncthreads();
    if (wave == 0 && kok) {
#pragma unroll
        for (int q = 0; q < MEAN_NQ; ++q)
            part[((long long)blockIdx.y * MEAN_NQ + q) * nxm + k] =
                ((red[0][q][lane] + red[1][q][lane]) + red[2][q][lane]) + red[3][q][lane];
    }
}

// sum of the groups in order and the normalisations; out[q * nxm + k]
__global__ void __launch_bounds__(256) k_mean_finish(KP p, const double *__restrict__ part, int groups, double *__restrict__ out) {
    const int t = blockIdx.x * 256 + threadIdx.x, nxm = p.nxm;
    if (t >= MEAN_NQ * nxm) return;
    const int q = t / nxm, k = t - q * nxm;
    double s = 0.0;
    for (int g = 0; g < groups; ++g) s += part[((long long)g * MEAN_NQ + q) * nxm + k];
    s = s * (1.0 / double(p.ny)) * (1.0 / double(p.nz));     // *inym*inzm, left to right
    if (q >= 6 && q <= 9) s = sqrt(s);            // Trms, vxrms, vyrms, vzrms (:95-98)
    else if (q == 13) s /= p.pect;                 // chiT (:220)
    else if (q == 14) s /= p.ren;                  // epsilon (:222)
    out[t] = s;
}

// CalcSalStats + CalcDissipationSal (salinity.F90:510-577, :611-690, non-IBM
// branches) on the refined grid: Sbar, Srms, vxS, vyS, vzS, chiS (each nxmr
// values, the write order of :583-597).  RT is the storage type of the
// refined fields (FP32 in the mixed form); every sum is FP64.  sdxr[k] =
// dxr / g3rmr(k+1) (CreateMgrdGrid).  Same two-kernel structure as the base
// profiles.
constexpr int SAL_NQ = 6;
template <typename RT>
__global__ void __launch_bounds__(256) k_sal_stats_partials(KP p, const double *__restrict__ sdxr, double *__restrict__ part) {
    __shared__ double red[4][SAL_NQ][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane, n = p.nxmr;
    const bool kok = k < n;
    double a[SAL_NQ];
#pragma unroll
    for (int q = 0; q < SAL_NQ; ++q) a[q] = 0.0;
    if (kok) {
        const RT *sal = sel<RT>(p.sal, p.salf), *vxr = sel<RT>(p.vxr, p.vxrf), *vyr = sel<RT>(p.vyr, p.vyrf),
                 *vzr = sel<RT>(p.vzr, p.vzrf);
        const double tdx = 0.5 * sdxr[k], hdz = 0.5 * p.dzr, hdy = 0.5 * p.dyr;
        const int j = blockIdx.y * 4 + wave;
        const int jp = wrapp(j, p.nyr), jm = wrapm(j, p.nyr);
        for (int i = 0; i < p.nzr && j < p.nyr; ++i) {
            const int ip = wrapp(i, p.nzr), im = wrapm(i, p.nzr);
            auto at = [&](const RT *f, int kk, int jj, int ii) { return double(f[ridx(p, kk, jj, ii)]); };
            const double v = at(sal, k, j, i);
            a[0] += v;                                                      // :544-552
            a[1] += v * v;
            a[2] += 0.5 * (at(vxr, k, j, i) + at(vxr, k + 1, j, i)) * v;    // :555-563
            a[3] += 0.5 * (at(vyr, k, j, i) + at(vyr, k, jp, i)) * v;
            a[4] += 0.5 * (at(vzr, k, j, i) + at(vzr, k, j, ip)) * v;
            const double gz = (at(sal, k, j, ip) - at(sal, k, j, im)) * hdz;   // :677-678
            a[5] += gz * gz;
            const double gy = (at(sal, k, jp, i) - at(sal, k, jm, i)) * hdy;
            a[5] += gy * gy;
            if (k == 0) {                                                   // :680-682
                const double w = (at(sal, 1, j, i) - v + 2.0 * p.sfixs * (v - p.salbp[(long long)i * p.nyr + j])) * tdx;
                a[5] += w * w;
            }
            if (k >= 1 && k <= n - 2) {                                     // :683-685
                const double d = (at(sal, k + 1, j, i) - at(sal, k - 1, j, i)) * tdx;
                a[5] += d * d;
            }
            if (k == n - 1) {                                               // :686-688
                const double w = (v - at(sal, k - 1, j, i) + 2.0 * p.sfixn * (p.saltp[(long long)i * p.nyr + j] - v)) * tdx;
                a[5] += w * w;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < SAL_NQ; ++q) red[wave][q][lane] = a[q];
    __syncthreads();
    if (wave == 0 && kok) {
#pragma unroll
        for (int q = 0; q < SAL_NQ; ++q)
            part[((long long)blockIdx.y * SAL_NQ + q) * n + k] =
                ((red[0][q][lane] + red[1][q][lane]) + red[2][q][lane]) + red[3][q][lane];
    }
}

__global__ void __launch_bounds__(256) k_sal_stats_finish(KP p, const double *__restrict__ part, int groups, double *__restrict__ out) {
    const int t = blockIdx.x * 256 + threadIdx.x, n = p.nxmr;
    if (t >= SAL_NQ * n) return;
    const int q = t / n, k = t - q * n;
    double s = 0.0;
    for (int g = 0; g < groups; ++g) s += part[((long long)g * SAL_NQ + q) * n + k];
    const double inyzmr = 1.0 / double(p.nyr) / double(p.nzr);     // :537
    if (q == 1) s = sqrt(s * inyzmr);                               // :574-581
    else if (q == 5) s = s / p.pecs * inyzmr;
    else s = s * inyzmr;
    out[t] = s;
}

// CalcPhiStats (phasefield.F90:750-798): phibar and phirms of the refined
// phase field (each nxmr values, the write order of :787-791), RT its storage
// type (FP32 in the mixed form), FP64 sums.
template <typename RT>
__global__ void __launch_bounds__(256) k_phi_stats_partials(KP p, double *__restrict__ part) {
    __shared__ double red[4][2][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane, n = p.nxmr;
    const bool kok = k < n;
    double s0 = 0.0, s1 = 0.0;
    if (kok) {
        const RT *phi = sel<RT>(p.phi, p.phif);
        const int j = blockIdx.y * 4 + wave;
        for (int i = 0; i < p.nzr && j < p.nyr; ++i) {
            const double v = double(phi[ridx(p, k, j, i)]);
            s0 += v;                                                        // :770-776
            s1 += v * v;
        }
    }
    red[wave][0][lane] = s0;
    red[wave][1][lane] = s1;
    __syncthreads();
    if (wave == 0 && kok)
        for (int q = 0; q < 2; ++q)
            part[((long long)blockIdx.y * 2 + q) * n + k] =
                ((red[0][q][lane] + red[1][q][lane]) + red[2][q][lane]) + red[3][q][lane];
}
__global__ void __launch_bounds__(256) k_phi_stats_finish(KP p, const double *__restrict__ part, int groups, double *__restrict__ out) {
    const int t = blockIdx.x * 256 + threadIdx.x, n = p.nxmr;
    if (t >= 2 * n) return;
    const int q = t / n, k = t - q * n;
    double s = 0.0;
    for (int g = 0; g < groups; ++g) s += part[((long long)g * 2 + q) * n + k];
    const double inyzm = 1.0 / double(p.nyr) / double(p.nzr);      // :765
    out[t] = q == 0 ? s * inyzm : sqrt(s * inyzm);                  // :781-784
}

}  // namespace afid
