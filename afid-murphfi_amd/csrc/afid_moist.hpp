// afid_moist.hpp -- the specific humidity of the moist convection model
// (src/moisture.F90): a scalar on the base grid advected by u, diffused with
// kappa_q = 1/pecq, relaxed towards saturation qsat = exp(alpha_q (T -
// beta_q x)) where supersaturated, with the latent heat of that condensation
// added to the buoyancy T.
//
// Per-routine kernels (one per reference routine) and the fused column
// kernel of the step: explicit term, implicit RHS, Crank-Nicolson x-solve
// and update of one humidity column per wavefront, written to a second
// buffer while the other wavefronts still read the old column as a
// neighbour.  FP64 in both arithmetic forms.
#pragma once
#include <hip/hip_runtime.h>
#include "afid_params.hpp"
#include "afid_stencils.hpp"
#include "afid_kernels.hpp"

namespace afid {

// humidity column (j, i), its y/z neighbours, the velocity columns its
// fluxes read and the wall values
struct NbQ {
    const double *q, *q_jm, *q_jp, *q_im, *q_ip;
    const double *vx, *vy, *vy_jp, *vz, *vz_ip, *qs;
    double hbp, htp;
};
__device__ __forceinline__ NbQ make_nbq(const KP &p, int j, int i) {
    const int jm = wrapm(j, p.ny), jp = wrapp(j, p.ny), im = wrapm(i, p.nz), ip = wrapp(i, p.nz);
    auto o = [&](int jj, int ii) { return (long long)ii * p.plane + (long long)jj * p.ldx; };
    const long long c = o(j, i);
    NbQ n;
    n.q = p.humid + c; n.q_jm = p.humid + o(jm, i); n.q_jp = p.humid + o(jp, i);
    n.q_im = p.humid + o(j, im); n.q_ip = p.humid + o(j, ip);
    n.vx = p.vx + c; n.vy = p.vy + c; n.vy_jp = p.vy + o(jp, i); n.vz = p.vz + c; n.vz_ip = p.vz + o(j, ip);
    n.qs = p.qsat + c;
    n.hbp = p.humbp[(long long)i * p.ny + j];
    n.htp = p.humtp[(long long)i * p.ny + j];
    return n;
}

// ExplicitHumidity, moisture.F90:234-289, row k = kc-1 of 0..nxm-1
__device__ __forceinline__ double ex_humid(const KP &p, const NbQ &n, unsigned k) {
    const double udy = 0.5 * p.dy, udz = 0.5 * p.dz;
    const double qc = n.q[k];
    double hqx;
    if (k == 0)
        hqx = (n.vx[k + 1] * (n.q[k + 1] + qc) - n.vx[k] * 2.0 * n.hbp) * 0.5 * p.udx3m[k];
    else if (k == unsigned(p.nxm - 1))
        hqx = (n.vx[k + 1] * 2.0 * n.htp - n.vx[k] * (qc + n.q[k - 1])) * 0.5 * p.udx3m[k];
    else
        hqx = (n.vx[k + 1] * (n.q[k + 1] + qc) - n.vx[k] * (qc + n.q[k - 1])) * 0.5 * p.udx3m[k];
    const double hqy = (n.vy_jp[k] * (n.q_jp[k] + qc) - n.vy[k] * (qc + n.q_jm[k])) * udy;
    const double hqz = (n.vz_ip[k] * (n.q_ip[k] + qc) - n.vz[k] * (qc + n.q_im[k])) * udz;
    const double dyyq = (n.q_jp[k] - 2.0 * qc + n.q_jm[k]) * p.dyq;
    const double dzzq = (n.q_ip[k] - 2.0 * qc + n.q_im[k]) * p.dzq;
    double r = p.kappa_q * (dyyq + dzzq) - (hqx + hqy + hqz);
    const double qs = n.qs[k];
    if (qc > qs) r = r + (qs - qc) * p.q_itau;          // :274-278
    return r;
}

// ImplicitHumidity, moisture.F90:344-368: RHS of row k from the explicit
// term h and its history hk
__device__ __forceinline__ double rhs_humid(const KP &p, const NbQ &n, double h, double hk, unsigned k) {
    double dxxq;
    if (k == 0)
        dxxq = n.q[k + 1] * p.ap3ssk[k] + n.q[k] * p.ac3ssk[k] - (p.ap3ssk[k] + p.ac3ssk[k]) * n.hbp * p.q_fixs;
    else if (k == unsigned(p.nxm - 1))
        dxxq = -(p.am3ssk[k] + p.ac3ssk[k]) * n.htp * p.q_fixn + n.q[k] * p.ac3ssk[k] + n.q[k - 1] * p.am3ssk[k];
    else
        dxxq = n.q[k + 1] * p.ap3ssk[k] + n.q[k] * p.ac3ssk[k] + n.q[k - 1] * p.am3ssk[k];
    return (p.ga * h + p.ro * hk + (p.al / p.pecq) * dxxq) * p.dt;
}

// ---------------------------------------------------------------------------
// per-routine kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_explicit_humid(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const NbQ n = make_nbq(p, j, i);
    p.hhumid[bidx(p, k, j, i)] = ex_humid(p, n, k);
}
__global__ void __launch_bounds__(256) k_add_condensation(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c0 = (long long)i * p.plane + (long long)j * p.ldx;
    const double q = p.humid[c0 + k], qs = p.qsat[c0 + k];
    if (q > qs) p.hro[c0 + k] = p.hro[c0 + k] + p.q_gtau * (q - qs);
}
__global__ void __launch_bounds__(256) k_rhs_humid(KP p, double *rhs) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c = bidx(p, k, j, i);
    const NbQ n = make_nbq(p, j, i);
    const double h = p.hhumid[c];
    rhs[c] = rhs_humid(p, n, h, p.ruhumid[c], k);
    p.ruhumid[c] = h;
}
// UpdateSaturation, moisture.F90:321-333
__global__ void __launch_bounds__(256) k_update_saturation(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c = bidx(p, k, j, i);
    p.qsat[c] = exp(p.q_alpha * (p.temp[c] - p.q_beta * p.xm[k]));
}
// SetHumidityBCs, moisture.F90:129-148 (unsat_top: qfixN = 2 in humid.in)
__global__ void __launch_bounds__(256) k_humidity_bcs(KP p, double *hbp, double *htp, int unsat_top) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)p.ny * p.nz) return;
    hbp[t] = exp(p.q_alpha * p.tempbp[t]);
    htp[t] = unsat_top ? 0.0 : exp(p.q_alpha * (p.temptp[t] - p.q_beta));
}

// ---------------------------------------------------------------------------
// fused column kernel of the step: ExplicitHumidity, ImplicitHumidity and
// SolveImpEqnUpdate_Humidity for one column per wavefront; rows r = lane,
// lane+64, ... in the stencil phase, the RHS through LDS to the solver's
// layout, the new column to out (the other buffer).
// ---------------------------------------------------------------------------
template <int CH>
__global__ void __launch_bounds__(256) k_column_humid(KP p, TriFactors f, double *out) {
    __shared__ double lds[4][lds_col_size<CH>()];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned lane = threadIdx.x & 63;
    int j, i;
    const bool live = strip_column(xcd_block(0), wave, p.ny, p.nz, j, i);
    double *buf = lds[wave];
    const NbQ n = make_nbq(p, j, i);
    const long long c0 = (long long)i * p.plane + (long long)j * p.ldx;
    double *hk = p.ruhumid + c0;
    const unsigned nxm = p.nxm, nx = p.nx;
#pragma unroll 1
    for (unsigned r = lane; r < 64u * CH; r += 64) {
        double rhs = 0.0;
        if (live && r < nxm) {
            const double h = ex_humid(p, n, r);
            rhs = rhs_humid(p, n, h, hk[r], r);
            hk[r] = h;
        }
        buf[lds_row<CH>(r)] = rhs;
    }
    __syncthreads();
    column_solve_lds<CH>(buf, f);
    __syncthreads();
    if (!live) return;
    double *o = out + c0;
#pragma unroll 1
    for (unsigned r = lane; r < nx; r += 64) {
        const double q = n.q[r];
        o[r] = r < nxm ? q + buf[lds_row<CH>(r)] : q;
    }
}

// ---------------------------------------------------------------------------
// CalcMoistStats, moisture.F90:422-505: qbar, qsbar, qrms, qrel, vxq, vyq,
// vzq (nxm values each, the order of means.h5).  Same two-kernel structure
// as the mean profiles: a workgroup takes 64 x points and 4 y columns, each
// wavefront marches along z summing in FP64; the groups meet in LDS, then a
// second kernel sums the groups in a fixed order and normalises.
// ---------------------------------------------------------------------------
constexpr int MOIST_NQ = 7;
__global__ void __launch_bounds__(256) k_moist_partials(KP p, double *__restrict__ part) {
    __shared__ double red[4][MOIST_NQ][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int k = blockIdx.x * 64 + lane, nxm = p.nxm;
    const int j = blockIdx.y * 4 + wave;
    double a[MOIST_NQ];
#pragma unroll
    for (int q = 0; q < MOIST_NQ; ++q) a[q] = 0.0;
    if (k < nxm && j < p.ny) {
        const int jp = wrapp(j, p.ny);
        for (int i = 0; i < p.nz; ++i) {
            const int ip = wrapp(i, p.nz);
            const long long c = bidx(p, k, j, i);
            const double h = p.humid[c], qs = p.qsat[c];
            a[0] += h;
            a[1] += qs;
            a[2] += h * h;
            a[3] += h / qs;
            a[4] += 0.5 * (p.vx[c] + p.vx[c + 1]) * h;
            a[5] += 0.5 * (p.vy[c] + p.vy[bidx(p, k, jp, i)]) * h;
            a[6] += 0.5 * (p.vz[c] + p.vz[bidx(p, k, j, ip)]) * h;
        }
    }
#pragma unroll
    for (int q = 0; q < MOIST_NQ; ++q) red[wave][q][lane] = a[q];
    __syncthreads();
    if (wave == 0 && k < nxm) {
#pragma unroll
        for (int q = 0; q < MOIST_NQ; ++q)
            part[((long long)blockIdx.y * MOIST_NQ + q) * nxm + k] =
                ((red[0][q][lane] + red[1][q][lane]) + red[2][q][lane]) + red[3][q][lane];
    }
}
__global__ void __launch_bounds__(256) k_moist_finish(KP p, const double *__restrict__ part, int groups, double *__restrict__ out) {
    const int t = blockIdx.x * 256 + threadIdx.x, nxm = p.nxm;
    if (t >= MOIST_NQ * nxm) return;
    const int q = t / nxm, k = t - q * nxm;
    double s = 0.0;
    for (int g = 0; g < groups; ++g) s += part[((long long)g * MOIST_NQ + q) * nxm + k];
    const double inyzm = 1.0 / double(p.ny) / double(p.nz);       // :441
    out[t] = q == 2 ? sqrt(s * inyzm) : s * inyzm;                  // :465-473
}

}  // namespace afid
