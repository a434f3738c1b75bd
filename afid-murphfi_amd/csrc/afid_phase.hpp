// afid_phase.hpp -- phase-field kernels (src/phasefield.F90): the explicit
// and implicit phi terms on the refined grid, the fused phi column, the
// masked velocity solves, the coarse -> refined temperature transfer and the
// seeded sphere.  The refined -> coarse transfers (phic, latent heat) share
// k_interp_coarse_tile with InterpSalMultigrid.
#pragma once
#include <hip/hip_runtime.h>
#include "afid_params.hpp"
#include "afid_stencils.hpp"
#include "afid_kernels.hpp"

namespace afid {

// ExplicitPhase, phasefield.F90:345-383, row k of a refined column: lateral
// diffusion D (phi_yy + phi_zz) minus the Allen-Cahn source
// pf_B phi (1 - phi) (1 - 2 phi + A (T_r - T_m)), in arithmetic T
template <typename T>
__device__ __forceinline__ T ex_phase_val(const KP &p, T pc, T pym, T pyp, T pzm, T pzp, T tr) {
    const T dyyp = (pyp - T(2.0) * pc + pym) * T(p.pf_d * p.dyqr);
    const T dzzp = (pzp - T(2.0) * pc + pzm) * T(p.pf_d * p.dzqr);
    const T nlphi = T(p.pf_b) * pc * (T(1.0) - pc) * (T(1.0) - T(2.0) * pc + T(p.pf_a) * (tr - T(p.pf_tm)));
    return dyyp + dzzp - nlphi;
}
// ImplicitPhase, phasefield.F90:386-419: zero-gradient x second derivative
__device__ __forceinline__ double rhs_phase_val(const KP &p, unsigned k, double h, double ru, double pm, double pc,
                                                double pp) {
    double dxxp;
    if (k == 0) dxxp = pp * p.ap3spkr[k] + pc * p.ac3spkr[k];
    else if (k == unsigned(p.nxmr - 1)) dxxp = pc * p.ac3spkr[k] + pm * p.am3spkr[k];
    else dxxp = pp * p.ap3spkr[k] + pc * p.ac3spkr[k] + pm * p.am3spkr[k];
    return (p.ga * h + p.ro * ru + (p.al * p.pf_d) * dxxp) * p.dt;
}

// per-routine kernels (point layout of afid_kernels.hpp)
__global__ void __launch_bounds__(256) k_explicit_phase(KP p) {
    AFID_POINT_PROLOGUE(p.nxmr, p.nyr, p.nzr)
    const double *f = p.phi;
    p.hphi[ridx(p, k, j, i)] = ex_phase_val<double>(p, f[ridx(p, k, j, i)], f[ridx(p, k, jm, i)], f[ridx(p, k, jp, i)],
                                            f[ridx(p, k, j, im)], f[ridx(p, k, j, ip)], p.tempr[ridx(p, k, j, i)]);
}
__global__ void __launch_bounds__(256) k_rhs_phase(KP p) {
    AFID_POINT_PROLOGUE(p.nxmr, p.nyr, p.nzr)
    const long long c = ridx(p, k, j, i);
    const double *f = p.phi;
    const double pm = k > 0 ? f[c - 1] : 0.0, pp = k < p.nxmr - 1 ? f[c + 1] : 0.0;
    const double h = p.hphi[c];
    p.rhsr[c] = rhs_phase_val(p, k, h, p.ruphi[c], pm, f[c], pp);
    p.ruphi[c] = h;
}
// AddVolumePenalty, phasefield.F90:525-566
__global__ void __launch_bounds__(256) k_volume_penalty(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c = bidx(p, k, j, i);
    const double *pc = p.phic;
    if (k >= 1) p.qcap[c] -= (pc[c] + pc[c - 1]) * p.vx[c] * p.heta;
    p.dph[c] -= (pc[c] + pc[bidx(p, k, jm, i)]) * p.vy[c] * p.heta;
    p.dq[c] -= (pc[c] + pc[bidx(p, k, j, im)]) * p.vz[c] * p.heta;
}

// ExplicitPhase + ImplicitPhase + SolveImpEqnUpdate_Phi (phasefield.F90:
// 345-466) for one refined column per wavefront, rows in pairs as in
// k_column_sal.  Writes the new phi to phi_out, the explicit term to ruphi
// and the increment phi(l+1) - phi(l) to rhsr (read by AddLatentHeat).
// RT: storage and explicit-term arithmetic of the refined arrays; the
// implicit RHS, the solve and the update run in FP64.
template <int CH, typename RT>
__global__ void __launch_bounds__(256) k_column_phi(KP p, TriFactors f, RT *phi_out) {
    using V = typename Pair<RT>::T;
    constexpr int M2 = (64 * CH + 127) / 128;
    __shared__ double lds[4][lds_col_size<CH>()];
    // the solver's folded factors, read by every column of the workgroup
    __shared__ double fac[2][64 * CH];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned lane = threadIdx.x & 63;
    for (int q = threadIdx.x; q < 64 * CH; q += 256) { fac[0][q] = f.ln[q]; fac[1][q] = f.e[q]; }
    int j, i;
    const bool live = strip_column(xcd_block(0), wave, p.nyr, p.nzr, j, i);
    double *buf = lds[wave];
    const int jm = wrapm(j, p.nyr), jp = wrapp(j, p.nyr), im = wrapm(i, p.nzr), ip = wrapp(i, p.nzr);
    auto col = [&](int jj, int ii) { return (long long)ii * p.planer + (long long)jj * p.ldxr; };
    const long long c0 = col(j, i);
    const RT *phi = sel<RT>(p.phi, p.phif);
    const RT *ph = phi + c0, *pym = phi + col(jm, i), *pyp = phi + col(jp, i);
    const RT *pzm = phi + col(j, im), *pzp = phi + col(j, ip), *tr = sel<RT>(p.tempr, p.temprf) + c0;
    RT *ru = sel<RT>(p.ruphi, p.ruphif) + c0;
    const unsigned nxmr = p.nxmr;
    V own[M2];
#pragma unroll
    for (int m = 0; m < M2; ++m) {
        const unsigned r0 = 2 * lane + 128 * m, r1 = r0 + 1;
        double rhs0 = 0.0, rhs1 = 0.0;
        own[m] = V{RT(0), RT(0)};
        if (live && r0 < nxmr) {
            // r0 even, rows 8-aligned: aligned pairs; r0 + 1 <= nxmr < ldxr
            const V pc = *reinterpret_cast<const V *>(ph + r0);
            const V a = *reinterpret_cast<const V *>(pym + r0), b = *reinterpret_cast<const V *>(pyp + r0);
            const V cz = *reinterpret_cast<const V *>(pzm + r0), d = *reinterpret_cast<const V *>(pzp + r0);
            const V t = *reinterpret_cast<const V *>(tr + r0), u = *reinterpret_cast<const V *>(ru + r0);
            const RT pm = r0 > 0 ? ph[r0 - 1] : RT(0);
            const RT pp2 = r1 + 1 < nxmr ? ph[r1 + 1] : RT(0);
            const RT h0 = ex_phase_val<RT>(p, pc.x, a.x, b.x, cz.x, d.x, t.x);
            rhs0 = rhs_phase_val(p, r0, double(h0), double(u.x), double(pm), double(pc.x), double(pc.y));
            own[m] = pc;
            if (r1 < nxmr) {
                const RT h1 = ex_phase_val<RT>(p, pc.y, a.y, b.y, cz.y, d.y, t.y);
                rhs1 = rhs_phase_val(p, r1, double(h1), double(u.y), double(pc.x), double(pc.y), double(pp2));
                *reinterpret_cast<V *>(ru + r0) = V{h0, h1};
            } else {
                ru[r0] = h0;
            }
        }
        if (r0 < 64u * CH) buf[lds_row<CH>(r0)] = rhs0 * f.sn[r0];
        if (r1 < 64u * CH) buf[lds_row<CH>(r1)] = rhs1 * f.sn[r1];
    }
    __syncthreads();
    column_solve_lds_folded<CH>(buf, fac[0], fac[1]);
    __syncthreads();
    if (!live) return;
    RT *out = phi_out + c0, *dphi = sel<RT>(p.rhsr, p.rhsrf) + c0;
#pragma unroll
    for (int m = 0; m < M2; ++m) {
        const unsigned r0 = 2 * lane + 128 * m, r1 = r0 + 1;
        if (r0 >= nxmr) continue;
        const double d0 = buf[lds_row<CH>(r0)];
        if (r1 < nxmr) {
            const double d1 = buf[lds_row<CH>(r1)];
            *reinterpret_cast<V *>(out + r0) = V{RT(double(own[m].x) + d0), RT(double(own[m].y) + d1)};
            *reinterpret_cast<V *>(dphi + r0) = V{RT(d0), RT(d1)};
        } else {
            out[r0] = RT(double(own[m].x) + d0);
            dphi[r0] = RT(d0);
        }
    }
    if (lane == 0) out[nxmr] = ph[nxmr];        // the unused index nxr is carried over
}

// SolveImpEqnUpdate_X_pf / _YZ_pf (phasefield.F90:821-930) for the
// per-routine path: one column per wavefront, rhs as assembled by
// ImplicitAndUpdateV*, q updated in place (rows 2..nxm of vx, 1..nxm of
// vy / vz).  AX 0: vx (rows 1 and nx are identity rows with rhs -vx, rows
// kc test 0.5 (phic(kc) + phic(kc+1))); AX 1 / 2: vy / vz (phic averaged
// with the -y / -z neighbour).
template <int CH, int AX>
__global__ void __launch_bounds__(256) k_tri_update_pf(KP p, TriFactors f, const double *rhs, double *q) {
    const long long col = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (col >= (long long)p.ny * p.nz) return;
    const int lane = threadIdx.x & 63;
    const int j = int(col % p.ny), i = int(col / p.ny);
    const long long base = (long long)i * p.plane + (long long)j * p.ldx;
    const long long nbr = AX == 1 ? (long long)i * p.plane + (long long)wrapm(j, p.ny) * p.ldx
                        : AX == 2 ? (long long)wrapm(i, p.nz) * p.plane + (long long)j * p.ldx : base;
    const int n = f.n, nxm = p.nxm;
    double v[CH], a[CH], c[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;
        v[m] = 0.0; a[m] = 0.0; c[m] = 0.0;
        if (r >= n) continue;
        bool solid;
        if (AX == 0) solid = (r == 0 || r == nxm) ? true : 0.5 * (p.phic[base + r] + p.phic[base + r + 1]) > 0.9;
        else solid = 0.5 * (p.phic[base + r] + p.phic[nbr + r]) > 0.9;
        if (solid) { v[m] = -q[base + r]; }
        else { v[m] = rhs[base + r] * f.s[r]; a[m] = f.a[r]; c[m] = f.c[r]; }
    }
    wave_tri_solve_general<CH>(v, a, c);
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;
        if (AX == 0 ? (r >= 1 && r < nxm) : (r < n)) q[base + r] += v[m];
    }
}

// InterpTempMultigrid (phasefield.F90:498-520): T with the ghosts
// tpdv(0) = (1-2 TfixS) T(1) + 2 TfixS tempbp and tpdv(nx) likewise with
// TfixN, zero beyond, interpolated to the refined cell centres
// (interpolate_xyz_to_refined, HermiteInterpolations.F90:182-214).  Tile of
// VXR_TJ x VXR_TI refined columns, x in chunks of 64; the coarse block is
// staged in LDS and contracted along z, y, then x.
__device__ __forceinline__ double temp_ghosted(const KP &p, int xc, int j, int i) {
    const double *c = p.temp + (long long)i * p.plane + (long long)j * p.ldx;
    if (xc >= 1 && xc <= p.nxm) return c[xc - 1];
    const long long pl = (long long)i * p.ny + j;
    if (xc == 0) return (1.0 - 2.0 * p.tfixs) * c[0] + 2.0 * p.tfixs * p.tempbp[pl];
    if (xc == p.nx) return (1.0 - 2.0 * p.tfixn) * c[p.nxm - 1] + 2.0 * p.tfixn * p.temptp[pl];
    return 0.0;
}
template <typename RT>
__global__ void __launch_bounds__(256) k_interp_temp_tile(KP p, int xcmax, int cjmax, int cimax) {
    extern __shared__ __align__(16) unsigned char smraw[];
    __shared__ int cjs[16], cis[16];                  // the tile's coarse rows (periodic), cjmax, cimax <= 16
    __shared__ RT yw[VXR_TJ][4], zw[VXR_TI][4];
    __shared__ int yo[VXR_TJ], zo[VXR_TI];
    RT *sm = reinterpret_cast<RT *>(smraw);
    RT *tempr = sel<RT>(p.tempr, p.temprf);
    const RT *xs_w = sel<RT>(p.xs_w, p.xs_wf), *ys_w = sel<RT>(p.ys_w, p.ys_wf), *zs_w = sel<RT>(p.zs_w, p.zs_wf);
    const int J0 = blockIdx.x * VXR_TJ, I0 = blockIdx.y * VXR_TI;
    const int nj = min(VXR_TJ, p.nyr - J0), ni = min(VXR_TI, p.nzr - I0);
    const int cj0 = p.ys_src[J0] - 1, CJ = p.ys_src[J0 + nj - 1] - 1 + 4 - cj0;
    const int ci0 = p.zs_src[I0] - 1, CI = p.zs_src[I0 + ni - 1] - 1 + 4 - ci0;
    RT *B = sm;                                       // [ci][cj][xc]
    RT *Z = B + xcmax * cjmax * cimax;                // [ti][cj][xc]
    RT *Q = Z + xcmax * cjmax * VXR_TI;               // [ti][tj][xc]
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t < CJ) cjs[t] = wrapf(cj0 + t, p.ny);
    if (t >= 32 && t - 32 < CI) cis[t - 32] = wrapf(ci0 + t - 32, p.nz);
    if (t < nj) {
        yo[t] = p.ys_src[J0 + t] - 1 - cj0;
        for (int a = 0; a < 4; ++a) yw[t][a] = ys_w[4 * (J0 + t) + a];
    }
    if (t >= 64 && t - 64 < ni) {
        zo[t - 64] = p.zs_src[I0 + t - 64] - 1 - ci0;
        for (int a = 0; a < 4; ++a) zw[t - 64][a] = zs_w[4 * (I0 + t - 64) + a];
    }
    for (int R0 = 0; R0 < p.nxmr; R0 += 64) {
        const int R1 = min(R0 + 64, p.nxmr);
        const int x0 = p.xs_src[R0], XC = p.xs_src[R1 - 1] + 4 - x0;
        __syncthreads();
        {
            Walk3 w(t, 256, XC, CJ);
            for (int q = t; q < XC * CJ * CI; q += 256, w.next())
                B[(w.c * cjmax + w.b) * xcmax + w.a] = RT(temp_ghosted(p, x0 + w.a, cjs[w.b], cis[w.c]));
        }
        __syncthreads();
        Walk3 wz(t, 256, XC, CJ);
        for (int q = t; q < XC * CJ * ni; q += 256, wz.next()) {
            const int xc = wz.a, cj = wz.b, ti = wz.c;
            const RT *b = B + (zo[ti] * cjmax + cj) * xcmax + xc;
            const int st = cjmax * xcmax;
            Z[(ti * cjmax + cj) * xcmax + xc] = b[0] * zw[ti][0] + b[st] * zw[ti][1] + b[2 * st] * zw[ti][2] + b[3 * st] * zw[ti][3];
        }
        __syncthreads();
        Walk3 wy(t, 256, XC, nj);
        for (int q = t; q < XC * nj * ni; q += 256, wy.next()) {
            const int xc = wy.a, tj = wy.b, ti = wy.c;
            const RT *z = Z + (ti * cjmax + yo[tj]) * xcmax + xc;
            Q[(ti * VXR_TJ + tj) * xcmax + xc] = z[0] * yw[tj][0] + z[xcmax] * yw[tj][1] + z[2 * xcmax] * yw[tj][2] + z[3 * xcmax] * yw[tj][3];
        }
        __syncthreads();
        const int kr = R0 + lane;
        if (kr < R1) {
            const int xa = p.xs_src[kr] - x0;
            const RT *wx = xs_w + 4 * kr;
            const RT w0 = wx[0], w1 = wx[1], w2 = wx[2], w3 = wx[3];
            for (int ti = 0; ti < ni; ++ti) {
                RT *orow = tempr + (long long)(I0 + ti) * p.planer + (long long)J0 * p.ldxr + kr;
                for (int tj = wave; tj < nj; tj += 4) {
                    const RT *qq = Q + (ti * VXR_TJ + tj) * xcmax + xa;
                    orow[(long long)tj * p.ldxr] = qq[0] * w0 + qq[1] * w1 + qq[2] * w2 + qq[3] * w3;
                }
            }
        }
    }
    const int ncol = nj * ni;
    // the unused index nxr of tempr is zero (phasefield.F90:502)
    for (int cl = t; cl < ncol; cl += 256) tempr[ridx(p, p.nxmr, J0 + cl % nj, I0 + cl / nj)] = RT(0);
}

// set_ice_sphere (phasefield.F90:311-341) with the radius r0 of the
// seeded initial state: refined phi = (1 - tanh((r - r0)/(2 eps)))/2 ...
template <typename RT>
__global__ void __launch_bounds__(256) k_init_phi_sphere(KP p, const double *xmr, double r0, double eps, double alx3,
                                                          double ylen, double zlen) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= p.nxr || j >= p.nyr) return;
    const long long o = ridx(p, k, j, i);
    RT *phi = sel<RT>(p.phi, p.phif);
    if (k >= p.nxmr) { phi[o] = RT(0); return; }
    const double y = (j + 0.5) * ylen / p.nyr, z = (i + 0.5) * zlen / p.nzr;
    const double dx = xmr[k] - 0.5 * alx3, dy = y - 0.5 * ylen, dz = z - 0.5 * zlen;
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    phi[o] = RT(0.5 * (1.0 - tanh(0.5 * (r - r0) / eps)));
}
// ... and T = 0 in the base cells inside the sphere
__global__ void __launch_bounds__(256) k_init_temp_sphere(KP p, const double *xm, double r0, double alx3, double ylen,
                                                           double zlen) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= p.nxm || j >= p.ny) return;
    const double y = (j + 0.5) * ylen / p.ny, z = (i + 0.5) * zlen / p.nz;
    const double dx = xm[k] - 0.5 * alx3, dy = y - 0.5 * ylen, dz = z - 0.5 * zlen;
    if (sqrt(dx * dx + dy * dy + dz * dz) <= r0) p.temp[bidx(p, k, j, i)] = 0.0;
}

}  // namespace afid
