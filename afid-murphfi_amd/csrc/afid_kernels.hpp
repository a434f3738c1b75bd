// afid_kernels.hpp -- HIP kernels of the substep (except the Poisson
// transforms, in afid_poisson.hpp).
//
// Point kernels: block (64,4), thread (k, j) with x fastest so a wavefront
// reads 64 consecutive doubles; grid (ceil(n/64), ceil(ny/4), nz).
// Column kernels: one 64-lane wavefront per wall-normal column, each lane
// owning CH consecutive rows; the tridiagonal recurrences run as two
// wavefront-wide affine scans (forward elimination, back substitution).
#pragma once
#include <hip/hip_runtime.h>
#include "afid_params.hpp"
#include "afid_stencils.hpp"

namespace afid {

#define AFID_POINT_PROLOGUE(NXLIM, NY, NZ)                                    \
    const int k = blockIdx.x * 64 + threadIdx.x;                               \
    const int j = blockIdx.y * 4 + threadIdx.y;                                \
    const int i = blockIdx.z;                                                  \
    if (k >= (NXLIM) || j >= (NY)) return;                                     \
    const int jm = wrapm(j, NY), jp = wrapp(j, NY), im = wrapm(i, NZ), ip = wrapp(i, NZ); \
    (void)jm; (void)jp; (void)im; (void)ip;

// ---------------------------------------------------------------------------
// per-routine explicit kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_explicit_vx(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    if (k == 0) return;   // qcap(1) is not computed (ExplicitTermsVX.F90:50)
    p.qcap[bidx(p, k, j, i)] = ex_vx(p, k, j, jm, jp, i, im, ip);
}
__global__ void __launch_bounds__(256) k_explicit_vy(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    p.dph[bidx(p, k, j, i)] = ex_vy(p, k, j, jm, jp, i, im, ip);
}
__global__ void __launch_bounds__(256) k_explicit_vz(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    p.dq[bidx(p, k, j, i)] = ex_vz(p, k, j, jm, jp, i, im, ip);
}
__global__ void __launch_bounds__(256) k_explicit_temp(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    p.hro[bidx(p, k, j, i)] = ex_temp(p, k, j, jm, jp, i, im, ip);
}
__global__ void __launch_bounds__(256) k_explicit_sal(KP p) {
    AFID_POINT_PROLOGUE(p.nxmr, p.nyr, p.nzr)
    p.hsal[ridx(p, k, j, i)] = ex_sal(p, k, j, jm, jp, i, im, ip);
}
// AddSalBuoyancy, salinity.F90:495-506
__global__ void __launch_bounds__(256) k_add_sal_buoyancy(KP p, double *rkv) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c = bidx(p, k, j, i);
    rkv[c] = rkv[c] - p.bycs * p.salc[c];
}

// ---------------------------------------------------------------------------
// per-routine implicit RHS kernels (ImplicitAndUpdate*.F90 loops)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_rhs_vx(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    if (k == 0) return;
    const long long c = bidx(p, k, j, i);
    const double h = p.qcap[c];
    p.rhsx[c] = rhs_vx(p, h, p.rux[c], k, j, i);
    p.rux[c] = h;
}
__global__ void __launch_bounds__(256) k_rhs_vy(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c = bidx(p, k, j, i);
    const double h = p.dph[c];
    p.rhsy[c] = rhs_vy(p, h, p.ruy[c], k, j, jm, i);
    p.ruy[c] = h;
}
__global__ void __launch_bounds__(256) k_rhs_vz(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c = bidx(p, k, j, i);
    const double h = p.dq[c];
    p.rhsz[c] = rhs_vz(p, h, p.ruz[c], k, j, i, im);
    p.ruz[c] = h;
}
__global__ void __launch_bounds__(256) k_rhs_temp(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c = bidx(p, k, j, i);
    const double h = p.hro[c];
    p.rhst[c] = rhs_temp(p, h, p.rut[c], k, j, i);
    p.rut[c] = h;
}
__global__ void __launch_bounds__(256) k_rhs_sal(KP p) {
    AFID_POINT_PROLOGUE(p.nxmr, p.nyr, p.nzr)
    const long long c = ridx(p, k, j, i);
    const double h = p.hsal[c];
    p.rhsr[c] = rhs_sal(p, h, p.rusal[c], k, j, i);
    p.rusal[c] = h;
}

// Fused explicit terms + implicit RHS of the four base fields (the
// ExplicitTerms* and ImplicitAndUpdate* loops of one substep read only
// pre-substep fields, so they can run in one sweep); AddSalBuoyancy folded in.
__global__ void __launch_bounds__(256) k_explicit_rhs_all(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c = bidx(p, k, j, i);
    const double sb = (p.active_s == 1) ? p.bycs * p.salc[c] : 0.0;
    if (k > 0) {
        double h = ex_vx(p, k, j, jm, jp, i, im, ip);
        if (p.gaxis == 1) h -= sb;
        p.rhsx[c] = rhs_vx(p, h, p.rux[c], k, j, i);
        p.rux[c] = h;
    }
    {
        double h = ex_vy(p, k, j, jm, jp, i, im, ip);
        if (p.gaxis == 2) h -= sb;
        p.rhsy[c] = rhs_vy(p, h, p.ruy[c], k, j, jm, i);
        p.ruy[c] = h;
    }
    {
        double h = ex_vz(p, k, j, jm, jp, i, im, ip);
        if (p.gaxis == 3) h -= sb;
        p.rhsz[c] = rhs_vz(p, h, p.ruz[c], k, j, i, im);
        p.ruz[c] = h;
    }
    {
        double h = ex_temp(p, k, j, jm, jp, i, im, ip);
        p.rhst[c] = rhs_temp(p, h, p.rut[c], k, j, i);
        p.rut[c] = h;
    }
}
// Fused ExplicitSalinity + ImplicitSalinity RHS
__global__ void __launch_bounds__(256) k_explicit_rhs_sal(KP p) {
    AFID_POINT_PROLOGUE(p.nxmr, p.nyr, p.nzr)
    const long long c = ridx(p, k, j, i);
    double h = ex_sal(p, k, j, jm, jp, i, im, ip);
    p.rhsr[c] = rhs_sal(p, h, p.rusal[c], k, j, i);
    p.rusal[c] = h;
}

// ---------------------------------------------------------------------------
// wavefront tridiagonal solve (SolveImpEqnUpdate_{X,YZ,Temp,Sal})
// ---------------------------------------------------------------------------
// Inclusive scan of affine maps y -> A y + B across the 64 lanes, in lane
// order (UP) or reverse lane order.  Composition keeps the later map outside.
template <bool UP>
__device__ __forceinline__ void wave_affine_scan(double &A, double &B) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double Ap = UP ? __shfl_up(A, off, 64) : __shfl_down(A, off, 64);
        double Bp = UP ? __shfl_up(B, off, 64) : __shfl_down(B, off, 64);
        bool ok = UP ? (lane >= off) : (lane + off < 64);
        if (ok) { B = A * Bp + B; A = A * Ap; }
    }
}

// Solve the normalised system on the CH rows a lane owns (v in: raw RHS d,
// v out: solution x), with factors f (see TriFactors).
template <int CH>
__device__ __forceinline__ void wave_tri_solve(double (&v)[CH], const TriFactors &f) {
    const int lane = threadIdx.x & 63;
    const int r0 = lane * CH;
    double l[CH], e[CH], g[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = r0 + m;
        l[m] = f.l[r]; e[m] = f.e[r]; g[m] = f.g[r];
        v[m] *= f.s[r];
    }
    // forward: y_r = v_r - l_r y_{r-1}
    double A = 1.0, B = 0.0;
#pragma unroll
    for (int m = 0; m < CH; ++m) { B = -l[m] * B + v[m]; A = -l[m] * A; }
    wave_affine_scan<true>(A, B);
    double y = __shfl_up(B, 1, 64);
    if (lane == 0) y = 0.0;
#pragma unroll
    for (int m = 0; m < CH; ++m) { y = -l[m] * y + v[m]; v[m] = y; }
    // backward: x_r = g_r y_r - e_r x_{r+1}
    A = 1.0; B = 0.0;
#pragma unroll
    for (int m = CH - 1; m >= 0; --m) { B = -e[m] * B + g[m] * v[m]; A = -e[m] * A; }
    wave_affine_scan<false>(A, B);
    double x = __shfl_down(B, 1, 64);
    if (lane == 63) x = 0.0;
#pragma unroll
    for (int m = CH - 1; m >= 0; --m) { x = -e[m] * x + g[m] * v[m]; v[m] = x; }
}

// One column per wavefront: q[r] += solve(rhs)[r] for r in [u0,u1); rows
// listed in zero0/zero1 (if >= 0) are set to zero afterwards (the wall rows
// of vx, ImplicitAndUpdateVX.F90:86-87).  rhs holds the raw RHS on entry and
// the increment on exit when keep_rhs is set (as the reference's rhs array
// does); the fused step skips that write.
template <int CH>
__global__ void __launch_bounds__(256) k_tri_update(TriFactors f, double *rhs, double *q, int ldx,
                                                    long long plane, int ny, long long ncols, int u0, int u1,
                                                    int zero0, int zero1, int keep_rhs) {
    const long long col = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (col >= ncols) return;
    const int lane = threadIdx.x & 63;
    const int j = int(col % ny), i = int(col / ny);
    const long long base = (long long)i * plane + (long long)j * ldx;
    double v[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;
        v[m] = r < f.n ? rhs[base + r] : 0.0;
    }
    wave_tri_solve<CH>(v, f);
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;
        if (r < f.n) {
            if (keep_rhs) rhs[base + r] = v[m];
            if (r >= u0 && r < u1) q[base + r] += v[m];
            if (r == zero0 || r == zero1) q[base + r] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// pressure projection point kernels
// ---------------------------------------------------------------------------
// CalcLocalDivergence, pressure.F90:147-161
__global__ void __launch_bounds__(256) k_local_divergence(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const double usdtal = 1.0 / (p.dt * p.al);
    const double dq = (p.vz[bidx(p, k, j, ip)] - p.vz[bidx(p, k, j, i)]) * p.dz
                    + (p.vy[bidx(p, k, jp, i)] - p.vy[bidx(p, k, j, i)]) * p.dy
                    + (p.vx[bidx(p, k + 1, j, i)] - p.vx[bidx(p, k, j, i)]) * p.udx3m[k];
    p.dph[bidx(p, k, j, i)] = dq * usdtal;
}

// RemoveDivergence (pressure.F90:245-266) and/or CorrectPressure
// (pressure.F90:276-300); both read only dph, so one sweep does both.
__global__ void __launch_bounds__(256) k_project(KP p, int do_vel, int do_pr) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const double *d = p.dph;
    const int km = k == 0 ? 0 : k - 1, kp = k == p.nxm - 1 ? k : k + 1;
    const long long c = bidx(p, k, j, i);
    const double dc = d[c];
    if (do_vel) {
        const double adt = p.al * p.dt;
        p.vx[c] -= (dc - d[bidx(p, km, j, i)]) * (adt * p.udx3c[k]);
        p.vy[c] -= (dc - d[bidx(p, k, jm, i)]) * (adt * p.dy);
        p.vz[c] -= (dc - d[bidx(p, k, j, im)]) * (adt * p.dz);
    }
    if (do_pr) {
        const double be = p.al * p.beta;
        p.pr[c] += dc - be * ((d[bidx(p, k, j, ip)] - 2.0 * dc + d[bidx(p, k, j, im)]) * p.dzq +
                              (d[bidx(p, k, jp, i)] - 2.0 * dc + d[bidx(p, k, jm, i)]) * p.dyq +
                              (d[bidx(p, kp, j, i)] * p.apphk[k] + dc * p.acphk[k] + d[bidx(p, km, j, i)] * p.amphk[k]));
    }
}

// ---------------------------------------------------------------------------
// multiresolution transfers
// ---------------------------------------------------------------------------
// Gradient of one velocity component at coarse cell centres with the wall
// ghosts of InterpVelMgrd.F90:37-51 (vx), :75-86 (vy), :161-172 (vz).
// tpdv x index kc = -1..nx+1 is stored at kc+1; -1 and nx+1 are zero.
__global__ void __launch_bounds__(256) k_mgrd_gradient(KP p, int comp) {
    const int kk = blockIdx.x * 64 + threadIdx.x;   // storage index 0..nx+2
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int i = blockIdx.z;
    if (kk > p.nx + 2 || j >= p.ny) return;
    const int kc = kk - 1;                          // reference index
    auto grad = [&](int kc1) -> double {            // kc1 in 1..nxm
        const int k = kc1 - 1;
        if (comp == 0) return (p.vx[bidx(p, k + 1, j, i)] - p.vx[bidx(p, k, j, i)]) * p.udx3m[k];
        if (comp == 1) return (p.vy[bidx(p, k, wrapp(j, p.ny), i)] - p.vy[bidx(p, k, j, i)]) * p.dy;
        return (p.vz[bidx(p, k, j, wrapp(i, p.nz))] - p.vz[bidx(p, k, j, i)]) * p.dz;
    };
    double v;
    if (kc >= 1 && kc <= p.nxm) v = grad(kc);
    else if (kc == 0) v = (1 - 2 * p.inslws) * grad(1);
    else if (kc == p.nx) v = (1 - 2 * p.inslwn) * grad(p.nxm);
    else v = 0.0;
    p.tpdv[(long long)i * p.ny * p.ldt + (long long)j * p.ldt + kk] = v;
}

__device__ __forceinline__ double tpdv_at(const KP &p, int kc, int j, int i) {
    return p.tpdv[(long long)i * p.ny * p.ldt + (long long)j * p.ldt + (kc + 1)];
}
__device__ __forceinline__ int wrapn(int j, int n) { j %= n; return j < 0 ? j + n : j; }

// vxr: Hermite transfer of dvx/dx (interpolate_xyz_to_refined,
// HermiteInterpolations.F90:182-214) integrated along x (InterpVelMgrd.F90:58-66).
// One wavefront per refined column (jr, ir); coarse-x partial sums staged in LDS.
template <int CH>
__global__ void __launch_bounds__(256) k_vxr(KP p) {
    extern __shared__ double lds[];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long col = (long long)blockIdx.x * 4 + wave;
    const long long ncol = (long long)p.nyr * p.nzr;
    const int nq = p.nx + 3;                       // coarse x = -1..nx+1
    double *Q = lds + wave * nq;
    const bool live = col < ncol;
    const int jr = live ? int(col % p.nyr) : 0, ir = live ? int(col / p.nyr) : 0;
    if (live) {
        const int js = p.ys_src[jr], is = p.zs_src[ir];
        const double *wy = p.ys_w + 4 * jr, *wz = p.zs_w + 4 * ir;
        int jj[4], ii[4];
        for (int b = 0; b < 4; ++b) { jj[b] = wrapn(js - 1 + b, p.ny); ii[b] = wrapn(is - 1 + b, p.nz); }
        for (int q = lane; q < nq; q += 64) {
            const int kc = q - 1;
            double acc = 0.0;
            for (int b = 0; b < 4; ++b) {
                double s2 = tpdv_at(p, kc, jj[b], ii[0]) * wz[0] + tpdv_at(p, kc, jj[b], ii[1]) * wz[1]
                          + tpdv_at(p, kc, jj[b], ii[2]) * wz[2] + tpdv_at(p, kc, jj[b], ii[3]) * wz[3];
                acc = b == 0 ? s2 * wy[0] : acc + s2 * wy[b];
            }
            Q[q] = acc;
        }
    }
    __syncthreads();
    if (!live) return;
    // refined increments inc[r] = tpdvr(r) * (xcr(r+1) - xcr(r)), r = 0..nxmr-1
    double inc[CH];
    double run = 0.0;
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;
        double v = 0.0;
        if (r < p.nxmr) {
            const int xs = p.xs_src[r];          // reference coarse index of the first point
            const double *w = p.xs_w + 4 * r;
            const double *q = Q + (xs + 1);
            v = (q[0] * w[0] + q[1] * w[1] + q[2] * w[2] + q[3] * w[3]) * p.lxa[r];
        }
        run += v;
        inc[m] = run;                            // chunk-local inclusive prefix
    }
    double A = 1.0, B = run;                     // scan of sums: maps y -> y + run
    wave_affine_scan<true>(A, B);
    double off = __shfl_up(B, 1, 64);
    if (lane == 0) off = 0.0;
    const long long base = ridx(p, 0, jr, ir);
    if (lane == 0) p.vxr[base] = 0.0;            // vxr(1) = 0
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;
        if (r < p.nxmr) p.vxr[base + r + 1] = off + inc[m];
    }
}

// vyr (InterpVelMgrd.F90:71-151) and vzr (:157-237): first-guess plane by 2-D
// Hermite transfer of the coarse plane, then a running sum of the transferred
// gradient along y (resp. z).  One thread per (refined x, refined z) (resp.
// (refined x, refined y)); the thread walks the line and keeps the four
// coarse values its Hermite stencil needs in registers.
__device__ __forceinline__ double plane_ghost(const KP &p, const double *f, int kc, long long off2, int sS, int sN) {
    // vyxzc / vzxyc of InterpVelMgrd.F90:99-106 and :185-192: kc = -1..nxm+2
    if (kc >= 1 && kc <= p.nxm) return f[off2 + (kc - 1)];
    if (kc == 0) return sS * f[off2 + 0];
    if (kc == p.nx) return sN * f[off2 + (p.nxm - 1)];
    return 0.0;
}

__global__ void __launch_bounds__(64) k_vyr(KP p) {
    const int kr = blockIdx.x * 64 + threadIdx.x;   // 0..nxr-1
    const int ir = blockIdx.y;
    if (kr >= p.nxr) return;
    double *out = p.vyr + ridx(p, kr, 0, ir);
    const long long sy = p.ldxr;
    if (kr == p.nxmr) {                              // vyr(nxr) is never written: stays zero
        for (int jr = 0; jr < p.nyr; ++jr) out[jr * sy] = 0.0;
        return;
    }
    const int xs = p.xs_src[kr], is = p.zs_src[ir];
    double wx[4], wz[4];
    int ii[4];
    for (int a = 0; a < 4; ++a) { wx[a] = p.xs_w[4 * kr + a]; wz[a] = p.zs_w[4 * ir + a]; ii[a] = wrapn(is - 1 + a, p.nz); }
    const int sS = 1 - 2 * p.inslws, sN = 1 - 2 * p.inslwn;
    // first guess at jr = 0 from the coarse plane j = 0 (czvy = czrs, cxvy = cxrs)
    double v = 0.0;
    for (int a = 0; a < 4; ++a) {
        const int kc = xs + a;
        double q1 = 0.0;
        for (int c = 0; c < 4; ++c) {
            double f = plane_ghost(p, p.vy, kc, (long long)ii[c] * p.plane, sS, sN);
            q1 = c == 0 ? f * wz[0] : q1 + f * wz[c];
        }
        v = a == 0 ? q1 * wx[0] : v + q1 * wx[a];
    }
    out[0] = v;
    // coarse window: W[b] = sum_a wx[a] sum_c wz[c] tpdv(xs+a, jc, is-1+c)
    auto coarse = [&](int jc) -> double {
        const int jw = wrapn(jc, p.ny);
        double s = 0.0;
        for (int a = 0; a < 4; ++a) {
            const int kc = xs + a;
            double q = tpdv_at(p, kc, jw, ii[0]) * wz[0] + tpdv_at(p, kc, jw, ii[1]) * wz[1]
                     + tpdv_at(p, kc, jw, ii[2]) * wz[2] + tpdv_at(p, kc, jw, ii[3]) * wz[3];
            s = a == 0 ? q * wx[0] : s + q * wx[a];
        }
        return s;
    };
    int cur = p.ys_src[0] - 1;                    // coarse j of W[0]
    double W[4];
    for (int b = 0; b < 4; ++b) W[b] = coarse(cur + b);
    const double lya = 1.0 / p.dyr;
    for (int jr = 0; jr < p.nyr - 1; ++jr) {
        const int want = p.ys_src[jr] - 1;
        while (cur < want) { W[0] = W[1]; W[1] = W[2]; W[2] = W[3]; W[3] = coarse(cur + 4); ++cur; }
        const double *wy = p.ys_w + 4 * jr;
        const double t = W[0] * wy[0] + W[1] * wy[1] + W[2] * wy[2] + W[3] * wy[3];
        v = v + t * lya;
        out[(jr + 1) * sy] = v;
    }
}

__global__ void __launch_bounds__(64) k_vzr(KP p) {
    const int kr = blockIdx.x * 64 + threadIdx.x;
    const int jr = blockIdx.y;
    if (kr >= p.nxr) return;
    double *out = p.vzr + ridx(p, kr, jr, 0);
    const long long sz = p.planer;
    if (kr == p.nxmr) {
        for (int ir = 0; ir < p.nzr; ++ir) out[ir * sz] = 0.0;
        return;
    }
    const int xs = p.xs_src[kr], js = p.ys_src[jr];
    double wx[4], wy[4];
    int jj[4];
    for (int a = 0; a < 4; ++a) { wx[a] = p.xs_w[4 * kr + a]; wy[a] = p.ys_w[4 * jr + a]; jj[a] = wrapn(js - 1 + a, p.ny); }
    const int sS = 1 - 2 * p.inslws, sN = 1 - 2 * p.inslwn;
    double v = 0.0;
    if (p.nz > 1) {   // InterpVelMgrd.F90:184 guard ic0 < xend(3)
        for (int a = 0; a < 4; ++a) {
            const int kc = xs + a;
            double q1 = 0.0;
            for (int b = 0; b < 4; ++b) {
                double f = plane_ghost(p, p.vz, kc, (long long)jj[b] * p.ldx, sS, sN);
                q1 = b == 0 ? f * wy[0] : q1 + f * wy[b];
            }
            v = a == 0 ? q1 * wx[0] : v + q1 * wx[a];
        }
    }
    out[0] = v;
    auto coarse = [&](int ic) -> double {
        const int iw = wrapn(ic, p.nz);
        double s = 0.0;
        for (int a = 0; a < 4; ++a) {
            const int kc = xs + a;
            double q = tpdv_at(p, kc, jj[0], iw) * wy[0] + tpdv_at(p, kc, jj[1], iw) * wy[1]
                     + tpdv_at(p, kc, jj[2], iw) * wy[2] + tpdv_at(p, kc, jj[3], iw) * wy[3];
            s = a == 0 ? q * wx[0] : s + q * wx[a];
        }
        return s;
    };
    int cur = p.zs_src[0] - 1;
    double W[4];
    for (int c = 0; c < 4; ++c) W[c] = coarse(cur + c);
    const double lza = 1.0 / p.dzr;
    for (int ir = 0; ir < p.nzr - 1; ++ir) {
        const int want = p.zs_src[ir] - 1;
        while (cur < want) { W[0] = W[1]; W[1] = W[2]; W[2] = W[3]; W[3] = coarse(cur + 4); ++cur; }
        const double *wz = p.zs_w + 4 * ir;
        const double t = W[0] * wz[0] + W[1] * wz[1] + W[2] * wz[2] + W[3] * wz[3];
        v = v + t * lza;
        out[(long long)(ir + 1) * sz] = v;
    }
}

// InterpSalMultigrid (salinity.F90:458-491) through the fast coarse-grid loop
// of interpolate_xyz_to_coarse_fast (HermiteInterpolations.F90:291-310).
__global__ void __launch_bounds__(256) k_interp_sal(KP p) {
    const int k = blockIdx.x * 64 + threadIdx.x;   // 0..nx-1 (salc(nx) = 0)
    const int j = blockIdx.y * 4 + threadIdx.y;
    const int i = blockIdx.z;
    if (k >= p.nx || j >= p.ny) return;
    if (k == p.nxm) { p.salc[bidx(p, k, j, i)] = 0.0; return; }
    const int kr0 = p.xb_src[k], jr0 = p.yb_src[j], ir0 = p.zb_src[i];   // reference indices of first point
    const double *wx = p.xb_w + 4 * k, *wy = p.yb_w + 4 * j, *wz = p.zb_w + 4 * i;
    int jj[4], ii[4];
    for (int b = 0; b < 4; ++b) { jj[b] = wrapn(jr0 - 1 + b, p.nyr); ii[b] = wrapn(ir0 - 1 + b, p.nzr); }
    double res = 0.0;
    for (int a = 0; a < 4; ++a) {
        const int kr = kr0 + a;   // tpdvr index -1..nxr+1
        double q1 = 0.0;
        for (int b = 0; b < 4; ++b) {
            double q2 = 0.0;
            for (int c = 0; c < 4; ++c) {
                const long long o2 = (long long)ii[c] * p.planer + (long long)jj[b] * p.ldxr;
                double f;
                if (kr >= 1 && kr <= p.nxmr) f = p.sal[o2 + kr - 1];
                else if (kr == 0) {
                    const double s1 = p.sal[o2];
                    f = p.sfixs == 1 ? 2.0 * p.salbp[(long long)ii[c] * p.nyr + jj[b]] - s1 : s1;
                } else if (kr == p.nxr) {
                    const double sn = p.sal[o2 + p.nxmr - 1];
                    f = p.sfixn == 1 ? 2.0 * p.saltp[(long long)ii[c] * p.nyr + jj[b]] - sn : sn;
                } else f = 0.0;
                q2 = c == 0 ? f * wz[0] : q2 + f * wz[c];
            }
            q1 = b == 0 ? q2 * wy[0] : q1 + q2 * wy[b];
        }
        res = a == 0 ? q1 * wx[0] : res + q1 * wx[a];
    }
    p.salc[bidx(p, k, j, i)] = res;
}

// ---------------------------------------------------------------------------
// seeded initial condition (CreateInitialConditions.F90:175-201 RBC branch)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double splitmix_uniform(unsigned long long seed, unsigned long long idx) {
    unsigned long long z = (idx ^ seed) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return double(z >> 11) * (1.0 / 9007199254740992.0);
}
// scalar = bottom + (top - bottom) x/alx3 + noise, noise amplitude clamped near
// the walls as the reference does (CreateInitialConditions.F90:187-201); the
// noise of point (k,j,i) is U[0,1) of splitmix64(seed ^ linear index) over the
// (nxf, ny, nz) array, x fastest.
__global__ void __launch_bounds__(256) k_init_scalar(double *f, const double *bot, const double *top, const double *xm,
                                                     int nxm, int nxf, int ldx, int ny, int nz, double alx3,
                                                     double eps, unsigned long long seed) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= nxf || j >= ny) return;
    const long long o = ((long long)i * ny + j) * ldx + k;
    if (k >= nxm) { f[o] = 0.0; return; }
    const long long p2 = (long long)i * ny + j;
    const double x = xm[k];
    double v = bot[p2] + (top[p2] - bot[p2]) * x / alx3;
    const double r = splitmix_uniform(seed, (unsigned long long)(((long long)i * ny + j) * nxf + k));
    const double amp = (fabs(x - 0.5) + eps > 0.5) ? 0.5 - fabs(x - 0.5) : eps;
    f[o] = v + amp * (2.0 * r - 1.0);
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_max(double v) {
    __shared__ double sm[8];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    const int t = threadIdx.x + threadIdx.y * blockDim.x, w = t >> 6;
    if ((t & 63) == 0) sm[w] = v;
    __syncthreads();
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    if (t == 0) { double m = sm[0]; for (int q = 1; q < nw; ++q) m = fmax(m, sm[q]); v = m; }
    return v;
}
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double ss[8];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int t = threadIdx.x + threadIdx.y * blockDim.x, w = t >> 6;
    if ((t & 63) == 0) ss[w] = v;
    __syncthreads();
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    if (t == 0) { double m = ss[0]; for (int q = 1; q < nw; ++q) m += ss[q]; v = m; }
    __syncthreads();
    return v;
}

// CalcMaxCFL.F90:24-44 (what=0) / CheckDivergence.F90:24-43 (what=1) on the
// base grid, slot 0; refined grid versions write slot 1.
__global__ void __launch_bounds__(256) k_cfl_div(KP p, int what) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    double v = 0.0;
    if (k < p.nxm && j < p.ny) {
        const int jp = wrapp(j, p.ny), ip = wrapp(i, p.nz);
        if (what == 0)
            v = fabs((p.vz[bidx(p, k, j, i)] + p.vz[bidx(p, k, j, ip)]) * 0.5 * p.dz)
              + fabs((p.vy[bidx(p, k, j, i)] + p.vy[bidx(p, k, jp, i)]) * 0.5 * p.dy)
              + fabs((p.vx[bidx(p, k, j, i)] + p.vx[bidx(p, k + 1, j, i)]) * 0.5 * p.udx3m[k]);
        else
            v = fabs((p.vz[bidx(p, k, j, ip)] - p.vz[bidx(p, k, j, i)]) * p.dz
                   + (p.vy[bidx(p, k, jp, i)] - p.vy[bidx(p, k, j, i)]) * p.dy
                   + (p.vx[bidx(p, k + 1, j, i)] - p.vx[bidx(p, k, j, i)]) * p.udx3m[k]);
    }
    v = block_max(v);
    if (threadIdx.x == 0 && threadIdx.y == 0) atomicMax(&p.redmax[2 * what], (unsigned long long)__double_as_longlong(v));
}
__global__ void __launch_bounds__(256) k_cfl_div_r(KP p, int what) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    double v = 0.0;
    if (k < p.nxmr && j < p.nyr) {
        const int jp = wrapp(j, p.nyr), ip = wrapp(i, p.nzr);
        if (what == 0)
            v = fabs((p.vzr[ridx(p, k, j, i)] + p.vzr[ridx(p, k, j, ip)]) * 0.5 * p.dzr)
              + fabs((p.vyr[ridx(p, k, j, i)] + p.vyr[ridx(p, k, jp, i)]) * 0.5 * p.dyr)
              + fabs((p.vxr[ridx(p, k, j, i)] + p.vxr[ridx(p, k + 1, j, i)]) * 0.5 * p.udx3mr[k]);
        else
            v = fabs((p.vzr[ridx(p, k, j, ip)] - p.vzr[ridx(p, k, j, i)]) * p.dzr
                   + (p.vyr[ridx(p, k, jp, i)] - p.vyr[ridx(p, k, j, i)]) * p.dyr
                   + (p.vxr[ridx(p, k + 1, j, i)] - p.vxr[ridx(p, k, j, i)]) * p.udx3mr[k]);
    }
    v = block_max(v);
    if (threadIdx.x == 0 && threadIdx.y == 0) atomicMax(&p.redmax[2 * what + 1], (unsigned long long)__double_as_longlong(v));
}

// GlobalQuantities.F90:39-60 partial sums per block: {T vxcen g3rc, T g3rc, (u^2+v^2+w^2) g3rc}
__global__ void __launch_bounds__(256) k_global_sums(KP p) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    double a = 0.0, b = 0.0, c = 0.0;
    if (k < p.nxm && j < p.ny) {
        const long long o = bidx(p, k, j, i);
        const double fac2 = p.g3rc[k];
        const double vxcen = (p.vx[o] + p.vx[o + 1]) * 0.5, t = p.temp[o];
        a = t * vxcen * fac2;
        b = t * fac2;
        c = fac2 * (p.vz[o] * p.vz[o] + p.vy[o] * p.vy[o] + p.vx[o] * p.vx[o]);
    }
    a = block_sum(a); b = block_sum(b); c = block_sum(c);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const long long blk = ((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        p.redsum[3 * blk + 0] = a; p.redsum[3 * blk + 1] = b; p.redsum[3 * blk + 2] = c;
    }
}
__global__ void __launch_bounds__(256) k_sum_partials(const double *part, long long nblk, double *out) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (long long q = threadIdx.x; q < nblk; q += blockDim.x) { a += part[3 * q]; b += part[3 * q + 1]; c += part[3 * q + 2]; }
    a = block_sum(a); b = block_sum(b); c = block_sum(c);
    if (threadIdx.x == 0) { out[0] = a; out[1] = b; out[2] = c; }
}

}  // namespace afid
