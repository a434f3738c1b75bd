// afid_restart.hpp -- restart onto a new grid (ReadFlowInterp.F90:90-141,
// src/multires/IC_interpolation/InterpInputVel.F90, InterpInputSal.F90,
// InterpInputPhi.F90): fields of a checkpoint on another grid are moved to
// this context's grids by the 4-point Hermite transfer of the multires code.
//
// Old fields arrive compact, [nzo][nyo][nxo] with x fastest (the checkpoint
// layout).  A staging kernel builds the old array extended in x by its wall
// ghosts, [nzo][nyo][nxo+3] for x index -1..nxo+1 (a gradient of the old
// velocity for u, the scalar itself for T, S, phi); the transfer kernel then
// gathers, for every new cell centre, the 4x4x4 old block of the old cell
// that owns it and contracts it axis by axis (z, y, x: the reference's
// order).  The velocity components are integrated from the transferred
// gradients (vx from the wall, vy / vz from a first-guess plane), so a
// divergence-free old velocity stays divergence-free.  This runs once, at
// start-up: plain one-thread-per-point kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace afid {

// per-axis old -> new tables: for new index n (0-based), the first old
// source (1-based reference index: x from -1, periodic y/z from -1) and its
// 4 weights
struct RestartTabs {
    const int *xs, *ys, *zs;
    const double *xw, *yw, *zw;
};

// x ghost of the extended old array: (1 - 2 inslw) times the first / last
// cell for a velocity gradient or plane (InterpInputVel.F90:60-61, :112-113),
// 2 wall - value (Dirichlet) or value (zero gradient) for a scalar
struct RestartGhost {
    int kind;                   // 0: velocity (factor), 1: scalar
    double lo_fac, up_fac;      // velocity: 1 - 2 inslwS / N
    int fix_lo, fix_up;         // scalar: Dirichlet (1) or zero gradient (0)
    double lo_val, up_val;      // scalar: wall values
};

// old value of component comp at reference x index kc (1..nxmo) used by the
// gradient stage: comp 0 d(vx)/dx at the old centres, 1 d(vy)/dy, 2 d(vz)/dz
// (InterpInputVel.F90:56-58, :109-110, :201-202); comp 3 the value itself
__device__ __forceinline__ double old_value(const double *f, int comp, int kc, int j, int i, int nxo, int nyo, int nzo,
                                            const double *xco, double dyo, double dzo) {
    const long long row = ((long long)i * nyo + j) * nxo;
    if (comp == 0) return (f[row + kc] - f[row + kc - 1]) / (xco[kc] - xco[kc - 1]);
    if (comp == 1) {
        const long long rp = ((long long)i * nyo + (j + 1 == nyo ? 0 : j + 1)) * nxo;
        return (f[rp + kc - 1] - f[row + kc - 1]) * dyo;
    }
    if (comp == 2) {
        const long long rp = ((long long)(i + 1 == nzo ? 0 : i + 1) * nyo + j) * nxo;
        return (f[rp + kc - 1] - f[row + kc - 1]) * dzo;
    }
    return f[row + kc - 1];
}

// extended old array ext[i][j][kc+1], kc = -1..nxo+1 (xco: old faces, 0-based)
__global__ void __launch_bounds__(256) k_restart_stage(const double *f, int comp, int nxo, int nyo, int nzo,
                                                       const double *xco, double dyo, double dzo, RestartGhost gh,
                                                       double *ext) {
    const int q = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    const int le = nxo + 3, nxmo = nxo - 1;
    if (q >= le) return;
    const int kc = q - 1;
    double v = 0.0;
    if (kc >= 1 && kc <= nxmo) v = old_value(f, comp, kc, j, i, nxo, nyo, nzo, xco, dyo, dzo);
    if (kc == 0 || kc == nxo) {
        const int ki = kc == 0 ? 1 : nxmo;
        const double inner = old_value(f, comp, ki, j, i, nxo, nyo, nzo, xco, dyo, dzo);
        if (gh.kind == 0) v = (kc == 0 ? gh.lo_fac : gh.up_fac) * inner;
        else if (kc == 0) v = gh.fix_lo == 1 ? 2.0 * gh.lo_val - inner : inner;
        else v = gh.fix_up == 1 ? 2.0 * gh.up_val - inner : inner;
    }
    ext[((long long)i * nyo + j) * le + q] = v;
}

__device__ __forceinline__ int wrap_old(int j1, int n) {      // 1-based periodic index -> 0-based
    int r = (j1 - 1) % n;
    return r < 0 ? r + n : r;
}

// interpolate_xyz_old_to_new(_ref), HermiteInterpolations.F90:314-380, one
// thread per new cell centre (k, j, i): out[i][j][k] of stride ldn / planen
__global__ void __launch_bounds__(256) k_restart_transfer(const double *ext, int nxo, int nyo, int nzo, RestartTabs t,
                                                          double *out, int nxn, int nyn, int nzn, int ldn,
                                                          long long planen) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= nxn || j >= nyn || i >= nzn) return;
    const int le = nxo + 3;
    const int x0 = t.xs[k] + 1;           // offset of the first source in an extended row
    const double *wx = t.xw + 4 * k, *wy = t.yw + 4 * j, *wz = t.zw + 4 * i;
    double q1[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < 4; ++b) {
        const int jo = wrap_old(t.ys[j] + b, nyo);
        double q2[4];
        for (int a = 0; a < 4; ++a) {
            double s = 0.0;
            for (int c = 0; c < 4; ++c) {
                const int io = wrap_old(t.zs[i] + c, nzo);
                const double v = ext[((long long)io * nyo + jo) * le + x0 + a];
                s = c == 0 ? v * wz[0] : s + v * wz[c];
            }
            q2[a] = s;
        }
        for (int a = 0; a < 4; ++a) q1[a] = b == 0 ? q2[a] * wy[0] : q1[a] + q2[a] * wy[b];
    }
    out[(long long)i * planen + (long long)j * ldn + k] = q1[0] * wx[0] + q1[1] * wx[1] + q1[2] * wx[2] + q1[3] * wx[3];
}

// first-guess plane of vy (axis 1, y face 1 from the old plane j = 1,
// InterpInputVel.F90:125-156) or of vz (axis 2, z face 1 from i = 1,
// :217-248): 2-D Hermite in (x, z) or (x, y) of the old plane with its wall
// ghosts (1 - 2 inslw) v.  Thread per new (k, m), m the other periodic axis.
__global__ void __launch_bounds__(256) k_restart_plane(const double *f, int axis, int nxo, int nyo, int nzo,
                                                       RestartTabs t, RestartGhost gh, double *out, int nxn, int nmn,
                                                       int ldn, long long planen) {
    const int k = blockIdx.x * 64 + threadIdx.x, m = blockIdx.y * 4 + threadIdx.y;
    if (k >= nxn || m >= nmn) return;
    const int nxmo = nxo - 1;
    const int *ms = axis == 1 ? t.zs : t.ys;
    const double *wm = (axis == 1 ? t.zw : t.yw) + 4 * m, *wx = t.xw + 4 * k;
    const int nmo = axis == 1 ? nzo : nyo;
    double q1[4];
    for (int a = 0; a < 4; ++a) {
        const int kc = t.xs[k] + a;                   // reference x index -1..nxo+1
        double s = 0.0;
        for (int c = 0; c < 4; ++c) {
            const int mo = wrap_old(ms[m] + c, nmo);
            const long long row = axis == 1 ? (long long)mo * nyo * nxo : (long long)mo * nxo;   // j = 0 / i = 0
            double v = 0.0;
            if (kc >= 1 && kc <= nxmo) v = f[row + kc - 1];
            else if (kc == 0) v = gh.lo_fac * f[row];
            else if (kc == nxo) v = gh.up_fac * f[row + nxmo - 1];
            s = c == 0 ? v * wm[0] : s + v * wm[c];
        }
        q1[a] = s;
    }
    const double v = q1[0] * wx[0] + q1[1] * wx[1] + q1[2] * wx[2] + q1[3] * wx[3];
    out[axis == 1 ? (long long)m * planen + k : (long long)m * ldn + k] = v;
}

// integration of a transferred gradient g (new cell centres, same layout as
// the field): axis 0 vx(k+1) = vx(k) + g(k) lxa(k) from vx(1) = 0
// (InterpInputVel.F90:71-79); axis 1 vy(j+1) = vy(j) + g(j) / dy from the
// plane j = 1 (:161-168); axis 2 likewise along z (:253-261).  One thread per
// line; the value carried past the last periodic cell is dropped (halo).
__global__ void __launch_bounds__(256) k_restart_integrate(double *v, const double *g, int axis, int nxm, int ny, int nz,
                                                           int ldx, long long plane, const double *lxa, double step) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (axis == 0) {
        if (t >= ny * nz) return;
        const long long o = (long long)(t / ny) * plane + (long long)(t % ny) * ldx;
        double a = 0.0;
        v[o] = 0.0;
        for (int k = 0; k < nxm; ++k) {
            a = a + g[o + k] * lxa[k];
            v[o + k + 1] = a;
        }
    } else if (axis == 1) {
        if (t >= nxm * nz) return;
        const int k = t % nxm, i = t / nxm;
        const long long o = (long long)i * plane + k;
        for (int j = 1; j < ny; ++j) v[o + (long long)j * ldx] = v[o + (long long)(j - 1) * ldx] + g[o + (long long)(j - 1) * ldx] * step;
    } else {
        if (t >= nxm * ny) return;
        const int k = t % nxm, j = t / nxm;
        const long long o = (long long)j * ldx + k;
        for (int i = 1; i < nz; ++i) v[o + (long long)i * plane] = v[o + (long long)(i - 1) * plane] + g[o + (long long)(i - 1) * plane] * step;
    }
}

// FP64 staging of a refined field -> its FP32 storage (mixed form)
__global__ void __launch_bounds__(256) k_restart_to_float(const double *a, float *b, long long n) {
    const long long q = (long long)blockIdx.x * 256 + threadIdx.x;
    if (q < n) b[q] = float(a[q]);
}

}  // namespace afid
