// afid_ibm.hpp -- the immersed-boundary forms of the wall-normal solves
// (src/ibm/SolveImpEqnUpdate_X_ibm.F90, _YZ_ibm, _Temp_ibm, _Sal_ibm).
//
// A mask per solve marks each row: 2 liquid (the plain normalised row),
// 0 solid (identity with right-hand side -q, + Tb for T), 1 / -1 upper /
// lower boundary point (row (1, -d) / (-d, 1) with right-hand side
// d q(k+1) - q(k) / d q(k-1) - q(k), + (1 - d) Tb for T).  The reference
// takes d from a list in loop order; here it is a per-cell array the host
// expanded from that list.  The rows differ from column to column, so the
// solve is the Moebius-scan form of the masked phase-field solve: one
// wavefront per column, CH rows per lane.
#pragma once
#include <hip/hip_runtime.h>
#include "afid_kernels.hpp"

namespace afid {

// slot: 0 vx (rows 1 and nx identity with rhs -vx, update 2..nxm, q(k+1)
// unclamped), 1 / 2 vy / vz, 3 T (solid value tsolid), 4 S (solid rows as
// liquid, d = 1).  Columns of length n (rows used nxm or nxmr), layout
// [nz][ny][ld] as the field.
struct IbmJob {
    const signed char *mask;
    const double *dist;
    double tsolid;
    int slot, nrow, ld, ny, nz;
    long long plane;
};

template <int CH>
__global__ void __launch_bounds__(256) k_tri_update_ibm(TriFactors f, IbmJob jb, const double *rhs, double *q) {
    const long long col = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (col >= (long long)jb.ny * jb.nz) return;
    const int lane = threadIdx.x & 63;
    const int j = int(col % jb.ny), i = int(col / jb.ny);
    const long long base = (long long)i * jb.plane + (long long)j * jb.ld;
    const int nm = jb.nrow;                           // rows 1..nm carry a mask (nxm, nxmr)
    const int n = jb.slot == 0 ? nm + 1 : nm;         // rows of the solve
    const double tb = jb.slot == 3 ? jb.tsolid : 0.0;
    double v[CH], a[CH], c[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;                  // 0-based row, kc = r + 1
        v[m] = 0.0; a[m] = 0.0; c[m] = 0.0;
        if (r >= n) continue;
        if (jb.slot == 0 && (r == 0 || r == nm)) {    // SolveImpEqnUpdate_X_ibm.F90:66-70
            v[m] = -q[base + r];
            continue;
        }
        const int mk = jb.mask[base + r];
        const bool liquid = jb.slot == 4 ? (mk % 2 == 0) : (mk == 2);
        if (liquid) {
            v[m] = rhs[base + r] * f.s[r]; a[m] = f.a[r]; c[m] = f.c[r];
        } else if (mk == 0) {
            v[m] = -q[base + r] + tb;
        } else if (mk == 1 || mk == -1) {
            const double d = jb.slot == 4 ? 1.0 : jb.dist[base + r];
            int nb;
            if (mk == 1) nb = jb.slot == 0 ? r + 1 : min(r + 1, nm - 1);
            else nb = max(r - 1, 0);
            v[m] = d * q[base + nb] - q[base + r] + (1.0 - d) * tb;
            if (mk == 1) c[m] = -d; else a[m] = -d;
        } else {
            v[m] = 0.0;                               // not a mask value: identity with zero increment
        }
        if (r == 0) a[m] = 0.0;                       // dgttrf reads sub-diagonals of rows 2..n only
        if (r == n - 1) c[m] = 0.0;                   // and super-diagonals of rows 1..n-1
    }
    wave_tri_solve_general<CH>(v, a, c);
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;
        if (jb.slot == 0 ? (r >= 1 && r < nm) : (r < nm)) q[base + r] += v[m];
    }
}

}  // namespace afid
