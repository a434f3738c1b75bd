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

// Refinement gate of the mixed form: a launch of a refinement sweep whose
// gate slot holds a divergence at or below tol returns at once.  Slots hold
// the bit pattern of a non-negative double (ordered like the value).
struct Gate {
    const unsigned long long *slot;   // null: always run
    double tol;
};
__device__ __forceinline__ bool gate_closed(const Gate &g) {
    return g.slot && __longlong_as_double((long long)*g.slot) <= g.tol;
}

// block-wide reductions (result valid in thread 0)
__device__ __forceinline__ double block_max(double v) {
    __shared__ double sm[16];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    const int t = threadIdx.x + threadIdx.y * blockDim.x, w = t >> 6;
    if ((t & 63) == 0) sm[w] = v;
    __syncthreads();
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    if (t == 0) { double m = sm[0]; for (int q = 1; q < nw; ++q) m = fmax(m, sm[q]); v = m; }
    return v;
}
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double ss[16];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int t = threadIdx.x + threadIdx.y * blockDim.x, w = t >> 6;
    if ((t & 63) == 0) ss[w] = v;
    __syncthreads();
    const int nw = (blockDim.x * blockDim.y + 63) >> 6;
    if (t == 0) { double m = ss[0]; for (int q = 1; q < nw; ++q) m += ss[q]; v = m; }
    __syncthreads();
    return v;
}

// thread 0 of a workgroup folds its maximum into a slot; the atomic is
// skipped when the slot already holds at least that value (the plain read
// may be stale, which only costs an atomic), so a grid of many workgroups
// does not serialise on one address
__device__ __forceinline__ void max_to_slot(unsigned long long *slot, double v) {
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
        if (bits > *slot) atomicMax(slot, bits);
    }
}

// ---------------------------------------------------------------------------
// per-routine explicit kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_explicit_vx(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    if (k == 0) return;   // qcap(1) is not computed (ExplicitTermsVX.F90:50)
    const Nb n = make_nb(p, j, i);
    p.qcap[bidx(p, k, j, i)] = ex_vx<double>(p, n, k);
}
__global__ void __launch_bounds__(256) k_explicit_vy(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const Nb n = make_nb(p, j, i);
    p.dph[bidx(p, k, j, i)] = ex_vy<double>(p, n, k);
}
__global__ void __launch_bounds__(256) k_explicit_vz(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const Nb n = make_nb(p, j, i);
    p.dq[bidx(p, k, j, i)] = ex_vz<double>(p, n, k);
}
__global__ void __launch_bounds__(256) k_explicit_temp(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const Nb n = make_nb(p, j, i);
    p.hro[bidx(p, k, j, i)] = ex_temp<double>(p, n, k);
}
__global__ void __launch_bounds__(256) k_explicit_sal(KP p) {
    AFID_POINT_PROLOGUE(p.nxmr, p.nyr, p.nzr)
    const NbS<double> n = make_nbs<double>(p, j, i);
    p.hsal[ridx(p, k, j, i)] = ex_sal<double, double>(p, n, k);
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
    const Nb n = make_nb(p, j, i);
    p.rhsx[c] = rhs_vx(p, n, h, p.rux[c], k);
    p.rux[c] = h;
}
__global__ void __launch_bounds__(256) k_rhs_vy(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c = bidx(p, k, j, i);
    const double h = p.dph[c];
    const Nb n = make_nb(p, j, i);
    p.rhsy[c] = rhs_vy(p, n, h, p.ruy[c], k);
    p.ruy[c] = h;
}
__global__ void __launch_bounds__(256) k_rhs_vz(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c = bidx(p, k, j, i);
    const double h = p.dq[c];
    const Nb n = make_nb(p, j, i);
    p.rhsz[c] = rhs_vz(p, n, h, p.ruz[c], k);
    p.ruz[c] = h;
}
__global__ void __launch_bounds__(256) k_rhs_temp(KP p) {
    AFID_POINT_PROLOGUE(p.nxm, p.ny, p.nz)
    const long long c = bidx(p, k, j, i);
    const double h = p.hro[c];
    const Nb n = make_nb(p, j, i);
    p.rhst[c] = rhs_temp(p, n, h, p.rut[c], k);
    p.rut[c] = h;
}
__global__ void __launch_bounds__(256) k_rhs_sal(KP p) {
    AFID_POINT_PROLOGUE(p.nxmr, p.nyr, p.nzr)
    const long long c = ridx(p, k, j, i);
    const double h = p.hsal[c];
    const NbS<double> n = make_nbs<double>(p, j, i);
    p.rhsr[c] = rhs_sal(p, n, h, p.rusal[c], k);
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
    const double *l = f.l + lane * CH, *e = f.e + lane * CH, *g = f.g + lane * CH, *sc = f.s + lane * CH;
#pragma unroll
    for (int m = 0; m < CH; ++m) v[m] *= sc[m];
    // forward: y_r = v_r - l_r y_{r-1}
    double A = 1.0, B = 0.0;
#pragma unroll
    for (int m = 0; m < CH; ++m) { const double lm = l[m]; B = -lm * B + v[m]; A = -lm * A; }
    wave_affine_scan<true>(A, B);
    double y = __shfl_up(B, 1, 64);
    if (lane == 0) y = 0.0;
#pragma unroll
    for (int m = 0; m < CH; ++m) { y = -l[m] * y + v[m]; v[m] = y; }
    // backward: x_r = g_r y_r - e_r x_{r+1}
    A = 1.0; B = 0.0;
#pragma unroll
    for (int m = CH - 1; m >= 0; --m) { const double em = e[m]; B = -em * B + g[m] * v[m]; A = -em * A; }
    wave_affine_scan<false>(A, B);
    double x = __shfl_down(B, 1, 64);
    if (lane == 63) x = 0.0;
#pragma unroll
    for (int m = CH - 1; m >= 0; --m) { x = -e[m] * x + g[m] * v[m]; v[m] = x; }
}

// The same solve with the folded factors (TriFactors::ln, e): v in is the
// right-hand side already scaled by sn, v out the solution.
template <int CH>
__device__ __forceinline__ void wave_tri_solve_folded(double (&v)[CH], const double *ln, const double *e) {
    const int lane = threadIdx.x & 63;
    const double *l = ln + lane * CH, *ee = e + lane * CH;
    // forward: y'_r = v_r - ln_r y'_{r-1}
    double A = 1.0, B = 0.0;
#pragma unroll
    for (int m = 0; m < CH; ++m) { const double lm = l[m]; B = -lm * B + v[m]; A = -lm * A; }
    wave_affine_scan<true>(A, B);
    double y = __shfl_up(B, 1, 64);
    if (lane == 0) y = 0.0;
#pragma unroll
    for (int m = 0; m < CH; ++m) { y = -l[m] * y + v[m]; v[m] = y; }
    // backward: x_r = y'_r - e_r x_{r+1}
    A = 1.0; B = 0.0;
#pragma unroll
    for (int m = CH - 1; m >= 0; --m) { const double em = ee[m]; B = -em * B + v[m]; A = -em * A; }
    wave_affine_scan<false>(A, B);
    double x = __shfl_down(B, 1, 64);
    if (lane == 63) x = 0.0;
#pragma unroll
    for (int m = CH - 1; m >= 0; --m) { x = -ee[m] * x + v[m]; v[m] = x; }
}

// 2x2 projective matrix (normalised so its largest entry is 1)
struct Mob { double m00, m01, m10, m11; };
__device__ __forceinline__ Mob mob_mul(const Mob &x, const Mob &y) {   // x * y
    Mob r{x.m00 * y.m00 + x.m01 * y.m10, x.m00 * y.m01 + x.m01 * y.m11,
          x.m10 * y.m00 + x.m11 * y.m10, x.m10 * y.m01 + x.m11 * y.m11};
    const double mx = fmax(fmax(fabs(r.m00), fabs(r.m01)), fmax(fabs(r.m10), fabs(r.m11)));
    if (mx > 0.0) { const double q = 1.0 / mx; r.m00 *= q; r.m01 *= q; r.m10 *= q; r.m11 *= q; }
    return r;
}

// Solve a tridiagonal system whose rows differ from column to column (row
// r: a_r x_{r-1} + x_r + c_r x_{r+1} = v_r), CH rows per lane.  The Thomas
// recurrence c'_r = c_r / (1 - a_r c'_{r-1}) is a Moebius map of c'_{r-1};
// its prefix over the wavefront is a scan of 2x2 matrix products.  With c'
// known, forward elimination and back substitution are affine scans as in
// wave_tri_solve.
// c' of the row before this lane's first row: the Moebius prefix of the
// earlier lanes' rows applied to c' = 0
template <int CH>
__device__ __forceinline__ double wave_mobius_prefix(const double (&a)[CH], const double (&c)[CH]) {
    const int lane = threadIdx.x & 63;
    Mob L{1.0, 0.0, 0.0, 1.0};
#pragma unroll
    for (int m = 0; m < CH; ++m) L = mob_mul(Mob{0.0, c[m], -a[m], 1.0}, L);
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        Mob P{__shfl_up(L.m00, off, 64), __shfl_up(L.m01, off, 64), __shfl_up(L.m10, off, 64), __shfl_up(L.m11, off, 64)};
        if (lane >= off) L = mob_mul(L, P);
    }
    // the exclusive prefix applied to (0, 1)
    const double p0 = __shfl_up(L.m01, 1, 64), q0 = __shfl_up(L.m11, 1, 64);
    return lane == 0 ? 0.0 : p0 / q0;
}

template <int CH>
__device__ __forceinline__ void wave_tri_solve_general(double (&v)[CH], const double (&a)[CH], const double (&c)[CH]) {
    const int lane = threadIdx.x & 63;
    double t = wave_mobius_prefix<CH>(a, c);
    double cp[CH], g[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) { g[m] = 1.0 / (1.0 - a[m] * t); cp[m] = c[m] * g[m]; t = cp[m]; }
    // forward: y_r = g_r v_r - g_r a_r y_{r-1}
    double A = 1.0, B = 0.0;
#pragma unroll
    for (int m = 0; m < CH; ++m) { const double am = -g[m] * a[m]; B = am * B + g[m] * v[m]; A = am * A; }
    wave_affine_scan<true>(A, B);
    double y = __shfl_up(B, 1, 64);
    if (lane == 0) y = 0.0;
#pragma unroll
    for (int m = 0; m < CH; ++m) { y = -g[m] * a[m] * y + g[m] * v[m]; v[m] = y; }
    // backward: x_r = y_r - c'_r x_{r+1}
    A = 1.0; B = 0.0;
#pragma unroll
    for (int m = CH - 1; m >= 0; --m) { B = -cp[m] * B + v[m]; A = -cp[m] * A; }
    wave_affine_scan<false>(A, B);
    double x = __shfl_down(B, 1, 64);
    if (lane == 63) x = 0.0;
#pragma unroll
    for (int m = CH - 1; m >= 0; --m) { x = -cp[m] * x + v[m]; v[m] = x; }
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
// fused column kernels: explicit terms + implicit RHS + tridiagonal solve +
// update of one wall-normal column per wavefront.  Stencil phase: lane l
// handles rows l, l+64, ... (coalesced loads of the column and its y/z
// neighbours); the RHS goes through LDS to the chunked layout of the scan
// solver and back.  Fields are read from the pre-substep arrays and the
// updated column is written to a second buffer (other wavefronts still read
// the old values as neighbours).
// ---------------------------------------------------------------------------
// Consecutive workgroups are dealt round-robin to the 8 XCDs; give each XCD
// a contiguous range of columns so y/z neighbours share its L2.
// The grid has a multiple of 8 workgroups (padded by the host); the map is a
// bijection of [0, gridDim.x) and indices >= the real count are idle.
__device__ __forceinline__ long long xcd_block(long long) {
    const long long b = blockIdx.x;
    const long long per = gridDim.x / 8;
    return (b % 8) * per + b / 8;
}
// Tile grids: a 1-D grid over the tiles (tile = tx + nx_tiles * (ty +
// ny_tiles * tz)), padded to a multiple of 8 workgroups.  Through
// xcd_block each XCD takes a contiguous range of tiles, so neighbouring
// tiles, which read the same halo cache lines, run on one L2 at about the
// same time.  False for the padding workgroups.
__host__ __device__ __forceinline__ unsigned tile_grid(long long ntiles) { return unsigned((ntiles + 7) / 8 * 8); }
// The vxr tiles (ny_tiles x nz_tiles) in groups of TILE_GY y tiles walked
// z-fastest: y-neighbour tiles are 1 apart and z-neighbour tiles, which
// share 3 of the 5 coarse planes of their block, TILE_GY apart, so both are
// in flight on one XCD together.
constexpr int TILE_GY = 4;
__device__ __forceinline__ bool tile_yz_grouped(int nyt, int nzt, int &ty, int &tz) {
    const long long t = xcd_block(0);
    if (t >= (long long)nyt * nzt) return false;
    const long long per = (long long)TILE_GY * nzt;
    const int g = int(t / per), r = int(t - g * per);
    const int w = min(TILE_GY, nyt - g * TILE_GY);      // the last group may be narrower
    ty = g * TILE_GY + r % w;
    tz = r / w;
    return true;
}
__device__ __forceinline__ bool tile_3d(int nxt, int nyt, int nzt, int &tx, int &ty, int &tz) {
    const long long t = xcd_block(0);
    if (t >= (long long)nxt * nyt * nzt) return false;
    tx = int(t % nxt);
    const long long r = t / nxt;
    ty = int(r % nyt);
    tz = int(r / nyt);
    return true;
}
// Column walk: a workgroup owns 4 consecutive y columns, one per
// wavefront; workgroups form strips of STRIP_BLOCKS*4 columns in y, walked
// along z, so the y and z neighbour columns a wavefront reads belong to its
// own workgroup or to workgroups that ran just before it on the same XCD
// (still in its L2).
constexpr int STRIP_BLOCKS = 4;
__host__ __device__ __forceinline__ long long strip_blocks(int ny, int nz) {
    return (long long)((ny + 4 * STRIP_BLOCKS - 1) / (4 * STRIP_BLOCKS)) * STRIP_BLOCKS * nz;
}
// (j, i) of this wavefront's column; false if the slot is idle
__device__ __forceinline__ bool strip_column(long long blk, int wave, int ny, int nz, int &j, int &i) {
    const long long per_strip = (long long)STRIP_BLOCKS * nz;
    const long long strip = blk / per_strip, rem = blk % per_strip;
    i = int(rem / STRIP_BLOCKS);
    j = int(strip * 4 * STRIP_BLOCKS + (rem % STRIP_BLOCKS) * 4 + wave);
    const bool ok = blk < strip_blocks(ny, nz) && j < ny && i < nz;
    if (!ok) { j = 0; i = 0; }
    return ok;
}

// Row r of a column lives at lds_row<CH>(r): chunks of CH rows at an odd
// stride of CHP >= CH+1 doubles, so the solver's lane-strided 64-bit reads
// (lane * CHP) hit distinct bank pairs.
template <int CH>
constexpr int lds_chunk() { return (CH + 1) % 2 ? CH + 1 : CH + 2; }
// Tight form: CH itself when odd (lane strides of 2*CH dwords reach 16
// distinct bank pairs per 16 lanes), else CH+1.
template <int CH>
constexpr int lds_chunk_tight() { return CH % 2 ? CH : CH + 1; }
template <int CH, int CHP = lds_chunk<CH>()>
__device__ __forceinline__ unsigned lds_row(unsigned r) { return (r / CH) * CHP + r % CH; }
template <int CH, int CHP = lds_chunk<CH>()>
constexpr int lds_col_size() { return 64 * CHP; }

template <int CH, int CHP = lds_chunk<CH>()>
__device__ __forceinline__ void column_solve_lds(double *buf, const TriFactors &f) {
    const int lane = threadIdx.x & 63;
    double v[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) v[m] = buf[lane * CHP + m];
    wave_tri_solve<CH>(v, f);
#pragma unroll
    for (int m = 0; m < CH; ++m) buf[lane * CHP + m] = v[m];
}

// column_solve_lds with the folded factors (two tables, e.g. staged in LDS)
template <int CH, int CHP = lds_chunk<CH>()>
__device__ __forceinline__ void column_solve_lds_folded(double *buf, const double *ln, const double *e) {
    const int lane = threadIdx.x & 63;
    double v[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) v[m] = buf[lane * CHP + m];
    wave_tri_solve_folded<CH>(v, ln, e);
#pragma unroll
    for (int m = 0; m < CH; ++m) buf[lane * CHP + m] = v[m];
}

// two-element vector of the storage type (64-bit or 128-bit lane accesses)
// Masked solve of one column held in LDS (SolveImpEqnUpdate_X_pf /
// _YZ_pf, phasefield.F90:821-930): rows flagged in msk are the identity with
// the right-hand side already in buf; the others are the rows of f.
template <int CH>
__device__ __forceinline__ void column_solve_masked_lds(double *buf, const double *msk, const TriFactors &f) {
    const int lane = threadIdx.x & 63;
    double v[CH], a[CH], c[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;
        const bool solid = msk[lane * lds_chunk<CH>() + m] != 0.0;
        const double sv = f.s[r];
        v[m] = buf[lane * lds_chunk<CH>() + m];
        if (!solid) v[m] *= sv;
        a[m] = solid ? 0.0 : f.a[r];
        c[m] = solid ? 0.0 : f.c[r];
    }
    wave_tri_solve_general<CH>(v, a, c);
#pragma unroll
    for (int m = 0; m < CH; ++m) buf[lane * lds_chunk<CH>() + m] = v[m];
}

template <typename RT> struct Pair;
template <> struct Pair<float> { using T = float2; };
template <> struct Pair<double> { using T = double2; };
// 16-byte vector of the storage type: 4 floats or 2 doubles per lane access
template <typename RT> struct Vec16;
template <> struct Vec16<float> { using T = float4; static constexpr int n = 4; };
template <> struct Vec16<double> { using T = double2; static constexpr int n = 2; };
template <typename RT>
__device__ __forceinline__ RT vget(const typename Vec16<RT>::T &v, int m) { return reinterpret_cast<const RT *>(&v)[m]; }

// ExplicitSalinity + ImplicitSalinity + SolveImpEqnUpdate_Sal (salinity.F90:268-454).
// RT is the storage (and explicit-stencil arithmetic) type of the refined
// fields; the implicit RHS, the solve and the update run in FP64.
// Stencil phase: each lane takes SUB adjacent rows (r0 = SUB*lane + 64*SUB*m
// .. r0+SUB-1, 16 bytes) with one vector access per array, so a wavefront
// access moves 1 KiB; the x-neighbours outside the group come from two
// scalar loads that hit the same cache lines.  The RHS goes through LDS to
// the solver's layout (CH rows per lane).
template <int CH, typename RT>
__global__ void __launch_bounds__(256) k_column_sal(KP p, TriFactors f, RT *sal_out) {
    using V = typename Vec16<RT>::T;
    constexpr int SUB = Vec16<RT>::n;
    constexpr int MG = (64 * CH + 64 * SUB - 1) / (64 * SUB);   // group iterations covering the solver's 64*CH rows
    __shared__ double lds[4][lds_col_size<CH>()];
    // the solver's folded factors, read by every column of the workgroup
    // (two tables: 39 KB of LDS in all at CH = 12, four workgroups per CU)
    __shared__ double fac[2][64 * CH];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned lane = threadIdx.x & 63;
    for (int q = threadIdx.x; q < 64 * CH; q += 256) { fac[0][q] = f.ln[q]; fac[1][q] = f.e[q]; }
    int j, i;
    const bool live = strip_column(xcd_block(0), wave, p.nyr, p.nzr, j, i);
    double *buf = lds[wave];
    const NbS<RT> n = make_nbs<RT>(p, j, i);
    const long long c0 = (long long)i * p.planer + (long long)j * p.ldxr;
    RT *rus = sel<RT>(p.rusal, p.rusalf) + c0;
    const unsigned nxmr = p.nxmr;
    V own[MG];
#pragma unroll
    for (int m = 0; m < MG; ++m) {
        const unsigned r0 = SUB * lane + 64 * SUB * m;
        double rhs[SUB];
#pragma unroll
        for (int q = 0; q < SUB; ++q) { rhs[q] = 0.0; reinterpret_cast<RT *>(&own[m])[q] = RT(0); }
        if (live && r0 + SUB <= nxmr) {
            // r0 a multiple of SUB and every row 8-aligned: the group is one
            // aligned 16-byte access; r0 + SUB <= nxmr < ldxr stays in the row
            auto ld = [&](const RT *a) { return *reinterpret_cast<const V *>(a + r0); };
            const V sc = ld(n.s), sjm = ld(n.s_jm), sjp = ld(n.s_jp), sim = ld(n.s_im), sip = ld(n.s_ip);
            const V vx = ld(n.vxr), vy = ld(n.vyr), vyp = ld(n.vyr_jp), vz = ld(n.vzr), vzp = ld(n.vzr_ip), ru = ld(rus);
            const RT sm = r0 > 0 ? n.s[r0 - 1] : RT(0);
            const RT sn = r0 + SUB < nxmr ? n.s[r0 + SUB] : RT(0);
            const RT vn = n.vxr[r0 + SUB];
            V h;
#pragma unroll
            for (int q = 0; q < SUB; ++q) {
                const unsigned k = r0 + q;
                const RT a = q == 0 ? sm : vget<RT>(sc, q - 1), b = vget<RT>(sc, q);
                const RT c = q == SUB - 1 ? sn : vget<RT>(sc, q + 1);
                const RT x1 = q == SUB - 1 ? vn : vget<RT>(vx, q + 1);
                const RT hq = ex_sal_val<RT>(p, k, a, b, c, vget<RT>(sjm, q), vget<RT>(sjp, q), vget<RT>(sim, q),
                                             vget<RT>(sip, q), vget<RT>(vx, q), x1, vget<RT>(vy, q), vget<RT>(vyp, q),
                                             vget<RT>(vz, q), vget<RT>(vzp, q), n.sbp, n.stp);
                reinterpret_cast<RT *>(&h)[q] = hq;
                rhs[q] = rhs_sal_val(p, k, double(hq), double(vget<RT>(ru, q)), double(a), double(b), double(c), n.sbp, n.stp);
            }
            own[m] = sc;
            *reinterpret_cast<V *>(rus + r0) = h;
        } else if (live) {
            // the group holding row nxmr - 1 when nxmr is not a multiple of SUB
#pragma unroll
            for (int q = 0; q < SUB; ++q) {
                const unsigned k = r0 + q;
                if (k < nxmr) {
                    const RT hq = ex_sal<RT, RT>(p, n, k);
                    rhs[q] = rhs_sal<RT>(p, n, double(hq), double(rus[k]), k);
                    reinterpret_cast<RT *>(&own[m])[q] = n.s[k];
                    rus[k] = hq;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < SUB; ++q)
            if (r0 + q < 64u * CH) buf[lds_row<CH>(r0 + q)] = rhs[q] * f.sn[r0 + q];
    }
    __syncthreads();
    column_solve_lds_folded<CH>(buf, fac[0], fac[1]);
    __syncthreads();
    if (!live) return;
    RT *out = sal_out + c0;
#pragma unroll
    for (int m = 0; m < MG; ++m) {
        const unsigned r0 = SUB * lane + 64 * SUB * m;
        if (r0 >= nxmr) continue;
        V o;
#pragma unroll
        for (int q = 0; q < SUB; ++q)
            reinterpret_cast<RT *>(&o)[q] = RT(double(vget<RT>(own[m], q)) + (r0 + q < 64u * CH ? buf[lds_row<CH>(r0 + q)] : 0.0));
        if (r0 + SUB <= nxmr) *reinterpret_cast<V *>(out + r0) = o;
        else
#pragma unroll
            for (int q = 0; q < SUB; ++q)
                if (r0 + q < nxmr) out[r0 + q] = vget<RT>(o, q);
    }
    if (lane == 0) out[nxmr] = n.s[nxmr];       // the unused index nxr is carried over
}

// Explicit terms of vx, vy, vz, T (+ AddSalBuoyancy) and ImplicitAndUpdate
// VX/VY/VZ/Temp with their solves (TimeMarcher.F90:45-95) for one column.
struct BaseOut { double *vx, *vy, *vz, *temp; };

// stencil phase of field FLD: RHS of each row into buf, old value into own.
// HT is the storage and arithmetic type of the explicit terms (and of salc);
// the RHS and the solve run in FP64.  With PF (phase field) the velocity
// terms lose the volume penalty (AddVolumePenalty, phasefield.F90:525-566),
// the temperature term gains the latent heat qlat (AddLatentHeat, :651-725),
// and a velocity row whose averaged phic exceeds pf_direct_force = 0.9
// (phasefield.F90:29) is flagged in msk with right-hand side -u, the
// identity row of SolveImpEqnUpdate_X_pf / _YZ_pf (:821-930).
template <int FLD, int CH, typename HT, bool PF>
__device__ __forceinline__ void column_base_rows(const KP &p, const Nb &n, long long c0, double *buf, double *own,
                                                 double *msk, unsigned lane, unsigned nrow, bool live) {
    const HT *salc = sel<HT>(p.salc, p.salcf) + c0;
    HT *ru = (FLD == 0 ? sel<HT>(p.rux, p.ruxf) : FLD == 1 ? sel<HT>(p.ruy, p.ruyf)
              : FLD == 2 ? sel<HT>(p.ruz, p.ruzf) : sel<HT>(p.rut, p.rutf)) + c0;
    const unsigned nxm = p.nxm;
#pragma unroll 1
    for (unsigned r = lane; r < nrow; r += 64) {
        double rhs = 0.0, old = 0.0;
        bool solid = false;
        if (live && r < unsigned(p.nx)) {
            const HT sb = (p.active_s == 1 && r < nxm) ? HT(p.bycs) * salc[r] : HT(0);
            if (FLD == 0) {
                old = n.vx[r];
                if (r >= 1 && r < nxm) {
                    HT h;
                    double bref = 0.0;
                    if (sizeof(HT) == sizeof(float) && p.bref) {
                        // FP32 terms: the buoyancy enters as its deviation from
                        // the row's FP64 reference, formed in FP64 and rounded
                        // once; the RHS adds the reference back in FP64
                        bref = p.bref[r];
                        h = ex_vx<HT, false>(p, n, r) + HT(p.byct * (p.active_t * 0.5 * (n.t[r] + n.t[r - 1])) - bref);
                    } else {
                        h = ex_vx<HT>(p, n, r);
                    }
                    if (p.gaxis == 1) h -= sb;
                    if (PF) {
                        h -= HT((n.pc[r] + n.pc[r - 1]) * n.vx[r] * p.heta);
                        solid = 0.5 * (n.pc[r] + n.pc[r + 1]) > 0.9;
                    }
                    rhs = rhs_vx(p, n, double(h) + bref, double(ru[r]) + bref, r);
                    ru[r] = h;
                } else if (PF) {
                    solid = true;                    // rows 1 and nx: identity with rhs -vx
                }
            } else if (FLD == 1) {
                old = n.vy[r];
                if (r < nxm) {
                    HT h = ex_vy<HT>(p, n, r);
                    if (p.gaxis == 2) h -= sb;
                    if (PF) {
                        const double pa = n.pc[r] + n.pc_jm[r];
                        h -= HT(pa * n.vy[r] * p.heta);
                        solid = 0.5 * pa > 0.9;
                    }
                    rhs = rhs_vy(p, n, double(h), double(ru[r]), r);
                    ru[r] = h;
                }
            } else if (FLD == 2) {
                old = n.vz[r];
                if (r < nxm) {
                    HT h = ex_vz<HT>(p, n, r);
                    if (p.gaxis == 3) h -= sb;
                    if (PF) {
                        const double pa = n.pc[r] + n.pc_im[r];
                        h -= HT(pa * n.vz[r] * p.heta);
                        solid = 0.5 * pa > 0.9;
                    }
                    rhs = rhs_vz(p, n, double(h), double(ru[r]), r);
                    ru[r] = h;
                }
            } else {
                old = n.t[r];
                if (r < nxm) {
                    HT h = ex_temp<HT>(p, n, r);
                    if (PF) h += HT(p.qlat[c0 + r]);
                    if (p.moist) h = HT(double(h) + condensation_heat(p, c0, r));
                    rhs = rhs_temp(p, n, double(h), double(ru[r]), r);
                    ru[r] = h;
                }
            }
        }
        if (PF && FLD < 3) {
            if (solid) rhs = -old;
            msk[lds_row<CH>(r)] = solid ? 1.0 : 0.0;
        }
        buf[lds_row<CH>(r)] = rhs;
        own[lds_row<CH>(r)] = old;
    }
}

// Four passes, one field each (the phase-field form, and wall-normal sizes
// above 512 cells).  CH = ceil(nx / 64).
template <int CH, typename HT, bool PF>
__global__ void __launch_bounds__(256) k_column_base(KP p, TriFactors fx, TriFactors fyz, TriFactors ft, BaseOut o) {
    __shared__ double lds[PF ? 3 : 2][4][lds_col_size<CH>()];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned lane = threadIdx.x & 63;
    int j, i;
    const bool live = strip_column(xcd_block(0), wave, p.ny, p.nz, j, i);
    double *buf = lds[0][wave], *own = lds[1][wave], *msk = lds[PF ? 2 : 1][wave];
    const Nb n = make_nb(p, j, i);
    const long long c0 = (long long)i * p.plane + (long long)j * p.ldx;
    const unsigned nrow = 64 * CH, nx = p.nx, nxm = p.nxm;
#pragma unroll 1
    for (int fld = 0; fld < 4; ++fld) {
        switch (fld) {
        case 0: column_base_rows<0, CH, HT, PF>(p, n, c0, buf, own, msk, lane, nrow, live); break;
        case 1: column_base_rows<1, CH, HT, PF>(p, n, c0, buf, own, msk, lane, nrow, live); break;
        case 2: column_base_rows<2, CH, HT, PF>(p, n, c0, buf, own, msk, lane, nrow, live); break;
        default: column_base_rows<3, CH, HT, PF>(p, n, c0, buf, own, msk, lane, nrow, live); break;
        }
        __syncthreads();
        const TriFactors &f = fld == 0 ? fx : (fld == 3 ? ft : fyz);
        if (PF && fld < 3) column_solve_masked_lds<CH>(buf, msk, f);
        else column_solve_lds<CH>(buf, f);
        __syncthreads();
        if (live) {
            double *out = (fld == 0 ? o.vx : fld == 1 ? o.vy : fld == 2 ? o.vz : o.temp) + c0;
            for (unsigned r = lane; r < nx; r += 64) {
                const unsigned q = lds_row<CH>(r);
                double v = own[q];
                if (fld == 0) v = (r >= 1 && r < nxm) ? v + buf[q] : 0.0;   // walls, ImplicitAndUpdateVX.F90:86-87
                else if (r < nxm) v += buf[q];
                out[r] = v;
            }
        }
        __syncthreads();
    }
}

// Right-hand sides of vx, vy, vz and T at row r < nx of one column (the
// stencil phase of column_base_rows for the four fields at once, no phase
// field): every neighbour value is loaded once for the four explicit terms,
// and the explicit terms are stored to their history arrays after all
// loads.  Rows without an equation (vx rows 0 and nxm, the unused index
// nxm of the others) give 0.
template <typename HT, bool PF = false>
__device__ __forceinline__ void base_rhs4(const KP &p, const Nb &n, long long c0, unsigned r, double (&rhs)[4],
                                          unsigned &solid) {
    const unsigned nxm = p.nxm;
    HT *rux = sel<HT>(p.rux, p.ruxf) + c0, *ruy = sel<HT>(p.ruy, p.ruyf) + c0;
    HT *ruz = sel<HT>(p.ruz, p.ruzf) + c0, *rut = sel<HT>(p.rut, p.rutf) + c0;
    const bool wx = r >= 1 && r < nxm, wc = r < nxm;
    const HT sb = (p.active_s == 1 && wc) ? HT(p.bycs) * sel<HT>(p.salc, p.salcf)[c0 + r] : HT(0);
    HT h0 = HT(0), h1 = HT(0), h2 = HT(0), h3 = HT(0);
    double bref = 0.0, r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
    // phase field: solid rows of vx, vy, vz (bits 0..2); vx rows 0 and nxm
    // are identity rows (SolveImpEqnUpdate_X_pf)
    bool sx = PF && !wx, sy = false, sz = false;
    if (wx) {
        if (sizeof(HT) == sizeof(float) && p.bref) {
            // FP32 terms: buoyancy as its deviation from the row's FP64
            // reference, which the RHS adds back in FP64
            bref = p.bref[r];
            h0 = ex_vx<HT, false>(p, n, r) + HT(p.byct * (p.active_t * 0.5 * (n.t[r] + n.t[r - 1])) - bref);
        } else {
            h0 = ex_vx<HT>(p, n, r);
        }
        if (p.gaxis == 1) h0 -= sb;
        if (PF) {
            // AddVolumePenalty (phasefield.F90:525-566); the row is solid where
            // the averaged phic exceeds pf_direct_force = 0.9 (:29, :821-930)
            h0 -= HT((n.pc[r] + n.pc[r - 1]) * n.vx[r] * p.heta);
            sx = 0.5 * (n.pc[r] + n.pc[r + 1]) > 0.9;
        }
        r0 = double(rux[r]);
    }
    if (wc) {
        h1 = ex_vy<HT>(p, n, r);
        if (p.gaxis == 2) h1 -= sb;
        h2 = ex_vz<HT>(p, n, r);
        if (p.gaxis == 3) h2 -= sb;
        h3 = ex_temp<HT>(p, n, r);
        if (PF) {
            const double py = n.pc[r] + n.pc_jm[r], pz = n.pc[r] + n.pc_im[r];
            h1 -= HT(py * n.vy[r] * p.heta);
            h2 -= HT(pz * n.vz[r] * p.heta);
            sy = 0.5 * py > 0.9;
            sz = 0.5 * pz > 0.9;
            h3 += HT(p.qlat[c0 + r]);            // AddLatentHeat (:651-725)
        }
        if (p.moist) h3 = HT(double(h3) + condensation_heat(p, c0, r));
        r1 = double(ruy[r]); r2 = double(ruz[r]); r3 = double(rut[r]);
    }
    rhs[0] = wx ? rhs_vx(p, n, double(h0) + bref, r0 + bref, r) : 0.0;
    rhs[1] = wc ? rhs_vy(p, n, double(h1), r1, r) : 0.0;
    rhs[2] = wc ? rhs_vz(p, n, double(h2), r2, r) : 0.0;
    rhs[3] = wc ? rhs_temp(p, n, double(h3), r3, r) : 0.0;
    if (PF) {
        // a solid row is the identity with right-hand side -u
        if (sx) rhs[0] = -n.vx[r];
        if (sy) rhs[1] = -n.vy[r];
        if (sz) rhs[2] = -n.vz[r];
    }
    solid = unsigned(sx) | unsigned(sy) << 1 | unsigned(sz) << 2;
    if (wx) rux[r] = h0;
    if (wc) { ruy[r] = h1; ruz[r] = h2; rut[r] = h3; }
}

// Masked solve of one column in LDS (tight layout CHP) with the solid flags
// of its rows as bytes in natural row order (mskb[r] != 0: identity row,
// right-hand side already in buf); the other rows are those of f
template <int CH, int CHP>
__device__ __forceinline__ void column_solve_masked_bytes(double *buf, const unsigned char *mskb, const TriFactors &f) {
    const int lane = threadIdx.x & 63;
    double v[CH], a[CH], c[CH];
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const int r = lane * CH + m;
        const bool solid = mskb[r] != 0;
        v[m] = buf[lane * CHP + m];
        if (!solid) v[m] *= f.s[r];
        a[m] = solid ? 0.0 : f.a[r];
        c[m] = solid ? 0.0 : f.c[r];
    }
    wave_tri_solve_general<CH>(v, a, c);
#pragma unroll
    for (int m = 0; m < CH; ++m) buf[lane * CHP + m] = v[m];
}

// Single-pass form of k_column_base.  One sweep over the rows forms the
// four right-hand sides, loading every neighbour value once for all four
// explicit terms.  Those of vx and vy go to two LDS columns of the
// wavefront, those of vz and T stay in registers.  The vx and vy solves and
// updates follow, then the vz and T right-hand sides move into the same two
// LDS columns for their solves and updates.  Two LDS columns per wavefront
// instead of four keep 4 wavefronts per SIMD; the updates read the old value
// of the own column again (from cache).
// With PF the velocity terms carry the volume penalty, T the latent heat,
// and solid velocity rows are identity rows: their flags go to LDS as bytes
// (vx, vy; vz's flags wait in a register bit mask) for the masked solves.
// CH = ceil(nxm / 64) is the size of the largest system (nxm rows); when
// nx = 64 CH + 1 the last face has no equation (the wall of vx, the unused
// index of the others) and only the update touches it.
template <int CH, typename HT, bool PF>
__global__ void __launch_bounds__(256) k_column_base1(KP p, TriFactors fx, TriFactors fyz, TriFactors ft, BaseOut o) {
    constexpr int CHP = lds_chunk_tight<CH>();
    constexpr unsigned NROW = 64u * CH;
    __shared__ double lds[4][2][lds_col_size<CH, CHP>()];
    __shared__ unsigned char mskb[PF ? 4 : 1][2][PF ? NROW : 1];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned lane = threadIdx.x & 63;
    int j, i;
    const bool live = strip_column(xcd_block(0), wave, p.ny, p.nz, j, i);
    const Nb n = make_nb(p, j, i);
    const long long c0 = (long long)i * p.plane + (long long)j * p.ldx;
    const unsigned nx = p.nx, nxm = p.nxm;
    double *b0 = lds[wave][0], *b1 = lds[wave][1];
    unsigned char *m0 = mskb[PF ? wave : 0][0], *m1 = mskb[PF ? wave : 0][1];
    double keep[2][CH];                          // right-hand sides of vz and T
    unsigned zsolid = 0;                         // vz's solid flags, bit m for row lane + 64 m
#pragma unroll
    for (int m = 0; m < CH; ++m) {
        const unsigned r = lane + 64u * m;
        double rhs[4] = {0.0, 0.0, 0.0, 0.0};
        unsigned solid = 0;
        if (live && r < nxm) base_rhs4<HT, PF>(p, n, c0, r, rhs, solid);
        const unsigned q = lds_row<CH, CHP>(r);
        b0[q] = rhs[0];
        b1[q] = rhs[1];
        keep[0][m] = rhs[2];
        keep[1][m] = rhs[3];
        if (PF) {
            m0[r] = (unsigned char)(solid & 1u);
            m1[r] = (unsigned char)((solid >> 1) & 1u);
            zsolid |= ((solid >> 2) & 1u) << m;
        }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < 2; ++g) {
        if (g == 1) {
#pragma unroll
            for (int m = 0; m < CH; ++m) {
                const unsigned q = lds_row<CH, CHP>(lane + 64u * m);
                b0[q] = keep[0][m];
                b1[q] = keep[1][m];
                if (PF) m0[lane + 64u * m] = (unsigned char)((zsolid >> m) & 1u);
            }
            __syncthreads();
        }
        if (PF) {
            column_solve_masked_bytes<CH, CHP>(b0, m0, g == 0 ? fx : fyz);
            if (g == 0) column_solve_masked_bytes<CH, CHP>(b1, m1, fyz);
            else column_solve_lds<CH, CHP>(b1, ft);
        } else {
            column_solve_lds<CH, CHP>(b0, g == 0 ? fx : fyz);
            column_solve_lds<CH, CHP>(b1, g == 0 ? fyz : ft);
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                const int fld = 2 * g + f;
                const double *old = fld == 0 ? n.vx : fld == 1 ? n.vy : fld == 2 ? n.vz : n.t;
                double *out = (fld == 0 ? o.vx : fld == 1 ? o.vy : fld == 2 ? o.vz : o.temp) + c0;
                const double *d = f == 0 ? b0 : b1;
#pragma unroll 1
                for (unsigned r = lane; r < nx; r += 64) {
                    const double u = old[r];
                    const double x = d[lds_row<CH, CHP>(r < NROW ? r : NROW - 1)];
                    if (fld == 0) out[r] = (r >= 1 && r < nxm) ? u + x : 0.0;     // walls, ImplicitAndUpdateVX.F90:86-87
                    else out[r] = r < nxm ? u + x : u;
                }
            }
        }
        if (g == 0) __syncthreads();
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
// DT: storage type of dph (FP32 in the mixed form's sweeps)
template <typename DT>
__global__ void __launch_bounds__(256) k_project(KP p, int do_vel, int do_pr, Gate gate) {
    if (gate_closed(gate)) return;
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    if (k >= p.nxm || j >= p.ny) return;
    const int jm = wrapm(j, p.ny), jp = wrapp(j, p.ny);
    const DT *d = sel<DT>(p.dph, p.dphf);
    const int km = k == 0 ? 0 : k - 1, kp = k == p.nxm - 1 ? k : k + 1;
    // z walked with the grid's z stride: a full grid (one plane each) or, for
    // a gated launch, a short grid that costs little when the gate is closed
    // every load of a cell ahead of its stores (the fields may alias as far
    // as the compiler knows; interleaved stores would serialise the loads)
    for (int i = blockIdx.z; i < p.nz; i += gridDim.z) {
        const int im = wrapm(i, p.nz), ip = wrapp(i, p.nz);
        const long long c = bidx(p, k, j, i);
        const double dc = d[c], dkm = d[bidx(p, km, j, i)], djm = d[bidx(p, k, jm, i)], dim = d[bidx(p, k, j, im)];
        if (do_vel) {
            const double adt = p.al * p.dt;
            const double ux = p.vx[c], uy = p.vy[c], uz = p.vz[c];
            const double nux = ux - (dc - dkm) * (adt * p.udx3c[k]);
            const double nuy = uy - (dc - djm) * (adt * p.dy);
            const double nuz = uz - (dc - dim) * (adt * p.dz);
            if (do_pr) {
                const double be = p.al * p.beta, prc = p.pr[c];
                const double npr = prc + (dc - be * ((d[bidx(p, k, j, ip)] - 2.0 * dc + dim) * p.dzq +
                                                     (d[bidx(p, k, jp, i)] - 2.0 * dc + djm) * p.dyq +
                                                     (d[bidx(p, kp, j, i)] * p.apphk[k] + dc * p.acphk[k] + dkm * p.amphk[k])));
                p.pr[c] = npr;
            }
            p.vx[c] = nux; p.vy[c] = nuy; p.vz[c] = nuz;
        } else if (do_pr) {
            const double be = p.al * p.beta;
            p.pr[c] += dc - be * ((d[bidx(p, k, j, ip)] - 2.0 * dc + dim) * p.dzq +
                                  (d[bidx(p, k, jp, i)] - 2.0 * dc + djm) * p.dyq +
                                  (d[bidx(p, kp, j, i)] * p.apphk[k] + dc * p.acphk[k] + dkm * p.amphk[k]));
        }
    }
}

// The same projection writing the corrected velocity to a second buffer
// (u_out), so that every cell can also form the divergence the corrected
// field will have (its own and its +1 neighbours' faces, from the old
// velocity and dph): the FP64 residual of the mixed form's iterative
// refinement, max-reduced into *slot (CheckDivergence units).  Each thread
// marches along z over PROJ_ZC planes and carries dph and vz of the planes
// i-1, i, i+1 in registers, so no plane is read twice.
constexpr int PROJ_ZC = 8;
__global__ void __launch_bounds__(256) k_project_res(KP p, double *vx_out, double *vy_out, double *vz_out,
                                                     double *partial) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const int i0 = blockIdx.z * PROJ_ZC, i1 = min(i0 + PROJ_ZC, p.nz);
    double dvm = 0.0;
    if (k < p.nxm && j < p.ny) {
        const int jm = wrapm(j, p.ny), jp = wrapp(j, p.ny);
        const int km = k == 0 ? 0 : k - 1, kp = k == p.nxm - 1 ? k : k + 1;
        const float *d = p.dphf;                  // the mixed form's FP32 sweep correction
        const double adt = p.al * p.dt, be = p.al * p.beta;
        const double ucx = adt * p.udx3c[k], ucx1 = k + 1 <= p.nxm - 1 ? adt * p.udx3c[k + 1] : 0.0;
        const long long row = (long long)j * p.ldx + k, rowjm = (long long)jm * p.ldx + k, rowjp = (long long)jp * p.ldx + k;
        double d_m = d[(long long)wrapm(i0, p.nz) * p.plane + row];
        double d_c = d[(long long)i0 * p.plane + row];
        double w_c = p.vz[(long long)i0 * p.plane + row];
        for (int i = i0; i < i1; ++i) {
            const long long pl = (long long)i * p.plane, plp = (long long)wrapp(i, p.nz) * p.plane;
            const long long c = pl + row;
            // every load of the cell ahead of its stores (the outputs may
            // alias the inputs as far as the compiler knows)
            const double d_p = d[plp + row], w_p = p.vz[plp + row];
            const double dkm = d[pl + row - k + km], dkp = d[pl + row - k + kp];
            const double djm = d[pl + rowjm], djp = d[pl + rowjp];
            const double vxc = p.vx[c], vyc = p.vy[c], vx1 = p.vx[c + 1], vy1 = p.vy[pl + rowjp];
            const double d1 = d[c + 1], prc = p.pr[c];
            // corrected faces of this cell (as k_project)
            const double ux = vxc - (d_c - dkm) * ucx;
            const double uy = vyc - (d_c - djm) * (adt * p.dy);
            const double uz = w_c - (d_c - d_m) * (adt * p.dz);
            // corrected +1 faces (owned by the neighbouring cells)
            const double ux1 = k + 1 <= p.nxm - 1 ? vx1 - (d1 - d_c) * ucx1 : vx1;
            const double uy1 = vy1 - (djp - d_c) * (adt * p.dy);
            const double uz1 = w_p - (d_p - d_c) * (adt * p.dz);
            dvm = fmax(dvm, fabs((uz1 - uz) * p.dz + (uy1 - uy) * p.dy + (ux1 - ux) * p.udx3m[k]));
            const double npr = prc + (d_c - be * ((d_p - 2.0 * d_c + d_m) * p.dzq + (djp - 2.0 * d_c + djm) * p.dyq +
                                                 (dkp * p.apphk[k] + d_c * p.acphk[k] + dkm * p.amphk[k])));
            if (k == p.nxm - 1) {      // index nx: wall face of vx (not corrected), unused for vy, vz
                const double wy = p.vy[c + 1], wz = p.vz[c + 1];
                vx_out[c + 1] = vx1; vy_out[c + 1] = wy; vz_out[c + 1] = wz;
            }
            vx_out[c] = ux; vy_out[c] = uy; vz_out[c] = uz;
            p.pr[c] = npr;
            d_m = d_c; d_c = d_p; w_c = w_p;
        }
    }
    dvm = block_max(dvm);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partial[((long long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = dvm;
}
// max of n per-workgroup partial maxima into *slot (grid-stride, one atomic
// per workgroup of a small grid)
__global__ void __launch_bounds__(256) k_max_partials(const double *partial, long long n, unsigned long long *slot) {
    double v = 0.0;
    for (long long q = (long long)blockIdx.x * 256 + threadIdx.x; q < n; q += (long long)gridDim.x * 256) v = fmax(v, partial[q]);
    v = block_max(v);
    if (threadIdx.x == 0) atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}

// ---------------------------------------------------------------------------
// multiresolution transfers
// ---------------------------------------------------------------------------
// tiled transfers of InterpVelMgrd: the coarse gradient is computed on the
// fly from the velocity (InterpVelMgrd.F90:37-51, :75-86, :161-172, with the
// wall ghosts), a coarse block is staged in LDS once per tile, contracted
// axis by axis (separable Hermite stencil), and the transferred gradient is
// integrated along its axis with the running sum carried across chunks.
// ---------------------------------------------------------------------------
// coarse gradient of component comp at reference x index xc (-1..nx+1) and
// 0-based periodic (j, i); the difference of the FP64 velocities is taken in
// FP64 and handed on in the transfer type RT
template <typename RT>
__device__ __forceinline__ RT mgrd_grad(const KP &p, int comp, int xc, int j, int i) {
    int kc = xc, sgn = 1;
    if (xc == 0) { kc = 1; sgn = 1 - 2 * p.inslws; }
    else if (xc == p.nx) { kc = p.nxm; sgn = 1 - 2 * p.inslwn; }
    else if (xc < 0 || xc > p.nx) return RT(0);
    const int k = kc - 1;
    double d;
    if (comp == 0) {
        const double *c = p.vx + (long long)i * p.plane + (long long)j * p.ldx;
        d = (c[k + 1] - c[k]) * p.udx3m[k];
    } else if (comp == 1) {
        const double *c = p.vy + (long long)i * p.plane;
        d = (c[(long long)wrapp(j, p.ny) * p.ldx + k] - c[(long long)j * p.ldx + k]) * p.dy;
    } else {
        const double *c = p.vz + (long long)j * p.ldx;
        d = (c[(long long)wrapp(i, p.nz) * p.plane + k] - c[(long long)i * p.plane + k]) * p.dz;
    }
    return RT(sgn * d);
}
__device__ __forceinline__ int wrapn(int j, int n) { j %= n; return j < 0 ? j + n : j; }
// periodic wrap of an index a few periods away at most, without a division
__device__ __forceinline__ int wrapf(int j, int n) {
    while (j < 0) j += n;
    while (j >= n) j -= n;
    return j;
}

// (c * B + b) * A + a for staging-block indices (all below 2^24): 24-bit
// multiplies, which issue at full rate
__device__ __forceinline__ int ix3(int c, int B, int b, int A, int a) { return __mul24(__mul24(c, B) + b, A) + a; }

// Components (a, b, c) of a linear index q = a + A (b + B c) that advances by a
// fixed step: the divisions happen once, each step costs a few adds.
struct Walk3 {
    int a, b, c, da, db, dc, A, B;
    __device__ __forceinline__ Walk3(int q, int step, int A_, int B_) : A(A_), B(B_) {
        a = q % A; b = (q / A) % B; c = q / (A * B);
        da = step % A; db = (step / A) % B; dc = step / (A * B);
    }
    __device__ __forceinline__ void next() {
        a += da; b += db; c += dc;
        if (a >= A) { a -= A; ++b; }
        if (b >= B) { b -= B; ++c; }
    }
};

// Tile sizes of the transfer kernels (host computes the LDS each needs).
constexpr int VXR_TJ = 16, VXR_TI = 4;      // refined columns per vxr tile
constexpr int VYZ_LINES = 4;                // refined lines (waves) per vyr/vzr tile
constexpr int VYZ_CHUNK = 16;               // refined points per scan chunk
constexpr int VXR_BU = 4, VYZ_BU = 4;       // coarse-block items loaded per batch and thread
constexpr int VYZ_NB = 4;                   // coarse planes per load batch of k_vyz_wave (8: slower)

// Inclusive prefix sum over the 64 lanes of a wavefront with DPP lane
// moves: Kogge-Stone within each row of 16 lanes (row_shr 1, 2, 4, 8),
// then the row totals carried forward (row_bcast 15 into rows 1 and 3,
// row_bcast 31 into rows 2 and 3).  No LDS traffic.
template <int CTRL, int ROWS, bool BC>
__device__ __forceinline__ int dpp32(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, BC); }
template <int CTRL, int ROWS, bool BC>
__device__ __forceinline__ float dpp_move(float v) { return __builtin_bit_cast(float, dpp32<CTRL, ROWS, BC>(__builtin_bit_cast(int, v))); }
template <int CTRL, int ROWS, bool BC>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = dpp32<CTRL, ROWS, BC>(int(b & 0xffffffffLL)), hi = dpp32<CTRL, ROWS, BC>(int(b >> 32));
    return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// value of lane 63 (the wavefront total of an inclusive scan), uniform
__device__ __forceinline__ float wave_lane63(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}
__device__ __forceinline__ double wave_lane63(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_readlane(int(b & 0xffffffffLL), 63), hi = __builtin_amdgcn_readlane(int(b >> 32), 63);
    return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    v += dpp_move<0x111, 0xf, true>(v);      // row_shr:1
    v += dpp_move<0x112, 0xf, true>(v);      // row_shr:2
    v += dpp_move<0x114, 0xf, true>(v);      // row_shr:4
    v += dpp_move<0x118, 0xf, true>(v);      // row_shr:8
    v += dpp_move<0x142, 0xa, false>(v);     // row_bcast:15
    v += dpp_move<0x143, 0xc, false>(v);     // row_bcast:31
    return v;
}

// vxr: tile of VXR_TJ x VXR_TI refined columns, x processed in chunks of 64
// refined points (one per lane); xcmax/cjmax/cimax bound the coarse block.
// RT: storage and arithmetic type of the transfer (double or float).  The
// tile's coarse column offsets and its y / z stencils are staged in LDS once.
template <typename RT>
__device__ __forceinline__ void vxr_tile(const KP &p, int xcmax, int cjmax, int cimax) {
    extern __shared__ __align__(16) unsigned char smraw[];
    __shared__ long long coff[16 * 16];                 // coarse column offsets [ci][cj] (cimax, cjmax <= 16)
    // per coarse x of the chunk: the row of its difference and the metric
    // with the wall-ghost sign (0 outside -1..nx+1)
    __shared__ int krow[128];
    __shared__ double kfac[128];
    __shared__ RT yw[VXR_TJ][4], zw[VXR_TI][4];
    __shared__ int yo[VXR_TJ], zo[VXR_TI];
    RT *sm = reinterpret_cast<RT *>(smraw);
    RT *vxr = sel<RT>(p.vxr, p.vxrf);
    const RT *xs_w = sel<RT>(p.xs_w, p.xs_wf), *ys_w = sel<RT>(p.ys_w, p.ys_wf), *zs_w = sel<RT>(p.zs_w, p.zs_wf);
    const RT *lxa = sel<RT>(p.lxa, p.lxaf);
    int tj0, ti0;
    if (!tile_yz_grouped((p.nyr + VXR_TJ - 1) / VXR_TJ, (p.nzr + VXR_TI - 1) / VXR_TI, tj0, ti0)) return;
    const int J0 = tj0 * VXR_TJ, I0 = ti0 * VXR_TI;
    const int nj = min(VXR_TJ, p.nyr - J0), ni = min(VXR_TI, p.nzr - I0);
    const int cj0 = p.ys_src[J0] - 1, CJ = p.ys_src[J0 + nj - 1] - 1 + 4 - cj0;
    const int ci0 = p.zs_src[I0] - 1, CI = p.zs_src[I0 + ni - 1] - 1 + 4 - ci0;
    RT *B = sm;                                       // [ci][cj][xc]
    RT *Z = B + xcmax * cjmax * cimax;                // [ti][cj][xc]
    RT *Q = Z + xcmax * cjmax * VXR_TI;               // [ti][tj][xc]
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    for (int q = t; q < CI * CJ; q += 256) {
        const int c = q / CJ, b = q - c * CJ;
        coff[c * CJ + b] = (long long)wrapf(ci0 + c, p.nz) * p.plane + (long long)wrapf(cj0 + b, p.ny) * p.ldx;
    }
    if (t < nj) {
        yo[t] = p.ys_src[J0 + t] - 1 - cj0;
        for (int a = 0; a < 4; ++a) yw[t][a] = ys_w[4 * (J0 + t) + a];
    }
    if (t < ni) {
        zo[t] = p.zs_src[I0 + t] - 1 - ci0;
        for (int a = 0; a < 4; ++a) zw[t][a] = zs_w[4 * (I0 + t) + a];
    }
    for (int q = t; q < nj * ni; q += 256) {
        const int ti = q / nj, tj = q - ti * nj;
        vxr[ridx(p, 0, J0 + tj, I0 + ti)] = RT(0);      // vxr(1) = 0
    }
    // running sums of the columns (ti, wave + 4u) this wavefront writes,
    // carried from one x chunk to the next
    RT run[VXR_TI][VXR_TJ / 4];
#pragma unroll
    for (int ti = 0; ti < VXR_TI; ++ti)
#pragma unroll
        for (int u = 0; u < VXR_TJ / 4; ++u) run[ti][u] = RT(0);
    const int sS = 1 - 2 * p.inslws, sN = 1 - 2 * p.inslwn;
    for (int R0 = 0; R0 < p.nxmr; R0 += 64) {
        const int R1 = min(R0 + 64, p.nxmr);
        const int x0 = p.xs_src[R0], XC = p.xs_src[R1 - 1] + 4 - x0;
        __syncthreads();
        for (int xo = t; xo < XC; xo += 256) {
            const int xc = x0 + xo;
            const bool ok = xc >= 0 && xc <= p.nx;
            const int k = !ok ? 0 : (xc == 0 ? 0 : (xc == p.nx ? p.nxm - 1 : xc - 1));
            krow[xo] = k;
            kfac[xo] = !ok ? 0.0 : (xc == 0 ? sS * p.udx3m[k] : (xc == p.nx ? sN * p.udx3m[k] : p.udx3m[k]));
        }
        __syncthreads();
        {
            // coarse gradient dvx/dx at reference x index xc with the wall
            // ghosts (InterpVelMgrd.F90:37-51); zero outside -1..nx+1
            // (kfac = sign * udx3m: the sign is +-1, so the product is
            // exactly the signed difference quotient).  The loads of VXR_BU
            // items are issued before any is used: an item outside the block
            // reads a valid address.
            const int nB = XC * CJ * CI;
            Walk3 w(t, 256, XC, CJ);
            for (int q0 = t; q0 < nB; q0 += 256 * VXR_BU) {
                double a0[VXR_BU], a1[VXR_BU], f[VXR_BU];
                int dst[VXR_BU];
#pragma unroll
                for (int r = 0; r < VXR_BU; ++r) {
                    const bool live = q0 + 256 * r < nB;
                    const int xo = live ? w.a : 0;
                    const int k = krow[xo];
                    f[r] = kfac[xo];
                    const double *col = p.vx + coff[min(w.c, CI - 1) * CJ + min(w.b, CJ - 1)];
                    a0[r] = col[k]; a1[r] = col[k + 1];
                    dst[r] = live ? ix3(w.c, cjmax, w.b, xcmax, w.a) : -1;
                    w.next();
                }
#pragma unroll
                for (int r = 0; r < VXR_BU; ++r)
                    if (dst[r] >= 0) B[dst[r]] = RT(f[r] * (a1[r] - a0[r]));
            }
        }
        __syncthreads();
        Walk3 wz(t, 256, XC, CJ);
        for (int q = t; q < XC * CJ * ni; q += 256, wz.next()) {
            const int xc = wz.a, cj = wz.b, ti = wz.c;
            const RT *b = B + ix3(zo[ti], cjmax, cj, xcmax, xc);
            const int st = __mul24(cjmax, xcmax);
            Z[ix3(ti, cjmax, cj, xcmax, xc)] = b[0] * zw[ti][0] + b[st] * zw[ti][1] + b[2 * st] * zw[ti][2] + b[3 * st] * zw[ti][3];
        }
        __syncthreads();
        Walk3 wy(t, 256, XC, nj);
        for (int q = t; q < XC * nj * ni; q += 256, wy.next()) {
            const int xc = wy.a, tj = wy.b, ti = wy.c;
            const RT *z = Z + ix3(ti, cjmax, yo[tj], xcmax, xc);
            Q[ix3(ti, VXR_TJ, tj, xcmax, xc)] = z[0] * yw[tj][0] + z[xcmax] * yw[tj][1] + z[2 * xcmax] * yw[tj][2] + z[3 * xcmax] * yw[tj][3];
        }
        __syncthreads();
        const int kr = R0 + lane;
        const bool in = kr < R1;
        int xa = 0;
        RT wx0 = 0, wx1 = 0, wx2 = 0, wx3 = 0, lx = 0;
        if (in) {
            xa = p.xs_src[kr] - x0;
            wx0 = xs_w[4 * kr]; wx1 = xs_w[4 * kr + 1]; wx2 = xs_w[4 * kr + 2]; wx3 = xs_w[4 * kr + 3];
            lx = lxa[kr];
        }
#pragma unroll
        for (int ti = 0; ti < VXR_TI; ++ti) {
            if (ti >= ni) break;
            RT *orow = vxr + (long long)(I0 + ti) * p.planer + (long long)J0 * p.ldxr + kr + 1;
#pragma unroll
            for (int u = 0; u < VXR_TJ / 4; ++u) {
                const int tj = wave + 4 * u;
                if (tj >= nj) break;
                const RT *qq = Q + ix3(ti, VXR_TJ, tj, xcmax, xa);
                RT v = in ? (qq[0] * wx0 + qq[1] * wx1 + qq[2] * wx2 + qq[3] * wx3) * lx : RT(0);
                v = wave_incl_scan(v);
                const RT base = run[ti][u];
                if (in) orow[(long long)tj * p.ldxr] = base + v;
                run[ti][u] = base + wave_lane63(v);
            }
        }
    }
}

template <typename RT>
__global__ void __launch_bounds__(256) k_vxr_tile(KP p, int xcmax, int cjmax, int cimax) { vxr_tile<RT>(p, xcmax, cjmax, cimax); }
// FP32: 96 registers leave room for the 5 workgroups per CU its LDS allows
template <>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8))) k_vxr_tile<float>(KP p, int xcmax, int cjmax, int cimax) {
    vxr_tile<float>(p, xcmax, cjmax, cimax);
}

// vyr (AX = 1) / vzr (AX = 2): lanes = 64 refined x, waves = VYZ_LINES lines
// of the other periodic axis; the scan axis is walked in chunks of VYZ_CHUNK.
// Per chunk: B = coarse gradient block [xc][cs][co], O = contraction over the
// other axis [xc][cs][line], X = contraction over x [cs][line][lane], then a
// running sum along the scan axis in registers.
template <int AX, typename RT>
__global__ void __launch_bounds__(256) k_vyz_tile(KP p, int xcmax, int csmax, int comax) {
    extern __shared__ __align__(16) unsigned char smraw[];
    RT *sm = reinterpret_cast<RT *>(smraw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, t = threadIdx.x;
    const int nS = AX == 1 ? p.nyr : p.nzr, nO = AX == 1 ? p.nzr : p.nyr;          // refined extents
    const int cS = AX == 1 ? p.ny : p.nz, cO = AX == 1 ? p.nz : p.ny;              // coarse extents
    const int *sS = AX == 1 ? p.ys_src : p.zs_src, *sO = AX == 1 ? p.zs_src : p.ys_src;
    const RT *ysw = sel<RT>(p.ys_w, p.ys_wf), *zsw = sel<RT>(p.zs_w, p.zs_wf), *xs_w = sel<RT>(p.xs_w, p.xs_wf);
    const RT *wS = AX == 1 ? ysw : zsw, *wO = AX == 1 ? zsw : ysw;
    const int R0 = blockIdx.x * 64, R1 = min(R0 + 64, p.nxmr);
    const int O0 = blockIdx.y * VYZ_LINES, nl = min(VYZ_LINES, nO - O0);
    const int kr = R0 + lane, line = O0 + wave;
    const bool live = wave < nl;
    RT *outp = AX == 1 ? sel<RT>(p.vyr, p.vyrf) : sel<RT>(p.vzr, p.vzrf);
    const long long sstride = AX == 1 ? (long long)p.ldxr : p.planer;
    const long long obase = AX == 1 ? (long long)line * p.planer : (long long)line * p.ldxr;
    // the unused x index nxr of vyr / vzr stays zero (InterpVelMgrd.F90:29-31)
    if (blockIdx.x == gridDim.x - 1 && lane == 0 && live)
        for (int s = 0; s < nS; ++s) outp[obase + s * sstride + p.nxmr] = RT(0);
    const int x0 = p.xs_src[R0], XC = p.xs_src[R1 - 1] + 4 - x0;
    const int co0 = sO[O0] - 1, CO = sO[O0 + nl - 1] - 1 + 4 - co0;
    RT wx[4] = {0, 0, 0, 0}, wo[4] = {0, 0, 0, 0};
    int xa = 0, oa = 0;
    if (kr < R1) { xa = p.xs_src[kr] - x0; for (int a = 0; a < 4; ++a) wx[a] = xs_w[4 * kr + a]; }
    if (live) { oa = sO[line] - 1 - co0; for (int a = 0; a < 4; ++a) wo[a] = wO[4 * line + a]; }
    // first guess on the plane s = 0 (InterpVelMgrd.F90:93-123 / :179-209):
    // velocity plane with wall ghosts, contracted over the other axis, then x
    RT v = RT(0);
    if (live && kr < R1 && (AX == 1 || p.nz > 1)) {
        const int sSg = 1 - 2 * p.inslws, sNg = 1 - 2 * p.inslwn;
        for (int a = 0; a < 4; ++a) {
            const int kc = x0 + xa + a;
            RT q1 = RT(0);
            for (int c = 0; c < 4; ++c) {
                const int o = wrapn(co0 + oa + c, cO);
                const double *col = AX == 1 ? p.vy + (long long)o * p.plane : p.vz + (long long)o * p.ldx;
                RT f;
                if (kc >= 1 && kc <= p.nxm) f = RT(col[kc - 1]);
                else if (kc == 0) f = RT(sSg * col[0]);
                else if (kc == p.nx) f = RT(sNg * col[p.nxm - 1]);
                else f = RT(0);
                q1 = c == 0 ? f * wo[0] : q1 + f * wo[c];
            }
            v = a == 0 ? q1 * wx[0] : v + q1 * wx[a];
        }
    }
    if (live && kr < R1) outp[obase + kr] = v;
    const RT dl = AX == 1 ? RT(1.0 / p.dyr) : RT(1.0 / p.dzr);
    RT *B = sm;                                               // [co][cs][xc]
    RT *X = B;                                                // [cs][line][lane], reuses B
    RT *O = B + max(xcmax * csmax * comax, csmax * VYZ_LINES * 64);   // [line][cs][xc]
    // coarse offsets of the tile: the other axis (fixed) and, per chunk, the
    // scan axis and its periodic successor; the lines' stencils
    __shared__ long long ooff[32], soff[32], snext[32];
    __shared__ RT lw[VYZ_LINES][4];
    __shared__ int lo[VYZ_LINES];
    const long long ostride = AX == 1 ? p.plane : (long long)p.ldx, sstr = AX == 1 ? (long long)p.ldx : p.plane;
    for (int q = t; q < CO; q += 256) ooff[q] = (long long)wrapf(co0 + q, cO) * ostride;
    if (t < nl) {
        lo[t] = sO[O0 + t] - 1 - co0;
        for (int a = 0; a < 4; ++a) lw[t][a] = wO[4 * (O0 + t) + a];
    }
    const double *vel = AX == 1 ? p.vy : p.vz;
    const double dgrid = AX == 1 ? p.dy : p.dz;
    const int sSg = 1 - 2 * p.inslws, sNg = 1 - 2 * p.inslwn;
    for (int S0 = 0; S0 < nS - 1; S0 += VYZ_CHUNK) {
        const int S1 = min(S0 + VYZ_CHUNK, nS - 1);
        const int cs0 = sS[S0] - 1, CS = sS[S1 - 1] - 1 + 4 - cs0;
        __syncthreads();
        for (int q = t; q < CS; q += 256) {
            const int js = wrapf(cs0 + q, cS);
            soff[q] = (long long)js * sstr;
            snext[q] = (long long)wrapp(js, cS) * sstr;
        }
        __syncthreads();
        {
            // coarse gradient d(vy)/dy or d(vz)/dz at reference x index xc
            // with the wall ghosts (InterpVelMgrd.F90:75-86, :161-172)
            // (loads of VYZ_BU items issued before any is used; an item
            // outside the block or the grid reads a valid address)
            const int nB = XC * CS * CO;
            Walk3 w(t, 256, XC, CS);
            for (int q0 = t; q0 < nB; q0 += 256 * VYZ_BU) {
                double a0[VYZ_BU], a1[VYZ_BU];
                int dst[VYZ_BU], xs[VYZ_BU];
#pragma unroll
                for (int r = 0; r < VYZ_BU; ++r) {
                    const int xc = x0 + w.a;
                    const bool live = q0 + 256 * r < nB, ok = live && xc >= 0 && xc <= p.nx;
                    const int k = !ok ? 0 : (xc == 0 ? 0 : (xc == p.nx ? p.nxm - 1 : xc - 1));
                    const int c = min(w.c, CO - 1), b = min(w.b, CS - 1);
                    const double *cb = vel + (ok ? ooff[c] : 0) + k;
                    a1[r] = cb[ok ? snext[b] : 0];
                    a0[r] = cb[ok ? soff[b] : 0];
                    dst[r] = live ? ix3(w.c, csmax, w.b, xcmax, w.a) : -1;
                    xs[r] = ok ? xc : -1;
                    w.next();
                }
#pragma unroll
                for (int r = 0; r < VYZ_BU; ++r)
                    if (dst[r] >= 0) {
                        const int xc = xs[r];
                        const double d = (a1[r] - a0[r]) * dgrid;
                        B[dst[r]] = xc < 0 ? RT(0) : RT(xc == 0 ? sSg * d : (xc == p.nx ? sNg * d : d));
                    }
            }
        }
        __syncthreads();
        // contraction over the other axis: one (cs, xc) per item, all lines
        {
            const int st = __mul24(csmax, xcmax);
            Walk3 wo(t, 256, XC, CS);
            for (int q = t; q < XC * CS; q += 256, wo.next()) {
                const int base = __mul24(wo.b, xcmax) + wo.a;
#pragma unroll
                for (int l = 0; l < VYZ_LINES; ++l)
                    if (l < nl) {
                        const RT *b = B + __mul24(lo[l], st) + base;
                        O[__mul24(l, st) + base] = b[0] * lw[l][0] + b[st] * lw[l][1] + b[2 * st] * lw[l][2] + b[3 * st] * lw[l][3];
                    }
            }
        }
        __syncthreads();
        if (live && kr < R1)
            for (int cs = 0; cs < CS; ++cs) {
                const RT *o = O + ix3(wave, csmax, cs, xcmax, xa);
                X[(cs * VYZ_LINES + wave) * 64 + lane] = o[0] * wx[0] + o[1] * wx[1] + o[2] * wx[2] + o[3] * wx[3];
            }
        __syncthreads();
        if (live && kr < R1)
            for (int s = S0; s < S1; ++s) {
                const int b0 = sS[s] - 1 - cs0;
                const RT *w = wS + 4 * s, *xx = X + (b0 * VYZ_LINES + wave) * 64 + lane;
                const int st = VYZ_LINES * 64;
                const RT tv = xx[0] * w[0] + xx[st] * w[1] + xx[2 * st] * w[2] + xx[3 * st] * w[3];
                v = v + tv * dl;
                outp[obase + (s + 1) * sstride + kr] = v;
            }
    }
}

// vyr (AX = 1) / vzr (AX = 2) with one wavefront per refined line of the
// other periodic axis and no LDS: lanes = 64 refined x, the scan axis walked
// point by point (InterpVelMgrd.F90:68-237).  Per coarse scan index cs the
// wavefront forms, for each coarse x of its chunk (one per lane; a second
// register when the chunk spans more than 64 coarse points), P(xc) = sum_c
// wo_c G(xc, cs, o_c), the coarse gradient G with the wall ghosts contracted
// over the other axis; then X(kr) = sum_a wx_a P(xa + a) by lane shuffles.
// The last four X stay in registers for the scan-axis stencil of each
// refined point, and the velocity row of plane cs + 1 is carried to the
// next step, so a step loads one row per o_c.  The arithmetic is that of
// k_vyz_tile term by term.  Workgroups of 4 wavefronts take 4 consecutive
// lines of one x chunk; consecutive workgroups on an XCD take consecutive
// line groups (shared coarse rows in its L2).
// (FP32: at least 6 wavefronts per SIMD, 71 VGPRs, which make 7; FP64 keeps
// its 94 VGPRs, since the 80 that 6 wavefronts allow spill and cost 30% on a
// 768 x 1536^2 refined grid; 8 wavefronts spill in FP32 too)
template <int AX, typename RT, bool TWO>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(RT) == 4 ? 6 : 1, 8))) k_vyz_wave(KP p, int nxchunk) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nS = AX == 1 ? p.nyr : p.nzr, nO = AX == 1 ? p.nzr : p.nyr;          // refined extents
    const int cS = AX == 1 ? p.ny : p.nz, cO = AX == 1 ? p.nz : p.ny;              // coarse extents
    const int *sS = AX == 1 ? p.ys_src : p.zs_src, *sO = AX == 1 ? p.zs_src : p.ys_src;
    const RT *ysw = sel<RT>(p.ys_w, p.ys_wf), *zsw = sel<RT>(p.zs_w, p.zs_wf), *xs_w = sel<RT>(p.xs_w, p.xs_wf);
    const RT *wS = AX == 1 ? ysw : zsw, *wO = AX == 1 ? zsw : ysw;
    const int ngroups = (nO + 3) / 4;
    const long long blk = xcd_block(0);
    if (blk >= (long long)nxchunk * ngroups) return;
    const int xchunk = int(blk / ngroups), line = int(blk % ngroups) * 4 + wave;
    if (line >= nO) return;
    const int R0 = xchunk * 64, R1 = min(R0 + 64, p.nxmr);
    const int kr = R0 + lane;
    const bool in = kr < R1;
    RT *outp = AX == 1 ? sel<RT>(p.vyr, p.vyrf) : sel<RT>(p.vzr, p.vzrf);
    const long long sstride = AX == 1 ? (long long)p.ldxr : p.planer;
    const long long obase = AX == 1 ? (long long)line * p.planer : (long long)line * p.ldxr;
    // the unused x index nxr of vyr / vzr stays zero (InterpVelMgrd.F90:29-31)
    if (xchunk == nxchunk - 1 && lane == 0)
        for (int s = 0; s < nS; ++s) outp[obase + s * sstride + p.nxmr] = RT(0);
    const int x0 = p.xs_src[R0];
    RT wx[4] = {0, 0, 0, 0};
    int xa = 0;
    if (in) { xa = p.xs_src[kr] - x0; for (int a = 0; a < 4; ++a) wx[a] = xs_w[4 * kr + a]; }
    // the line's stencil over the other axis (wavefront-uniform)
    const long long ostride = AX == 1 ? p.plane : (long long)p.ldx, sstr = AX == 1 ? (long long)p.ldx : p.plane;
    const int o0 = sO[line] - 1;
    RT wo[4];
    long long ooff[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) { wo[c] = wO[4 * line + c]; ooff[c] = (long long)wrapn(o0 + c, cO) * ostride; }
    const int sSg = 1 - 2 * p.inslws, sNg = 1 - 2 * p.inslwn;
    const double *vel = AX == 1 ? p.vy : p.vz;
    // first guess on the plane s = 0 (InterpVelMgrd.F90:93-123 / :179-209):
    // velocity plane with wall ghosts, contracted over the other axis, then x
    RT v = RT(0);
    if (in && (AX == 1 || p.nz > 1)) {
        for (int a = 0; a < 4; ++a) {
            const int kc = x0 + xa + a;
            RT q1 = RT(0);
            for (int c = 0; c < 4; ++c) {
                const double *col = vel + ooff[c];
                RT f;
                if (kc >= 1 && kc <= p.nxm) f = RT(col[kc - 1]);
                else if (kc == 0) f = RT(sSg * col[0]);
                else if (kc == p.nx) f = RT(sNg * col[p.nxm - 1]);
                else f = RT(0);
                q1 = c == 0 ? f * wo[0] : q1 + f * wo[c];
            }
            v = a == 0 ? q1 * wx[0] : v + q1 * wx[a];
        }
    }
    if (in) outp[obase + kr] = v;
    if (nS < 2) return;
    // this lane's coarse x rows (xc = x0 + lane, and + 64): row, ghost sign
    // (0 outside -1..nx+1 reads row 0)
    auto xrow = [&](int xc, int &k, double &sg) {
        const bool ok = xc >= 0 && xc <= p.nx;
        k = !ok ? 0 : (xc == 0 ? 0 : (xc == p.nx ? p.nxm - 1 : xc - 1));
        sg = !ok ? 0.0 : (xc == 0 ? double(sSg) : (xc == p.nx ? double(sNg) : 1.0));
    };
    // lanes past the chunk's XC coarse points are never gathered: they take
    // the last point's row, so their loads stay on cache lines in use
    const int XC = p.xs_src[R1 - 1] + 4 - x0;
    int k0, k1;
    double g0, g1;
    xrow(x0 + min(lane, XC - 1), k0, g0);
    xrow(x0 + min(64 + lane, XC - 1), k1, g1);
    const double dgrid = AX == 1 ? p.dy : p.dz;
    const RT dl = AX == 1 ? RT(1.0 / p.dyr) : RT(1.0 / p.dzr);
    int cs = sS[0] - 1;                       // unwrapped coarse scan index of the next X
    double va[4], vb[4];                      // velocity rows of plane cs (second register: vb)
    {
        const long long pl = (long long)wrapf(cs, cS) * sstr;
#pragma unroll
        for (int c = 0; c < 4; ++c) { va[c] = vel[ooff[c] + pl + k0]; vb[c] = TWO ? vel[ooff[c] + pl + k1] : 0.0; }
    }
    // X at the VYZ_NB coarse indices cs .. cs + VYZ_NB - 1 into W[3 ..]; the
    // rows of all VYZ_NB planes are loaded before any is used, so a wavefront
    // keeps 4 VYZ_NB loads in flight instead of waiting on each plane
    RT W[VYZ_NB + 3];
    auto batch = [&]() {
        double na[VYZ_NB][4], nb[VYZ_NB][4];
#pragma unroll
        for (int q = 0; q < VYZ_NB; ++q) {
            const long long pn = (long long)wrapf(cs + 1 + q, cS) * sstr;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                na[q][c] = vel[ooff[c] + pn + k0];
                nb[q][c] = TWO ? vel[ooff[c] + pn + k1] : 0.0;
            }
        }
#pragma unroll
        for (int q = 0; q < VYZ_NB; ++q) {
            RT b[4], b2[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double p0 = q == 0 ? va[c] : na[q - 1][c];
                b[c] = RT(g0 * ((na[q][c] - p0) * dgrid));
                if (TWO) {
                    const double p1 = q == 0 ? vb[c] : nb[q - 1][c];
                    b2[c] = RT(g1 * ((nb[q][c] - p1) * dgrid));
                } else {
                    b2[c] = RT(0);
                }
            }
            const RT P = b[0] * wo[0] + b[1] * wo[1] + b[2] * wo[2] + b[3] * wo[3];
            const RT P2 = b2[0] * wo[0] + b2[1] * wo[1] + b2[2] * wo[2] + b2[3] * wo[3];
            RT o[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int idx = xa + a;
                o[a] = __shfl(P, idx & 63, 64);
                if (TWO) {
                    const RT o2 = __shfl(P2, idx & 63, 64);
                    if (idx >= 64) o[a] = o2;
                }
            }
            W[3 + q] = o[0] * wx[0] + o[1] * wx[1] + o[2] * wx[2] + o[3] * wx[3];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) { va[c] = na[VYZ_NB - 1][c]; vb[c] = nb[VYZ_NB - 1][c]; }
        cs += VYZ_NB;
    };
    // W[0 .. nv-1] hold X at coarse indices lo .. lo + nv - 1; the window of
    // the refined point s is W[0 .. 3] once lo has reached its first source
    batch();
#pragma unroll
    for (int q = 0; q < 3; ++q) W[q] = W[3 + q];
#pragma unroll
    for (int q = 3; q < VYZ_NB; ++q) W[q] = W[3 + q];
    int lo = sS[0] - 1, nv = VYZ_NB;
#pragma unroll 1
    for (int s = 0; s < nS - 1; ++s) {
        const int b0 = sS[s] - 1;
        while (lo < b0) {
#pragma unroll
            for (int q = 0; q < VYZ_NB + 2; ++q) W[q] = W[q + 1];
            ++lo;
            if (--nv == 3) { batch(); nv += VYZ_NB; }
        }
        const RT *w = wS + 4 * s;
        const RT tv = W[0] * w[0] + W[1] * w[1] + W[2] * w[2] + W[3] * w[3];
        v = v + tv * dl;
        if (in) outp[obase + (s + 1) * sstride + kr] = v;
    }
}

// Refined -> coarse cell centres through the fast coarse-grid loop of
// interpolate_xyz_to_coarse_fast (HermiteInterpolations.F90:291-310):
// InterpSalMultigrid (salinity.F90:458-491), InterpPhiMultigrid and the
// transfer of AddLatentHeat (phasefield.F90:470-494, :651-725).
// Tile: 64 coarse x (lanes) x SALC_TJ coarse y (two rows per wave), marching
// through SALC_TZ coarse z planes.  Each refined z plane the tile needs is
// staged once in LDS (with the wall ghosts of salinity.F90:466-482, double
// buffered), contracted over x and y, and the last four contracted planes are
// kept in registers for the z contraction.
constexpr int SALC_TJ = 8, SALC_TZ = 16;
constexpr int SALC_BU = 4;      // plane items loaded per batch and thread
constexpr int SALC_IN = -(1 << 30);   // marks an interior point of a staged plane
// What a refined-to-coarse tile interpolates and where the result goes.
enum CoarseMode {
    CM_SAL = 0,     // S with the wall ghosts of salinity.F90:466-482 -> salc (salc(nx) = 0)
    CM_PHI = 1,     // phi with zero-gradient ghosts (phasefield.F90:479-487) -> phic (phic(nx) = 0)
    CM_LAT = 2      // d phi with zero-gradient ghosts (:674-683) -> out (+)= scale * value
};
// Per axis and coarse index: the first of the 4 refined sources (1-based
// reference index: x may name the wall ghosts 0 and nxr or the zero entries
// beyond them, y/z may name periodic images) and its 4 weights (RT).  The
// bracketing (fast) path of interpolate_xyz_to_coarse_fast and the general
// path of interpolate_xyz_to_coarse differ only in these tables.
struct CoarseJob {
    int mode;
    const void *src;            // refined field (RT)
    void *out;                  // coarse field (RT for CM_SAL, double otherwise)
    const int *x_src, *y_src, *z_src;
    const void *x_w, *y_w, *z_w;
    double scale;
    int accumulate;
};
template <typename RT>
__device__ __forceinline__ RT coarse_src(const KP &p, const CoarseJob &jb, const RT *f, int kr, int jj, int ii) {
    const long long o2 = (long long)ii * p.planer + (long long)jj * p.ldxr;
    if (kr >= 1 && kr <= p.nxmr) return f[o2 + kr - 1];
    if (kr == 0) {
        const RT s1 = f[o2];
        if (jb.mode != CM_SAL) return s1;
        return p.sfixs == 1 ? RT(2.0 * p.salbp[(long long)ii * p.nyr + jj]) - s1 : s1;
    }
    if (kr == p.nxr) {
        const RT sn = f[o2 + p.nxmr - 1];
        if (jb.mode != CM_SAL) return sn;
        return p.sfixn == 1 ? RT(2.0 * p.saltp[(long long)ii * p.nyr + jj]) - sn : sn;
    }
    return RT(0);
}
template <typename RT>
__global__ void __launch_bounds__(256) k_interp_coarse_tile(KP p, CoarseJob jb, int xrmax, int yrmax) {
    extern __shared__ __align__(16) unsigned char smraw[];
    RT *sm = reinterpret_cast<RT *>(smraw);
    const RT *src = static_cast<const RT *>(jb.src);
    const int *x_src = jb.x_src;
    const RT *xb_w = static_cast<const RT *>(jb.x_w);
    const RT *yb_w = static_cast<const RT *>(jb.y_w), *zb_w = static_cast<const RT *>(jb.z_w);
    const int *yb_src = jb.y_src, *zb_src = jb.z_src;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, t = threadIdx.x;
    const int nxt = (p.nxm + 63) / 64;
    int tx, ty, tz;
    if (!tile_3d(nxt, (p.ny + SALC_TJ - 1) / SALC_TJ, (p.nz + SALC_TZ - 1) / SALC_TZ, tx, ty, tz)) return;
    const int K0 = tx * 64, K1 = min(K0 + 64, p.nxm);
    const int J0 = ty * SALC_TJ, nj = min(SALC_TJ, p.ny - J0);
    const int I0 = tz * SALC_TZ, I1 = min(I0 + SALC_TZ, p.nz);
    const int x0 = x_src[K0], XR = x_src[K1 - 1] + 4 - x0;                  // reference refined x index
    const int y0 = yb_src[J0] - 1, YR = yb_src[J0 + nj - 1] - 1 + 4 - y0;   // 0-based refined y
    const int zfirst = zb_src[I0] - 1, zlast = zb_src[I1 - 1] - 1 + 3;      // 0-based refined z (unwrapped)
    const int k = K0 + lane;
    const bool kok = k < K1;
    RT wx[4] = {0, 0, 0, 0};
    int xa = 0;
    if (kok) { xa = x_src[k] - x0; for (int a = 0; a < 4; ++a) wx[a] = xb_w[4 * k + a]; }
    // two coarse rows per wave: tj = wave and wave + 4
    int ya[2] = {0, 0};
    RT wy[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    bool rok[2];
    for (int h = 0; h < 2; ++h) {
        const int tj = wave + 4 * h;
        rok[h] = tj < nj;
        if (rok[h]) { ya[h] = yb_src[J0 + tj] - 1 - y0; for (int b = 0; b < 4; ++b) wy[h][b] = yb_w[4 * (J0 + tj) + b]; }
    }
    RT ring[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};   // contracted refined planes zr-3..zr
    const int plane = XR * YR;
    // refined y rows of the tile (periodic): their indices and row offsets
    __shared__ int rowj[64];
    for (int q = t; q < YR; q += 256) rowj[q] = wrapf(y0 + q, p.nyr);
    __syncthreads();
    int ic = I0;
    for (int zr = zfirst; zr <= zlast; ++zr) {
        RT *B = sm + ((zr - zfirst) & 1) * xrmax * yrmax;     // [yr][xr], double buffered
        const int ii = wrapn(zr, p.nzr);
        const RT *fz = src + (long long)ii * p.planer;
        // SALC_BU loads per thread in flight: interior points read first
        // (an item outside the plane reads a valid address), wall ghosts and
        // points beyond them are filled in afterwards
        Walk3 w(t, 256, XR, YR);
        for (int q0 = t; q0 < plane; q0 += 256 * SALC_BU) {
            RT v[SALC_BU];
            int dst[SALC_BU], kx[SALC_BU], jy[SALC_BU];
#pragma unroll
            for (int r = 0; r < SALC_BU; ++r) {
                const int kr = x0 + w.a, jj = rowj[min(w.b, YR - 1)];
                const bool live = q0 + 256 * r < plane, in = live && kr >= 1 && kr <= p.nxmr;
                v[r] = fz[(long long)jj * p.ldxr + (in ? kr - 1 : 0)];
                dst[r] = live ? __mul24(w.b, xrmax) + w.a : -1;
                kx[r] = in ? SALC_IN : kr;
                jy[r] = jj;
                w.next();
            }
#pragma unroll
            for (int r = 0; r < SALC_BU; ++r)
                if (dst[r] >= 0) B[dst[r]] = kx[r] == SALC_IN ? v[r] : coarse_src<RT>(p, jb, src, kx[r], jy[r], ii);
        }
        __syncthreads();
        for (int h = 0; h < 2; ++h) {
            RT v = RT(0);
            if (kok && rok[h]) {
                for (int b = 0; b < 4; ++b) {
                    const RT *row = B + __mul24(ya[h] + b, xrmax) + xa;
                    const RT q1 = row[0] * wx[0] + row[1] * wx[1] + row[2] * wx[2] + row[3] * wx[3];
                    v = b == 0 ? q1 * wy[h][0] : v + q1 * wy[h][b];
                }
            }
            ring[h][0] = ring[h][1]; ring[h][1] = ring[h][2]; ring[h][2] = ring[h][3]; ring[h][3] = v;
        }
        // every coarse plane whose 4-plane stencil ends at zr is complete
        while (ic < I1 && zb_src[ic] - 1 + 3 == zr) {
            const RT *wz = zb_w + 4 * ic;
            for (int h = 0; h < 2; ++h)
                if (kok && rok[h]) {
                    const RT v = ring[h][0] * wz[0] + ring[h][1] * wz[1] + ring[h][2] * wz[2] + ring[h][3] * wz[3];
                    const long long o = bidx(p, k, J0 + wave + 4 * h, ic);
                    if (jb.mode == CM_SAL) static_cast<RT *>(jb.out)[o] = v;
                    else {
                        double *out = static_cast<double *>(jb.out);
                        const double val = jb.mode == CM_LAT ? jb.scale * double(v) : double(v);
                        out[o] = jb.accumulate ? out[o] + val : val;
                    }
                }
            ++ic;
        }
    }
    // salc(nx), phic(nx) stay zero (salinity.F90:463, phasefield.F90:476): written by the last x tile
    if (jb.mode != CM_LAT && tx == nxt - 1 && lane == 0)
        for (int h = 0; h < 2; ++h)
            if (rok[h])
                for (int iz = I0; iz < I1; ++iz) {
                    const long long o = bidx(p, p.nxm, J0 + wave + 4 * h, iz);
                    if (jb.mode == CM_SAL) static_cast<RT *>(jb.out)[o] = RT(0);
                    else static_cast<double *>(jb.out)[o] = 0.0;
                }
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
template <typename FT>
__global__ void __launch_bounds__(256) k_init_scalar(FT *f, const double *bot, const double *top, const double *xm,
                                                     int nxm, int nxf, int ldx, int ny, int nz, double alx3,
                                                     double eps, unsigned long long seed) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= nxf || j >= ny) return;
    const long long o = ((long long)i * ny + j) * ldx + k;
    if (k >= nxm) { f[o] = FT(0); return; }
    const long long p2 = (long long)i * ny + j;
    const double x = xm[k];
    double v = bot[p2] + (top[p2] - bot[p2]) * x / alx3;
    const double r = splitmix_uniform(seed, (unsigned long long)(((long long)i * ny + j) * nxf + k));
    const double amp = (fabs(x - 0.5) + eps > 0.5) ? 0.5 - fabs(x - 0.5) : eps;
    f[o] = FT(v + amp * (2.0 * r - 1.0));
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// CalcMaxCFL.F90:24-44 (what=0) / CheckDivergence.F90:24-43 (what=1) on the
// base grid, slot 0; refined grid versions write slot 1.
__global__ void __launch_bounds__(256) k_cfl_div(KP p, int what) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    double v = 0.0;
    for (int i = blockIdx.z; i < p.nz; i += gridDim.z) {
    if (k < p.nxm && j < p.ny) {
        const int jp = wrapp(j, p.ny), ip = wrapp(i, p.nz);
        double w;
        if (what == 0)
            w = fabs((p.vz[bidx(p, k, j, i)] + p.vz[bidx(p, k, j, ip)]) * 0.5 * p.dz)
              + fabs((p.vy[bidx(p, k, j, i)] + p.vy[bidx(p, k, jp, i)]) * 0.5 * p.dy)
              + fabs((p.vx[bidx(p, k, j, i)] + p.vx[bidx(p, k + 1, j, i)]) * 0.5 * p.udx3m[k]);
        else
            w = fabs((p.vz[bidx(p, k, j, ip)] - p.vz[bidx(p, k, j, i)]) * p.dz
                   + (p.vy[bidx(p, k, jp, i)] - p.vy[bidx(p, k, j, i)]) * p.dy
                   + (p.vx[bidx(p, k + 1, j, i)] - p.vx[bidx(p, k, j, i)]) * p.udx3m[k]);
        v = fmax(v, w);
    }
    }
    v = block_max(v);
    if (threadIdx.x == 0 && threadIdx.y == 0) atomicMax(&p.redmax[2 * what], (unsigned long long)__double_as_longlong(v));
}
template <typename RT>
__global__ void __launch_bounds__(256) k_cfl_div_r(KP p, int what) {
    const int k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y;
    const RT *vxr = sel<RT>(p.vxr, p.vxrf), *vyr = sel<RT>(p.vyr, p.vyrf), *vzr = sel<RT>(p.vzr, p.vzrf);
    double v = 0.0;
    for (int i = blockIdx.z; i < p.nzr; i += gridDim.z) {
    if (k < p.nxmr && j < p.nyr) {
        const int jp = wrapp(j, p.nyr), ip = wrapp(i, p.nzr);
        double w;
        if (what == 0)
            w = fabs((double(vzr[ridx(p, k, j, i)]) + double(vzr[ridx(p, k, j, ip)])) * 0.5 * p.dzr)
              + fabs((double(vyr[ridx(p, k, j, i)]) + double(vyr[ridx(p, k, jp, i)])) * 0.5 * p.dyr)
              + fabs((double(vxr[ridx(p, k, j, i)]) + double(vxr[ridx(p, k + 1, j, i)])) * 0.5 * p.udx3mr[k]);
        else
            w = fabs((double(vzr[ridx(p, k, j, ip)]) - double(vzr[ridx(p, k, j, i)])) * p.dzr
                   + (double(vyr[ridx(p, k, jp, i)]) - double(vyr[ridx(p, k, j, i)])) * p.dyr
                   + (double(vxr[ridx(p, k + 1, j, i)]) - double(vxr[ridx(p, k, j, i)])) * p.udx3mr[k]);
        v = fmax(v, w);
    }
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
