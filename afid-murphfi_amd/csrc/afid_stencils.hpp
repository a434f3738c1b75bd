// afid_stencils.hpp -- explicit and implicit-RHS stencils on the staggered
// grid, written once as device functions and used both by the per-routine
// kernels (one per reference subroutine) and by the fused column kernels.
//
// A stencil reads a wall-normal column and its y/z neighbour columns through
// a set of column base pointers (struct Nb / NbS).  In the column kernels
// those pointers are wavefront-uniform, so they live in scalar registers and
// every load is a scalar base plus a 32-bit lane offset.  Indices are 0-based:
// k = kc-1 (wall-normal); the arithmetic follows the reference term by term.
#pragma once
#include <hip/hip_runtime.h>
#include "afid_params.hpp"

namespace afid {

__device__ __forceinline__ long long bidx(const KP &p, int k, int j, int i) {
    return (long long)i * p.plane + (long long)j * p.ldx + k;
}
__device__ __forceinline__ long long ridx(const KP &p, int k, int j, int i) {
    return (long long)i * p.planer + (long long)j * p.ldxr + k;
}
__device__ __forceinline__ int wrapm(int j, int n) { return j == 0 ? n - 1 : j - 1; }
__device__ __forceinline__ int wrapp(int j, int n) { return j == n - 1 ? 0 : j + 1; }

// column (j,i) of the base grid and the neighbour columns its stencils read
struct Nb {
    const double *vx, *vx_jm, *vx_jp, *vx_im, *vx_ip;
    const double *vy, *vy_jm, *vy_jp, *vy_im, *vy_ip, *vy_jp_im;
    const double *vz, *vz_jm, *vz_jp, *vz_im, *vz_ip, *vz_jm_ip;
    const double *t, *t_jm, *t_jp, *t_im, *t_ip;
    const double *pr, *pr_jm, *pr_im;
    double tbp, ttp;
};
__device__ __forceinline__ Nb make_nb(const KP &p, int j, int i) {
    const int jm = wrapm(j, p.ny), jp = wrapp(j, p.ny), im = wrapm(i, p.nz), ip = wrapp(i, p.nz);
    auto o = [&](int jj, int ii) { return (long long)ii * p.plane + (long long)jj * p.ldx; };
    const long long c = o(j, i), cjm = o(jm, i), cjp = o(jp, i), cim = o(j, im), cip = o(j, ip);
    Nb n;
    n.vx = p.vx + c; n.vx_jm = p.vx + cjm; n.vx_jp = p.vx + cjp; n.vx_im = p.vx + cim; n.vx_ip = p.vx + cip;
    n.vy = p.vy + c; n.vy_jm = p.vy + cjm; n.vy_jp = p.vy + cjp; n.vy_im = p.vy + cim; n.vy_ip = p.vy + cip;
    n.vy_jp_im = p.vy + o(jp, im);
    n.vz = p.vz + c; n.vz_jm = p.vz + cjm; n.vz_jp = p.vz + cjp; n.vz_im = p.vz + cim; n.vz_ip = p.vz + cip;
    n.vz_jm_ip = p.vz + o(jm, ip);
    n.t = p.temp + c; n.t_jm = p.temp + cjm; n.t_jp = p.temp + cjp; n.t_im = p.temp + cim; n.t_ip = p.temp + cip;
    n.pr = p.pr + c; n.pr_jm = p.pr + cjm; n.pr_im = p.pr + cim;
    n.tbp = p.tempbp[(long long)i * p.ny + j];
    n.ttp = p.temptp[(long long)i * p.ny + j];
    return n;
}

// ExplicitTermsVX.F90:42-127, valid for k = 1..nxm-1 (kc = 2..nxm)
__device__ __forceinline__ double ex_vx(const KP &p, const Nb &n, unsigned k) {
    const double udy = p.dy * 0.25, udz = p.dz * 0.25, udyq = p.dyq / p.ren, udzq = p.dzq / p.ren;
    const double rkm = p.rudx3m[k - 1], rkc = p.rudx3m[k], u3 = p.udx3c[k];
    const double vxc = n.vx[k];
    const double vxzp = n.vx_ip[k], vxzm = n.vx_im[k], vxyp = n.vx_jp[k], vxym = n.vx_jm[k];
    double hxz = u3 * udz * ((rkm * n.vz_ip[k - 1] + rkc * n.vz_ip[k]) * (vxzp + vxc)
                           - (rkm * n.vz[k - 1] + rkc * n.vz[k]) * (vxc + vxzm));
    double hxy = u3 * udy * ((rkm * n.vy_jp[k - 1] + rkc * n.vy_jp[k]) * (vxyp + vxc)
                           - (rkm * n.vy[k - 1] + rkc * n.vy[k]) * (vxc + vxym));
    const double a = n.vx[k + 1] + vxc, b = vxc + n.vx[k - 1];
    double hxx = (a * a - b * b) * u3 * 0.25;
    double dzzvx = (vxzm - 2.0 * vxc + vxzp) * udzq;
    double dyyvx = (vxym - 2.0 * vxc + vxyp) * udyq;
    double q = -(hxx + hxy + hxz) + dyyvx + dzzvx;
    if (p.gaxis == 1) q += p.byct * (p.active_t * 0.5 * (n.t[k] + n.t[k - 1]));
    return q;
}

// ExplicitTermsVY.F90:50-118, k = 0..nxm-1
__device__ __forceinline__ double ex_vy(const KP &p, const Nb &n, unsigned k) {
    const double udyq = p.dyq / p.ren, udzq = p.dzq / p.ren, udy = p.dy * 0.25, udz = p.dz * 0.25;
    const unsigned kmm = k == 0 ? 0 : k - 1, kpp = k == unsigned(p.nxm - 1) ? k : k + 1;
    const double vyc = n.vy[k];
    const double vyyp = n.vy_jp[k], vyym = n.vy_jm[k], vyzp = n.vy_ip[k], vyzm = n.vy_im[k];
    double hyx = ((n.vx[k + 1] + n.vx_jm[k + 1]) * (n.vy[kpp] + vyc)
                - (n.vx[k] + n.vx_jm[k]) * (vyc + n.vy[kmm])) * p.udx3m[k] * 0.25;
    const double a = vyyp + vyc, b = vyym + vyc;
    double hyy = (a * a - b * b) * udy;
    double hyz = ((vyzp + vyc) * (n.vz_ip[k] + n.vz_jm_ip[k]) - (vyc + vyzm) * (n.vz[k] + n.vz_jm[k])) * udz;
    double dyyvy = (vyyp - 2.0 * vyc + vyym) * udyq;
    double dzzvy = (vyzp - 2.0 * vyc + vyzm) * udzq;
    double q = -(hyx + hyy + hyz) + dyyvy + dzzvy + p.Gy;
    if (p.gaxis == 2) q += p.byct * (p.active_t * 0.5 * (n.t[k] + n.t_jm[k]));
    return q;
}

// ExplicitTermsVZ.F90:45-114, k = 0..nxm-1
__device__ __forceinline__ double ex_vz(const KP &p, const Nb &n, unsigned k) {
    const double udyq = p.dyq / p.ren, udzq = p.dzq / p.ren, udy = p.dy * 0.25, udz = p.dz * 0.25;
    const unsigned kmm = k == 0 ? 0 : k - 1, kpp = k == unsigned(p.nxm - 1) ? k : k + 1;
    const double vzc = n.vz[k];
    const double vzzp = n.vz_ip[k], vzzm = n.vz_im[k], vzyp = n.vz_jp[k], vzym = n.vz_jm[k];
    const double a = vzzp + vzc, b = vzzm + vzc;
    double hzz = (a * a - b * b) * udz;
    double hzy = ((n.vy_jp[k] + n.vy_jp_im[k]) * (vzyp + vzc) - (n.vy[k] + n.vy_im[k]) * (vzc + vzym)) * udy;
    double hzx = ((n.vx[k + 1] + n.vx_im[k + 1]) * (n.vz[kpp] + vzc)
                - (n.vx[k] + n.vx_im[k]) * (vzc + n.vz[kmm])) * p.udx3m[k] * 0.25;
    double dzzvz = (vzzp - 2.0 * vzc + vzzm) * udzq;
    double dyyvz = (vzyp - 2.0 * vzc + vzym) * udyq;
    double q = -(hzx + hzy + hzz) + dyyvz + dzzvz;
    if (p.gaxis == 3) q += p.byct * (p.active_t * 0.5 * (n.t[k] + n.t_im[k]));
    return q;
}

// ExplicitTermsTemp.F90:39-77, k = 0..nxm-1
__device__ __forceinline__ double ex_temp(const KP &p, const Nb &n, unsigned k) {
    const double udz = p.dz * 0.5, udy = p.dy * 0.5, udzq = p.dzq / p.pect, udyq = p.dyq / p.pect;
    const double tc = n.t[k], tzp = n.t_ip[k], tzm = n.t_im[k], typ = n.t_jp[k], tym = n.t_jm[k];
    const double vxp = n.vx[k + 1], vxc = n.vx[k];
    double htx;
    if (k == 0)
        htx = (vxp * (n.t[k + 1] + tc) - vxc * 2.0 * n.tbp) * 0.5 * p.udx3m[k];
    else if (k == unsigned(p.nxm - 1))
        htx = (vxp * 2.0 * n.ttp - vxc * (tc + n.t[k - 1])) * 0.5 * p.udx3m[k];
    else
        htx = (vxp * (n.t[k + 1] + tc) - vxc * (tc + n.t[k - 1])) * 0.5 * p.udx3m[k];
    double htz = (n.vz_ip[k] * (tzp + tc) - n.vz[k] * (tc + tzm)) * udz;
    double hty = (n.vy_jp[k] * (typ + tc) - n.vy[k] * (tc + tym)) * udy;
    double dzzt = (tzp - 2.0 * tc + tzm) * udzq;
    double dyyt = (typ - 2.0 * tc + tym) * udyq;
    return -(htx + hty + htz) + dyyt + dzzt;
}

// x-second-derivative parts of the implicit RHS (ImplicitAndUpdate*.F90)
__device__ __forceinline__ double rhs_vx(const KP &p, const Nb &n, double h, double ru, unsigned k) {
    // ImplicitAndUpdateVX.F90:37-66, k = 1..nxm-1
    double dxxvx = n.vx[k + 1] * p.ap3ck[k] + n.vx[k] * p.ac3ck[k] + n.vx[k - 1] * p.am3ck[k];
    double dxp = (n.pr[k] - n.pr[k - 1]) * (p.al * p.udx3c[k]);
    return (p.ga * h + p.ro * ru + (p.al / p.ren) * dxxvx - dxp) * p.dt;
}
__device__ __forceinline__ double rhs_vy(const KP &p, const Nb &n, double h, double ru, unsigned k) {
    // ImplicitAndUpdateVY.F90:41-82
    const double amm = p.am3sk[k], acc = p.ac3sk[k], app = p.ap3sk[k];
    const unsigned kmm = k == 0 ? 0 : k - 1, kpp = k == unsigned(p.nxm - 1) ? k : k + 1;
    double dxxvy;
    if (k == 0)
        dxxvy = n.vy[kpp] * app + n.vy[k] * acc - (acc + app) * p.xminusu * p.inslws;
    else if (k == unsigned(p.nxm - 1))
        dxxvy = -(acc + amm) * p.xplusu * p.inslwn + n.vy[k] * acc + n.vy[kmm] * amm;
    else
        dxxvy = n.vy[kpp] * app + n.vy[k] * acc + n.vy[kmm] * amm;
    double dyp = (n.pr[k] - n.pr_jm[k]) * (p.dy * p.al);
    return (p.ga * h + p.ro * ru + (p.al / p.ren) * dxxvy - dyp) * p.dt;
}
__device__ __forceinline__ double rhs_vz(const KP &p, const Nb &n, double h, double ru, unsigned k) {
    // ImplicitAndUpdateVZ.F90:38-70
    const unsigned kmm = k == 0 ? 0 : k - 1, kpp = k == unsigned(p.nxm - 1) ? k : k + 1;
    double dxxvz = n.vz[kpp] * p.ap3sk[k] + n.vz[k] * p.ac3sk[k] + n.vz[kmm] * p.am3sk[k];
    double dzp = (n.pr[k] - n.pr_im[k]) * p.dz * p.al;
    return (p.ga * h + p.ro * ru + (p.al / p.ren) * dxxvz - dzp) * p.dt;
}
__device__ __forceinline__ double rhs_temp(const KP &p, const Nb &n, double h, double ru, unsigned k) {
    // ImplicitAndUpdateTemp.F90:33-62
    double dxxt;
    if (k == 0)
        dxxt = n.t[k + 1] * p.ap3ssk[k] + n.t[k] * p.ac3ssk[k] - (p.ap3ssk[k] + p.ac3ssk[k]) * n.tbp * p.tfixs;
    else if (k == unsigned(p.nxm - 1))
        dxxt = n.t[k] * p.ac3ssk[k] + n.t[k - 1] * p.am3ssk[k] - (p.am3ssk[k] + p.ac3ssk[k]) * n.ttp * p.tfixn;
    else
        dxxt = n.t[k + 1] * p.ap3ssk[k] + n.t[k] * p.ac3ssk[k] + n.t[k - 1] * p.am3ssk[k];
    return (p.ga * h + p.ro * ru + (p.al / p.pect) * dxxt) * p.dt;
}

// refined column (j,i) and neighbours for the salinity stencils
struct NbS {
    const double *s, *s_jm, *s_jp, *s_im, *s_ip;
    const double *vxr, *vyr, *vyr_jp, *vzr, *vzr_ip;
    double sbp, stp;
};
__device__ __forceinline__ NbS make_nbs(const KP &p, int j, int i) {
    const int jm = wrapm(j, p.nyr), jp = wrapp(j, p.nyr), im = wrapm(i, p.nzr), ip = wrapp(i, p.nzr);
    auto o = [&](int jj, int ii) { return (long long)ii * p.planer + (long long)jj * p.ldxr; };
    const long long c = o(j, i);
    NbS n;
    n.s = p.sal + c; n.s_jm = p.sal + o(jm, i); n.s_jp = p.sal + o(jp, i); n.s_im = p.sal + o(j, im); n.s_ip = p.sal + o(j, ip);
    n.vxr = p.vxr + c; n.vyr = p.vyr + c; n.vyr_jp = p.vyr + o(jp, i); n.vzr = p.vzr + c; n.vzr_ip = p.vzr + o(j, ip);
    n.sbp = p.salbp[(long long)i * p.nyr + j];
    n.stp = p.saltp[(long long)i * p.nyr + j];
    return n;
}

// ExplicitSalinity, salinity.F90:298-363 (non-IBM), refined k = 0..nxmr-1
__device__ __forceinline__ double ex_sal(const KP &p, const NbS &n, unsigned k) {
    const double udyr = 0.5 * p.dyr, udzr = 0.5 * p.dzr, udyrq = p.dyqr / p.pecs, udzrq = p.dzqr / p.pecs;
    const double sc = n.s[k], syp = n.s_jp[k], sym = n.s_jm[k], szp = n.s_ip[k], szm = n.s_im[k];
    const double vp = n.vxr[k + 1], vc = n.vxr[k];
    double hsx;
    if (k == 0)
        hsx = (vp * (n.s[k + 1] + sc) - vc * 2.0 * n.sbp) * p.udx3mr[k] * 0.5;
    else if (k == unsigned(p.nxmr - 1))
        hsx = (vp * 2.0 * n.stp - vc * (sc + n.s[k - 1])) * p.udx3mr[k] * 0.5;
    else
        hsx = (vp * (n.s[k + 1] + sc) - vc * (sc + n.s[k - 1])) * p.udx3mr[k] * 0.5;
    double hsy = (n.vyr_jp[k] * (syp + sc) - n.vyr[k] * (sc + sym)) * udyr;
    double hsz = (n.vzr_ip[k] * (szp + sc) - n.vzr[k] * (sc + szm)) * udzr;
    double dyys = (syp - 2.0 * sc + sym) * udyrq;
    double dzzs = (szp - 2.0 * sc + szm) * udzrq;
    return -(hsx + hsy + hsz) + dyys + dzzs;
}
// ImplicitSalinity, salinity.F90:378-400
__device__ __forceinline__ double rhs_sal(const KP &p, const NbS &n, double h, double ru, unsigned k) {
    double dxxs;
    if (k == 0)
        dxxs = n.s[k + 1] * p.ap3sskr[k] + n.s[k] * p.ac3sskr[k] - (p.ap3sskr[k] + p.ac3sskr[k]) * n.sbp * p.sfixs;
    else if (k == unsigned(p.nxmr - 1))
        dxxs = n.s[k] * p.ac3sskr[k] + n.s[k - 1] * p.am3sskr[k] - (p.am3sskr[k] + p.ac3sskr[k]) * n.stp * p.sfixn;
    else
        dxxs = n.s[k + 1] * p.ap3sskr[k] + n.s[k] * p.ac3sskr[k] + n.s[k - 1] * p.am3sskr[k];
    return (p.ga * h + p.ro * ru + (p.al / p.pecs) * dxxs) * p.dt;
}

}  // namespace afid
