// afid_stencils.hpp -- explicit and implicit-RHS stencils on the staggered
// grid, written once as device functions and used both by the per-routine
// kernels (one per reference subroutine) and by the fused column kernels.
//
// Indices are 0-based: k = kc-1 (wall-normal), j, i periodic.  jm/jp/im/ip
// are the wrapped neighbours.  Arithmetic follows the reference term by term.
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

// ExplicitTermsVX.F90:42-127, valid for k = 1..nxm-1 (kc = 2..nxm)
__device__ __forceinline__ double ex_vx(const KP &p, int k, int j, int jm, int jp, int i, int im, int ip) {
    const double *vx = p.vx, *vy = p.vy, *vz = p.vz;
    const double udy = p.dy * 0.25, udz = p.dz * 0.25, udyq = p.dyq / p.ren, udzq = p.dzq / p.ren;
    const double rkm = p.rudx3m[k - 1], rkc = p.rudx3m[k], u3 = p.udx3c[k];
    const double vxc = vx[bidx(p, k, j, i)];
    const double vxzp = vx[bidx(p, k, j, ip)], vxzm = vx[bidx(p, k, j, im)];
    const double vxyp = vx[bidx(p, k, jp, i)], vxym = vx[bidx(p, k, jm, i)];
    double hxz = u3 * udz * ((rkm * vz[bidx(p, k - 1, j, ip)] + rkc * vz[bidx(p, k, j, ip)]) * (vxzp + vxc)
                           - (rkm * vz[bidx(p, k - 1, j, i)] + rkc * vz[bidx(p, k, j, i)]) * (vxc + vxzm));
    double hxy = u3 * udy * ((rkm * vy[bidx(p, k - 1, jp, i)] + rkc * vy[bidx(p, k, jp, i)]) * (vxyp + vxc)
                           - (rkm * vy[bidx(p, k - 1, j, i)] + rkc * vy[bidx(p, k, j, i)]) * (vxc + vxym));
    const double a = vx[bidx(p, k + 1, j, i)] + vxc, b = vxc + vx[bidx(p, k - 1, j, i)];
    double hxx = (a * a - b * b) * u3 * 0.25;
    double dzzvx = (vxzm - 2.0 * vxc + vxzp) * udzq;
    double dyyvx = (vxym - 2.0 * vxc + vxyp) * udyq;
    double q = -(hxx + hxy + hxz) + dyyvx + dzzvx;
    if (p.gaxis == 1) q += p.byct * (p.active_t * 0.5 * (p.temp[bidx(p, k, j, i)] + p.temp[bidx(p, k - 1, j, i)]));
    return q;
}

// ExplicitTermsVY.F90:50-118, k = 0..nxm-1
__device__ __forceinline__ double ex_vy(const KP &p, int k, int j, int jm, int jp, int i, int im, int ip) {
    const double *vx = p.vx, *vy = p.vy, *vz = p.vz;
    const double udyq = p.dyq / p.ren, udzq = p.dzq / p.ren, udy = p.dy * 0.25, udz = p.dz * 0.25;
    const int kmm = k == 0 ? 0 : k - 1, kpp = k == p.nxm - 1 ? k : k + 1;
    const double vyc = vy[bidx(p, k, j, i)];
    const double vyyp = vy[bidx(p, k, jp, i)], vyym = vy[bidx(p, k, jm, i)];
    const double vyzp = vy[bidx(p, k, j, ip)], vyzm = vy[bidx(p, k, j, im)];
    double hyx = ((vx[bidx(p, k + 1, j, i)] + vx[bidx(p, k + 1, jm, i)]) * (vy[bidx(p, kpp, j, i)] + vyc)
                - (vx[bidx(p, k, j, i)] + vx[bidx(p, k, jm, i)]) * (vyc + vy[bidx(p, kmm, j, i)])) * p.udx3m[k] * 0.25;
    const double a = vyyp + vyc, b = vyym + vyc;
    double hyy = (a * a - b * b) * udy;
    double hyz = ((vyzp + vyc) * (vz[bidx(p, k, j, ip)] + vz[bidx(p, k, jm, ip)])
                - (vyc + vyzm) * (vz[bidx(p, k, j, i)] + vz[bidx(p, k, jm, i)])) * udz;
    double dyyvy = (vyyp - 2.0 * vyc + vyym) * udyq;
    double dzzvy = (vyzp - 2.0 * vyc + vyzm) * udzq;
    double q = -(hyx + hyy + hyz) + dyyvy + dzzvy + p.Gy;
    if (p.gaxis == 2) q += p.byct * (p.active_t * 0.5 * (p.temp[bidx(p, k, j, i)] + p.temp[bidx(p, k, jm, i)]));
    return q;
}

// ExplicitTermsVZ.F90:45-114, k = 0..nxm-1
__device__ __forceinline__ double ex_vz(const KP &p, int k, int j, int jm, int jp, int i, int im, int ip) {
    const double *vx = p.vx, *vy = p.vy, *vz = p.vz;
    const double udyq = p.dyq / p.ren, udzq = p.dzq / p.ren, udy = p.dy * 0.25, udz = p.dz * 0.25;
    const int kmm = k == 0 ? 0 : k - 1, kpp = k == p.nxm - 1 ? k : k + 1;
    const double vzc = vz[bidx(p, k, j, i)];
    const double vzzp = vz[bidx(p, k, j, ip)], vzzm = vz[bidx(p, k, j, im)];
    const double vzyp = vz[bidx(p, k, jp, i)], vzym = vz[bidx(p, k, jm, i)];
    const double a = vzzp + vzc, b = vzzm + vzc;
    double hzz = (a * a - b * b) * udz;
    double hzy = ((vy[bidx(p, k, jp, i)] + vy[bidx(p, k, jp, im)]) * (vzyp + vzc)
                - (vy[bidx(p, k, j, i)] + vy[bidx(p, k, j, im)]) * (vzc + vzym)) * udy;
    double hzx = ((vx[bidx(p, k + 1, j, i)] + vx[bidx(p, k + 1, j, im)]) * (vz[bidx(p, kpp, j, i)] + vzc)
                - (vx[bidx(p, k, j, i)] + vx[bidx(p, k, j, im)]) * (vzc + vz[bidx(p, kmm, j, i)])) * p.udx3m[k] * 0.25;
    double dzzvz = (vzzp - 2.0 * vzc + vzzm) * udzq;
    double dyyvz = (vzyp - 2.0 * vzc + vzym) * udyq;
    double q = -(hzx + hzy + hzz) + dyyvz + dzzvz;
    if (p.gaxis == 3) q += p.byct * (p.active_t * 0.5 * (p.temp[bidx(p, k, j, i)] + p.temp[bidx(p, k, j, im)]));
    return q;
}

// ExplicitTermsTemp.F90:39-77, k = 0..nxm-1
__device__ __forceinline__ double ex_temp(const KP &p, int k, int j, int jm, int jp, int i, int im, int ip) {
    const double *vx = p.vx, *vy = p.vy, *vz = p.vz, *t = p.temp;
    const double udz = p.dz * 0.5, udy = p.dy * 0.5, udzq = p.dzq / p.pect, udyq = p.dyq / p.pect;
    const double tc = t[bidx(p, k, j, i)];
    const double tzp = t[bidx(p, k, j, ip)], tzm = t[bidx(p, k, j, im)];
    const double typ = t[bidx(p, k, jp, i)], tym = t[bidx(p, k, jm, i)];
    const double vxp = vx[bidx(p, k + 1, j, i)], vxc = vx[bidx(p, k, j, i)];
    double htx;
    if (k == 0)
        htx = (vxp * (t[bidx(p, k + 1, j, i)] + tc) - vxc * 2.0 * p.tempbp[(long long)i * p.ny + j]) * 0.5 * p.udx3m[k];
    else if (k == p.nxm - 1)
        htx = (vxp * 2.0 * p.temptp[(long long)i * p.ny + j] - vxc * (tc + t[bidx(p, k - 1, j, i)])) * 0.5 * p.udx3m[k];
    else
        htx = (vxp * (t[bidx(p, k + 1, j, i)] + tc) - vxc * (tc + t[bidx(p, k - 1, j, i)])) * 0.5 * p.udx3m[k];
    double htz = (vz[bidx(p, k, j, ip)] * (tzp + tc) - vz[bidx(p, k, j, i)] * (tc + tzm)) * udz;
    double hty = (vy[bidx(p, k, jp, i)] * (typ + tc) - vy[bidx(p, k, j, i)] * (tc + tym)) * udy;
    double dzzt = (tzp - 2.0 * tc + tzm) * udzq;
    double dyyt = (typ - 2.0 * tc + tym) * udyq;
    return -(htx + hty + htz) + dyyt + dzzt;
}

// x-second-derivative parts of the implicit RHS (ImplicitAndUpdate*.F90)
__device__ __forceinline__ double rhs_vx(const KP &p, double h, double ru, int k, int j, int i) {
    // ImplicitAndUpdateVX.F90:37-66, k = 1..nxm-1
    const double *vx = p.vx;
    double dxxvx = vx[bidx(p, k + 1, j, i)] * p.ap3ck[k] + vx[bidx(p, k, j, i)] * p.ac3ck[k] + vx[bidx(p, k - 1, j, i)] * p.am3ck[k];
    double dxp = (p.pr[bidx(p, k, j, i)] - p.pr[bidx(p, k - 1, j, i)]) * (p.al * p.udx3c[k]);
    return (p.ga * h + p.ro * ru + (p.al / p.ren) * dxxvx - dxp) * p.dt;
}
__device__ __forceinline__ double rhs_vy(const KP &p, double h, double ru, int k, int j, int jm, int i) {
    // ImplicitAndUpdateVY.F90:41-82
    const double *vy = p.vy;
    const double amm = p.am3sk[k], acc = p.ac3sk[k], app = p.ap3sk[k];
    const int kmm = k == 0 ? 0 : k - 1, kpp = k == p.nxm - 1 ? k : k + 1;
    double dxxvy;
    if (k == 0)
        dxxvy = vy[bidx(p, kpp, j, i)] * app + vy[bidx(p, k, j, i)] * acc - (acc + app) * p.xminusu * p.inslws;
    else if (k == p.nxm - 1)
        dxxvy = -(acc + amm) * p.xplusu * p.inslwn + vy[bidx(p, k, j, i)] * acc + vy[bidx(p, kmm, j, i)] * amm;
    else
        dxxvy = vy[bidx(p, kpp, j, i)] * app + vy[bidx(p, k, j, i)] * acc + vy[bidx(p, kmm, j, i)] * amm;
    double dyp = (p.pr[bidx(p, k, j, i)] - p.pr[bidx(p, k, jm, i)]) * (p.dy * p.al);
    return (p.ga * h + p.ro * ru + (p.al / p.ren) * dxxvy - dyp) * p.dt;
}
__device__ __forceinline__ double rhs_vz(const KP &p, double h, double ru, int k, int j, int i, int im) {
    // ImplicitAndUpdateVZ.F90:38-70
    const double *vz = p.vz;
    const int kmm = k == 0 ? 0 : k - 1, kpp = k == p.nxm - 1 ? k : k + 1;
    double dxxvz = vz[bidx(p, kpp, j, i)] * p.ap3sk[k] + vz[bidx(p, k, j, i)] * p.ac3sk[k] + vz[bidx(p, kmm, j, i)] * p.am3sk[k];
    double dzp = (p.pr[bidx(p, k, j, i)] - p.pr[bidx(p, k, j, im)]) * p.dz * p.al;
    return (p.ga * h + p.ro * ru + (p.al / p.ren) * dxxvz - dzp) * p.dt;
}
__device__ __forceinline__ double rhs_temp(const KP &p, double h, double ru, int k, int j, int i) {
    // ImplicitAndUpdateTemp.F90:33-62
    const double *t = p.temp;
    double dxxt;
    if (k == 0)
        dxxt = t[bidx(p, k + 1, j, i)] * p.ap3ssk[k] + t[bidx(p, k, j, i)] * p.ac3ssk[k]
             - (p.ap3ssk[k] + p.ac3ssk[k]) * p.tempbp[(long long)i * p.ny + j] * p.tfixs;
    else if (k == p.nxm - 1)
        dxxt = t[bidx(p, k, j, i)] * p.ac3ssk[k] + t[bidx(p, k - 1, j, i)] * p.am3ssk[k]
             - (p.am3ssk[k] + p.ac3ssk[k]) * p.temptp[(long long)i * p.ny + j] * p.tfixn;
    else
        dxxt = t[bidx(p, k + 1, j, i)] * p.ap3ssk[k] + t[bidx(p, k, j, i)] * p.ac3ssk[k] + t[bidx(p, k - 1, j, i)] * p.am3ssk[k];
    return (p.ga * h + p.ro * ru + (p.al / p.pect) * dxxt) * p.dt;
}

// ExplicitSalinity, salinity.F90:298-363 (non-IBM), refined k = 0..nxmr-1
__device__ __forceinline__ double ex_sal(const KP &p, int k, int j, int jm, int jp, int i, int im, int ip) {
    const double *s = p.sal, *vxr = p.vxr, *vyr = p.vyr, *vzr = p.vzr;
    const double udyr = 0.5 * p.dyr, udzr = 0.5 * p.dzr, udyrq = p.dyqr / p.pecs, udzrq = p.dzqr / p.pecs;
    const double sc = s[ridx(p, k, j, i)];
    const double syp = s[ridx(p, k, jp, i)], sym = s[ridx(p, k, jm, i)];
    const double szp = s[ridx(p, k, j, ip)], szm = s[ridx(p, k, j, im)];
    const double vp = vxr[ridx(p, k + 1, j, i)], vc = vxr[ridx(p, k, j, i)];
    double hsx;
    if (k == 0)
        hsx = (vp * (s[ridx(p, k + 1, j, i)] + sc) - vc * 2.0 * p.salbp[(long long)i * p.nyr + j]) * p.udx3mr[k] * 0.5;
    else if (k == p.nxmr - 1)
        hsx = (vp * 2.0 * p.saltp[(long long)i * p.nyr + j] - vc * (sc + s[ridx(p, k - 1, j, i)])) * p.udx3mr[k] * 0.5;
    else
        hsx = (vp * (s[ridx(p, k + 1, j, i)] + sc) - vc * (sc + s[ridx(p, k - 1, j, i)])) * p.udx3mr[k] * 0.5;
    double hsy = (vyr[ridx(p, k, jp, i)] * (syp + sc) - vyr[ridx(p, k, j, i)] * (sc + sym)) * udyr;
    double hsz = (vzr[ridx(p, k, j, ip)] * (szp + sc) - vzr[ridx(p, k, j, i)] * (sc + szm)) * udzr;
    double dyys = (syp - 2.0 * sc + sym) * udyrq;
    double dzzs = (szp - 2.0 * sc + szm) * udzrq;
    return -(hsx + hsy + hsz) + dyys + dzzs;
}
// ImplicitSalinity, salinity.F90:378-400
__device__ __forceinline__ double rhs_sal(const KP &p, double h, double ru, int k, int j, int i) {
    const double *s = p.sal;
    double dxxs;
    if (k == 0)
        dxxs = s[ridx(p, k + 1, j, i)] * p.ap3sskr[k] + s[ridx(p, k, j, i)] * p.ac3sskr[k]
             - (p.ap3sskr[k] + p.ac3sskr[k]) * p.salbp[(long long)i * p.nyr + j] * p.sfixs;
    else if (k == p.nxmr - 1)
        dxxs = s[ridx(p, k, j, i)] * p.ac3sskr[k] + s[ridx(p, k - 1, j, i)] * p.am3sskr[k]
             - (p.am3sskr[k] + p.ac3sskr[k]) * p.saltp[(long long)i * p.nyr + j] * p.sfixn;
    else
        dxxs = s[ridx(p, k + 1, j, i)] * p.ap3sskr[k] + s[ridx(p, k, j, i)] * p.ac3sskr[k] + s[ridx(p, k - 1, j, i)] * p.am3sskr[k];
    return (p.ga * h + p.ro * ru + (p.al / p.pecs) * dxxs) * p.dt;
}

}  // namespace afid
