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
    const double *pc, *pc_jm, *pc_im;     // phic of the column and its -y / -z neighbours (phase field)
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
    n.pc = p.phic + c; n.pc_jm = p.phic + cjm; n.pc_im = p.phic + cim;
    n.tbp = p.tempbp[(long long)i * p.ny + j];
    n.ttp = p.temptp[(long long)i * p.ny + j];
    return n;
}

// ExplicitTermsVX.F90:42-127, valid for k = 1..nxm-1 (kc = 2..nxm); BUOY
// false leaves out the buoyancy of :117-127 (the caller adds it)
template <typename T, bool BUOY = true>
__device__ __forceinline__ T ex_vx(const KP &p, const Nb &n, unsigned k) {
    const T udy = T(p.dy) * T(0.25), udz = T(p.dz) * T(0.25), udyq = T(p.dyq) / T(p.ren), udzq = T(p.dzq) / T(p.ren);
    const T rkm = T(p.rudx3m[k - 1]), rkc = T(p.rudx3m[k]), u3 = T(p.udx3c[k]);
    const T vxc = T(n.vx[k]);
    const T vxzp = T(n.vx_ip[k]), vxzm = T(n.vx_im[k]), vxyp = T(n.vx_jp[k]), vxym = T(n.vx_jm[k]);
    T hxz = u3 * udz * ((rkm * T(n.vz_ip[k - 1]) + rkc * T(n.vz_ip[k])) * (vxzp + vxc)
                           - (rkm * T(n.vz[k - 1]) + rkc * T(n.vz[k])) * (vxc + vxzm));
    T hxy = u3 * udy * ((rkm * T(n.vy_jp[k - 1]) + rkc * T(n.vy_jp[k])) * (vxyp + vxc)
                           - (rkm * T(n.vy[k - 1]) + rkc * T(n.vy[k])) * (vxc + vxym));
    const T a = T(n.vx[k + 1]) + vxc, b = vxc + T(n.vx[k - 1]);
    T hxx = (a * a - b * b) * u3 * T(0.25);
    T dzzvx = (vxzm - T(2.0) * vxc + vxzp) * udzq;
    T dyyvx = (vxym - T(2.0) * vxc + vxyp) * udyq;
    T q = -(hxx + hxy + hxz) + dyyvx + dzzvx;
    if (BUOY && p.gaxis == 1) q += T(p.byct) * (T(p.active_t) * T(0.5) * (T(n.t[k]) + T(n.t[k - 1])));
    return q;
}

// ExplicitTermsVY.F90:50-118, k = 0..nxm-1
template <typename T>
__device__ __forceinline__ T ex_vy(const KP &p, const Nb &n, unsigned k) {
    const T udyq = T(p.dyq) / T(p.ren), udzq = T(p.dzq) / T(p.ren), udy = T(p.dy) * T(0.25), udz = T(p.dz) * T(0.25);
    const unsigned kmm = k == 0 ? 0 : k - 1, kpp = k == unsigned(p.nxm - 1) ? k : k + 1;
    const T vyc = T(n.vy[k]);
    const T vyyp = T(n.vy_jp[k]), vyym = T(n.vy_jm[k]), vyzp = T(n.vy_ip[k]), vyzm = T(n.vy_im[k]);
    T hyx = ((T(n.vx[k + 1]) + T(n.vx_jm[k + 1])) * (T(n.vy[kpp]) + vyc)
                - (T(n.vx[k]) + T(n.vx_jm[k])) * (vyc + T(n.vy[kmm]))) * T(p.udx3m[k]) * T(0.25);
    const T a = vyyp + vyc, b = vyym + vyc;
    T hyy = (a * a - b * b) * udy;
    T hyz = ((vyzp + vyc) * (T(n.vz_ip[k]) + T(n.vz_jm_ip[k])) - (vyc + vyzm) * (T(n.vz[k]) + T(n.vz_jm[k]))) * udz;
    T dyyvy = (vyyp - T(2.0) * vyc + vyym) * udyq;
    T dzzvy = (vyzp - T(2.0) * vyc + vyzm) * udzq;
    T q = -(hyx + hyy + hyz) + dyyvy + dzzvy + T(p.Gy);
    if (p.gaxis == 2) q += T(p.byct) * (T(p.active_t) * T(0.5) * (T(n.t[k]) + T(n.t_jm[k])));
    return q;
}

// ExplicitTermsVZ.F90:45-114, k = 0..nxm-1
template <typename T>
__device__ __forceinline__ T ex_vz(const KP &p, const Nb &n, unsigned k) {
    const T udyq = T(p.dyq) / T(p.ren), udzq = T(p.dzq) / T(p.ren), udy = T(p.dy) * T(0.25), udz = T(p.dz) * T(0.25);
    const unsigned kmm = k == 0 ? 0 : k - 1, kpp = k == unsigned(p.nxm - 1) ? k : k + 1;
    const T vzc = T(n.vz[k]);
    const T vzzp = T(n.vz_ip[k]), vzzm = T(n.vz_im[k]), vzyp = T(n.vz_jp[k]), vzym = T(n.vz_jm[k]);
    const T a = vzzp + vzc, b = vzzm + vzc;
    T hzz = (a * a - b * b) * udz;
    T hzy = ((T(n.vy_jp[k]) + T(n.vy_jp_im[k])) * (vzyp + vzc) - (T(n.vy[k]) + T(n.vy_im[k])) * (vzc + vzym)) * udy;
    T hzx = ((T(n.vx[k + 1]) + T(n.vx_im[k + 1])) * (T(n.vz[kpp]) + vzc)
                - (T(n.vx[k]) + T(n.vx_im[k])) * (vzc + T(n.vz[kmm]))) * T(p.udx3m[k]) * T(0.25);
    T dzzvz = (vzzp - T(2.0) * vzc + vzzm) * udzq;
    T dyyvz = (vzyp - T(2.0) * vzc + vzym) * udyq;
    T q = -(hzx + hzy + hzz) + dyyvz + dzzvz;
    if (p.gaxis == 3) q += T(p.byct) * (T(p.active_t) * T(0.5) * (T(n.t[k]) + T(n.t_im[k])));
    return q;
}

// ExplicitTermsTemp.F90:39-77, k = 0..nxm-1
template <typename T>
__device__ __forceinline__ T ex_temp(const KP &p, const Nb &n, unsigned k) {
    const T udz = T(p.dz) * T(0.5), udy = T(p.dy) * T(0.5), udzq = T(p.dzq) / T(p.pect), udyq = T(p.dyq) / T(p.pect);
    const T tc = T(n.t[k]), tzp = T(n.t_ip[k]), tzm = T(n.t_im[k]), typ = T(n.t_jp[k]), tym = T(n.t_jm[k]);
    const T vxp = T(n.vx[k + 1]), vxc = T(n.vx[k]);
    T htx;
    if (k == 0)
        htx = (vxp * (T(n.t[k + 1]) + tc) - vxc * T(2.0) * T(n.tbp)) * T(0.5) * T(p.udx3m[k]);
    else if (k == unsigned(p.nxm - 1))
        htx = (vxp * T(2.0) * T(n.ttp) - vxc * (tc + T(n.t[k - 1]))) * T(0.5) * T(p.udx3m[k]);
    else
        htx = (vxp * (T(n.t[k + 1]) + tc) - vxc * (tc + T(n.t[k - 1]))) * T(0.5) * T(p.udx3m[k]);
    T htz = (T(n.vz_ip[k]) * (tzp + tc) - T(n.vz[k]) * (tc + tzm)) * udz;
    T hty = (T(n.vy_jp[k]) * (typ + tc) - T(n.vy[k]) * (tc + tym)) * udy;
    T dzzt = (tzp - T(2.0) * tc + tzm) * udzq;
    T dyyt = (typ - T(2.0) * tc + tym) * udyq;
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

// latent heat of condensing vapour added to the T explicit term
// (AddCondensation, moisture.F90:304-311) at row k of the base column that
// starts at c0: gamma_q/tau_q (q - qsat) where q > qsat
__device__ __forceinline__ double condensation_heat(const KP &p, long long c0, unsigned k) {
    const double q = p.humid[c0 + k], qs = p.qsat[c0 + k];
    return q > qs ? p.q_gtau * (q - qs) : 0.0;
}

// storage-type selection of a field that exists in FP64 and FP32 forms
template <typename RT> __device__ __forceinline__ RT *sel(double *d, float *f);
template <> __device__ __forceinline__ double *sel<double>(double *d, float *) { return d; }
template <> __device__ __forceinline__ float *sel<float>(double *, float *f) { return f; }
template <typename RT> __device__ __forceinline__ const RT *sel(const double *d, const float *f);
template <> __device__ __forceinline__ const double *sel<double>(const double *d, const float *) { return d; }
template <> __device__ __forceinline__ const float *sel<float>(const double *, const float *f) { return f; }

// refined column (j,i) and neighbours for the salinity stencils
template <typename RT>
struct NbS {
    const RT *s, *s_jm, *s_jp, *s_im, *s_ip;
    const RT *vxr, *vyr, *vyr_jp, *vzr, *vzr_ip;
    double sbp, stp;
};
template <typename RT>
__device__ __forceinline__ NbS<RT> make_nbs(const KP &p, int j, int i) {
    const int jm = wrapm(j, p.nyr), jp = wrapp(j, p.nyr), im = wrapm(i, p.nzr), ip = wrapp(i, p.nzr);
    auto o = [&](int jj, int ii) { return (long long)ii * p.planer + (long long)jj * p.ldxr; };
    const long long c = o(j, i);
    const RT *sal = sel<RT>(p.sal, p.salf), *vxr = sel<RT>(p.vxr, p.vxrf), *vyr = sel<RT>(p.vyr, p.vyrf), *vzr = sel<RT>(p.vzr, p.vzrf);
    NbS<RT> n;
    n.s = sal + c; n.s_jm = sal + o(jm, i); n.s_jp = sal + o(jp, i); n.s_im = sal + o(j, im); n.s_ip = sal + o(j, ip);
    n.vxr = vxr + c; n.vyr = vyr + c; n.vyr_jp = vyr + o(jp, i); n.vzr = vzr + c; n.vzr_ip = vzr + o(j, ip);
    n.sbp = p.salbp[(long long)i * p.nyr + j];
    n.stp = p.saltp[(long long)i * p.nyr + j];
    return n;
}

// ExplicitSalinity, salinity.F90:298-363 (non-IBM), refined k = 0..nxmr-1,
// arithmetic in T, on the values of the row (sm, sc, sp: S at k-1, k, k+1;
// sym/syp/szm/szp: S at the y/z neighbours; vc/vp: vxr at k and k+1; vy/vyp,
// vz/vzp: vyr, vzr of the cell and of its +y/+z neighbour)
template <typename T>
__device__ __forceinline__ T ex_sal_val(const KP &p, unsigned k, T sm, T sc, T sp, T sym, T syp, T szm, T szp,
                                        T vc, T vp, T vy, T vyp, T vz, T vzp, double sbp, double stp) {
    const T udyr = T(0.5) * T(p.dyr), udzr = T(0.5) * T(p.dzr), udyrq = T(p.dyqr) / T(p.pecs), udzrq = T(p.dzqr) / T(p.pecs);
    const T u3 = T(p.udx3mr[k]);
    T hsx;
    if (k == 0)
        hsx = (vp * (sp + sc) - vc * T(2.0) * T(sbp)) * u3 * T(0.5);
    else if (k == unsigned(p.nxmr - 1))
        hsx = (vp * T(2.0) * T(stp) - vc * (sc + sm)) * u3 * T(0.5);
    else
        hsx = (vp * (sp + sc) - vc * (sc + sm)) * u3 * T(0.5);
    T hsy = (vyp * (syp + sc) - vy * (sc + sym)) * udyr;
    T hsz = (vzp * (szp + sc) - vz * (sc + szm)) * udzr;
    T dyys = (syp - T(2.0) * sc + sym) * udyrq;
    T dzzs = (szp - T(2.0) * sc + szm) * udzrq;
    return -(hsx + hsy + hsz) + dyys + dzzs;
}
template <typename T, typename RT>
__device__ __forceinline__ T ex_sal(const KP &p, const NbS<RT> &n, unsigned k) {
    const T sm = k > 0 ? T(n.s[k - 1]) : T(0), sp = k < unsigned(p.nxmr - 1) ? T(n.s[k + 1]) : T(0);
    return ex_sal_val<T>(p, k, sm, T(n.s[k]), sp, T(n.s_jm[k]), T(n.s_jp[k]), T(n.s_im[k]), T(n.s_ip[k]),
                         T(n.vxr[k]), T(n.vxr[k + 1]), T(n.vyr[k]), T(n.vyr_jp[k]), T(n.vzr[k]), T(n.vzr_ip[k]), n.sbp, n.stp);
}
// ImplicitSalinity, salinity.F90:378-400 (FP64 arithmetic in both forms)
__device__ __forceinline__ double rhs_sal_val(const KP &p, unsigned k, double h, double ru, double sm, double sc,
                                              double sp, double sbp, double stp) {
    double dxxs;
    if (k == 0)
        dxxs = sp * p.ap3sskr[k] + sc * p.ac3sskr[k] - (p.ap3sskr[k] + p.ac3sskr[k]) * sbp * p.sfixs;
    else if (k == unsigned(p.nxmr - 1))
        dxxs = sc * p.ac3sskr[k] + sm * p.am3sskr[k] - (p.am3sskr[k] + p.ac3sskr[k]) * stp * p.sfixn;
    else
        dxxs = sp * p.ap3sskr[k] + sc * p.ac3sskr[k] + sm * p.am3sskr[k];
    return (p.ga * h + p.ro * ru + (p.al / p.pecs) * dxxs) * p.dt;
}
template <typename RT>
__device__ __forceinline__ double rhs_sal(const KP &p, const NbS<RT> &n, double h, double ru, unsigned k) {
    const double sm = k > 0 ? double(n.s[k - 1]) : 0.0, sp = k < unsigned(p.nxmr - 1) ? double(n.s[k + 1]) : 0.0;
    return rhs_sal_val(p, k, h, ru, sm, double(n.s[k]), sp, n.sbp, n.stp);
}

}  // namespace afid
