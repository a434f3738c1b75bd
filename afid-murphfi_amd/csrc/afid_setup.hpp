// afid_setup.hpp -- host-side set-up of the wall-normal grid, metric
// coefficients, Poisson coefficients and Hermite transfer stencils.
//
// These are the reference's set-up routines (all FP64, run once at init):
//   CreateGrid           src/flow_solver/CreateGrid.F90:11-179
//   GridModule           src/grid.F90:12-307
//   InitPressureSolver   src/pressure.F90:61-114
//   CreateMgrdGrid       src/multires/CreateMgrdGrid.F90:11-166
//   interpolation_indices / extend_grid / construct_stencil
//                        src/multires/HermiteInterpolations.F90:28-178
//   CreateMgrdStencil    src/multires/CreateMgrdStencil.F90:1-81
// The formulas have to match the reference term by term, so the set-up
// follows the same steps; the data it produces is then flattened into the
// per-point tables the kernels read (one source index + four weights per
// target point and axis).
#pragma once
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace afid {

// A 1-D array addressed with Fortran bounds lo..hi.
struct FArr {
    int lo = 1;
    std::vector<double> v;
    FArr() = default;
    FArr(int lo_, int hi_) : lo(lo_), v(size_t(hi_ - lo_ + 1), 0.0) {}
    double &operator()(int i) { return v[size_t(i - lo)]; }
    double operator()(int i) const { return v[size_t(i - lo)]; }
};
struct IArr {
    int lo = 0;
    std::vector<int> v;
    IArr() = default;
    IArr(int lo_, int hi_) : lo(lo_), v(size_t(hi_ - lo_ + 1), 0) {}
    int &operator()(int i) { return v[size_t(i - lo)]; }
    int operator()(int i) const { return v[size_t(i - lo)]; }
};

static const double kPi = 2.0 * std::asin(1.0);

// Wall-normal face coordinates c(1..Nm+1) and centres m(1..Nm) for the
// clustering option istr3 (CreateGrid.F90:38-74, grid.F90).  Option 8 is
// the full-cosine clustering x_k = L/2 (1 - cos(pi k/Nm)), k = 0..Nm, which
// the reference does not offer.
inline void make_wall_grid(int istr3, int Nm, double L, double str, double sc_or_retau,
                           FArr &c, FArr &m) {
    c = FArr(1, Nm + 1);
    m = FArr(0, Nm + 1);
    switch (istr3) {
    case 0:  // uniform_grid, grid.F90:12-28
        for (int i = 1; i <= Nm + 1; ++i) c(i) = L * double(i - 1) / double(Nm);
        break;
    case 1:  // centre_focus_grid, grid.F90:133-152
        c(1) = 0.0;
        for (int i = 2; i <= Nm; ++i) {
            double q = 2.0 * (i - 1) / Nm - 1;
            c(i) = L / 2.0 * (str * std::pow(q, 7) + q) / (str + 1) + L / 2.0;
        }
        c(Nm + 1) = L;
        break;
    case 2: {  // natural_BL_grid, grid.F90:154-180
        const double jb = 16.0, ce = 0.8, al = 1.5, dyw = 0.1;
        for (int j = 1; j <= Nm + 1; ++j) {
            double jm = j - 1;
            c(j) = 1 / (1.0 + std::pow(jm / jb, 2)) * (dyw * jm + std::pow(0.75 * al * ce * jm, 4.0 / 3.0) * std::pow(jm / jb, 2));
        }
        double top = c(Nm + 1);
        for (int j = 1; j <= Nm + 1; ++j) c(j) = c(j) / top * L;
        break;
    }
    case 3: {  // sym_natural_BL_grid, grid.F90:182-215
        const double jb = 16.0, ce = 0.8, al = 1.5, dyw = 0.1;
        int nh = Nm / 2;
        for (int j = 1; j <= nh + 1; ++j) {
            double jm = j - 1;
            c(j) = 1 / (1.0 + std::pow(jm / jb, 2)) *
                   (dyw * jm + std::pow(0.75 * al * ce * jm / std::sqrt(sc_or_retau), 4.0 / 3.0) * std::pow(jm / jb, 2));
        }
        double Lh = 2.0 * c(nh + 1);
        for (int j = 1; j <= nh; ++j) c(Nm + 2 - j) = Lh - c(j);
        double top = c(Nm + 1);
        for (int j = 1; j <= Nm + 1; ++j) c(j) = c(j) / top * L;
        break;
    }
    case 4: {  // tanh_grid, grid.F90:30-53
        double tstr = std::tanh(str);
        c(1) = 0.0;
        for (int i = 2; i <= Nm + 1; ++i) {
            double z2 = double(2 * (i - 1) - Nm) / double(Nm);
            c(i) = 0.5 * (1 + std::tanh(str * z2) / tstr) * L;
        }
        break;
    }
    case 5: {  // scallop_grid, grid.F90:217-262
        double Retau = sc_or_retau, dw = 0.5;
        double kb = 0.2 * Retau / dw;
        double alpha = 8.0 / (Nm + 1 - kb) * std::sqrt(Retau) * (1.0 - std::pow(5.0, -0.5));
        c(1) = 0.0;
        for (int k = 1; k <= Nm; ++k) {
            double dxup = alpha / 4.0 * (alpha / 8.0 * (k - kb) + std::sqrt(Retau / 5.0));
            double sig = 0.5 * (1.0 + std::tanh((k - kb) / 20));
            c(k + 1) = c(k) + (1.0 - sig) * dw + sig * dxup;
        }
        double top = c(Nm + 1);
        for (int k = 1; k <= Nm + 1; ++k) c(k) = c(k) / top * L;
        break;
    }
    case 6: case 7: {  // cheb_grid grid.F90:55-94, asym_cheb_grid grid.F90:96-131
        int nclip = int(str);
        int Next = istr3 == 6 ? Nm + 1 + 2 * nclip : Nm + 1 + nclip;
        std::vector<double> ez(size_t(Next + 1)), e(size_t(Nm + 2));
        for (int i = 1; i <= Next; ++i)
            ez[i] = istr3 == 6 ? std::cos(kPi * (double(i) - 0.5) / double(Next)) : std::cos(kPi * double(i) / double(Next) / 2.0);
        for (int i = 1; i <= Nm + 1; ++i) e[i] = ez[size_t(i + nclip)];
        double delet = istr3 == 6 ? e[1] - e[size_t(Nm + 1)] : e[1];
        for (int i = 1; i <= Nm + 1; ++i) e[i] = istr3 == 6 ? 2.0 * e[i] / delet : e[i] / delet;
        c(1) = 0.0;
        for (int i = 2; i <= Nm; ++i) c(i) = istr3 == 6 ? 0.5 * L * (1.0 - e[i]) : L * (1.0 - e[i]);
        c(Nm + 1) = L;
        break;
    }
    case 8:
        c(1) = 0.0;
        for (int i = 2; i <= Nm; ++i) c(i) = 0.5 * L * (1.0 - std::cos(kPi * double(i - 1) / double(Nm)));
        c(Nm + 1) = L;
        break;
    default:
        throw std::invalid_argument("unsupported wall-normal grid option " + std::to_string(istr3));
    }
    for (int i = 1; i <= Nm; ++i) m(i) = 0.5 * (c(i) + c(i + 1));
}

// second_derivative_coeff, grid.F90:264-307, on x(1..n).
inline void second_derivative_coeff(FArr &ap, FArr &ac, FArr &am, const FArr &x, int n, double xlen,
                                    int fix_up, int fix_low) {
    ap = FArr(1, n); ac = FArr(1, n); am = FArr(1, n);
    double a33 = 2.0 / (x(1) + x(2));
    ap(1) = a33 / (x(2) - x(1));
    am(1) = 0.0;
    ac(1) = -ap(1) - fix_low * a33 / x(1);
    if (x(1) == 0.0) { ap(1) = 0.0; ac(1) = 1.0; }
    for (int k = 2; k <= n - 1; ++k) {
        a33 = 2.0 / (x(k + 1) - x(k - 1));
        ap(k) = a33 / (x(k + 1) - x(k));
        am(k) = a33 / (x(k) - x(k - 1));
        ac(k) = -ap(k) - am(k);
    }
    a33 = 2.0 / (2.0 * xlen - x(n) - x(n - 1));
    ap(n) = 0.0;
    am(n) = a33 / (x(n) - x(n - 1));
    ac(n) = -am(n) - fix_up * a33 / (xlen - x(n));
    if (x(n) == xlen) { am(n) = 0.0; ac(n) = 1.0; }
}

// interpolation_indices, HermiteInterpolations.F90:28-58: idx(0..n_old+1),
// idx(io) = first new index whose coordinate is >= x_old(io).
inline IArr interpolation_indices(const FArr &xo, int n_old, const FArr &xnew, int n_new, double L) {
    std::vector<double> xn(size_t(n_new + 2));
    for (int i = 1; i <= n_new; ++i) xn[i] = xnew(i);
    xn[0] = -xn[1];
    if (xn[0] == xn[1]) xn[0] = -xn[2];
    xn[size_t(n_new + 1)] = 2.0 * L - xn[size_t(n_new)];
    if (xn[size_t(n_new + 1)] == xn[size_t(n_new)]) xn[size_t(n_new + 1)] = 2.0 * L - xn[size_t(n_new - 1)];
    IArr idx(0, n_old + 1);
    idx(0) = 1;
    for (int io = 1; io <= n_old; ++io)
        for (int in = 0; in <= n_new; ++in)
            if (xn[size_t(in)] < xo(io) && xn[size_t(in + 1)] >= xo(io)) idx(io) = in + 1;
    idx(n_old + 1) = n_new + 1;
    return idx;
}

// extend_grid, HermiteInterpolations.F90:62-103.
inline FArr extend_grid(const FArr &x, int n, double L, int next) {
    FArr e(1 - next, n + next);
    for (int i = 1; i <= n; ++i) e(i) = x(i);
    for (int i = 1; i <= next; ++i) e(1 - i) = std::fabs(x(1)) < 1e-10 ? -x(1 + i) : -x(i);
    for (int i = 1; i <= next; ++i) e(n + i) = std::fabs(x(n) - L) < 1e-10 ? 2 * L - x(n - i) : 2 * L - x(n + 1 - i);
    return e;
}

// One axis of a Hermite transfer, flattened per target point: the target
// point in (1-based) reads source points src(in)..src(in)+3 with weights
// w[4*(in-1) .. 4*(in-1)+3].
struct Stencil {
    int n_new = 0;
    std::vector<int> src;     // size n_new, 1-based source index of the first point
    std::vector<double> w;    // 4*n_new
};

// construct_stencil, HermiteInterpolations.F90:107-178, plus the interval
// each target point belongs to (io), giving src = io - 1.
inline Stencil construct_stencil(const FArr &xold, int n_old, const FArr &xnew, int n_new, double L,
                                 const IArr &idx, char axis) {
    Stencil st;
    st.n_new = n_new;
    st.src.assign(size_t(n_new), INT32_MIN);
    st.w.assign(size_t(4 * n_new), 0.0);
    int next = axis == 'x' ? 1 : 2;
    FArr xo = extend_grid(xold, n_old, L, next), xn = extend_grid(xnew, n_new, L, next);
    for (int io = 0; io <= n_old; ++io) {
        int lo = std::max(idx(io), 1), hi = std::min(idx(io + 1) - 1, n_new);
        for (int in = lo; in <= hi; ++in) {
            double *w = &st.w[size_t(4 * (in - 1))];
            st.src[size_t(in - 1)] = io - 1;
            if (axis == 'x' && (io == 0 || io == n_old)) {
                double dlc = xo(io + 1) - xo(io);
                double t = (xn(in) - xo(io)) / dlc;
                w[0] = 0.0; w[1] = 1 - t; w[2] = t; w[3] = 0.0;
            } else {
                double dlm = xo(io) - xo(io - 1), dlc = xo(io + 1) - xo(io), dlp = xo(io + 2) - xo(io + 1);
                double t = (xn(in) - xo(io)) / dlc;
                if (t < 1e-12) {
                    w[0] = 0.0; w[1] = 1.0; w[2] = 0.0; w[3] = 0.0;
                } else {
                    double h00 = (1.0 + 2.0 * t) * (1.0 - t) * (1.0 - t);
                    double h10 = t * (1.0 - t) * (1.0 - t);
                    double h01 = (1.0 + 2.0 * (1.0 - t)) * t * t;
                    double h11 = -(1.0 - t) * t * t;
                    w[0] = -h10 * dlc * dlc / dlm / (dlc + dlm);
                    w[1] = h00 - h11 * dlp / (dlp + dlc) + h10 * (dlc - dlm) / dlm;
                    w[2] = h01 + h10 * dlm / (dlm + dlc) + h11 * (dlp - dlc) / dlp;
                    w[3] = h11 * dlc * dlc / dlp / (dlp + dlc);
                }
            }
        }
    }
    for (int in = 1; in <= n_new; ++in)
        if (st.src[size_t(in - 1)] == INT32_MIN) throw std::runtime_error("Hermite stencil leaves a target point uncovered");
    return st;
}

}  // namespace afid
