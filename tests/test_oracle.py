"""Pins of the FP64 CPU oracle (CPU only).

Each test checks the oracle against something it does not share code with:
numpy's FFT and dense solver, closed forms of the grid and of discrete
Fourier symbols, exactness and conservation properties of the discretisation,
and the reference's one published known answer (1-D diffusion,
docs/convergence.md:3-36).
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import random_state, load_state, relerr, uniform

po.set_threads(2)


# ------------------------------------------------------------------ numerics
@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8, 9, 12, 16, 24, 64, 96, 128, 384, 768, 7, 10])
def test_fft_matches_numpy(n):
    x = uniform(n, (n,)) + 1j * uniform(n + 100, (n,))
    assert relerr(po.fft(x, -1), np.fft.fft(x)) < 1e-14
    assert relerr(po.fft(x, +1), np.fft.ifft(x) * n) < 1e-14


@pytest.mark.parametrize("n", [5, 12, 30])
def test_naive_dft_definition(n):
    x = uniform(3 * n, (n,)) + 1j * uniform(5 * n, (n,))
    assert relerr(po.dft_naive(x, -1), np.fft.fft(x)) < 1e-14


def test_thomas_matches_dense_solve():
    n = 40
    a, c = uniform(1, (n,)), uniform(2, (n,))
    b = 2.5 + uniform(3, (n,))
    a[0] = 0.0; c[-1] = 0.0
    d = uniform(4, (n,))
    M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    assert relerr(po.tridiag_solve(a, b, c, d), np.linalg.solve(M, d)) < 1e-14


# ---------------------------------------------------------------------- grid
def test_cosine_grid_closed_form():
    n = 32
    o = po.Oracle(dict(nxm=n, nym=4, nzm=4, istr3=8))
    xc = o.grid("xc")
    k = np.arange(n + 1)
    assert np.allclose(xc, 0.5 * (1 - np.cos(np.pi * k / n)), atol=1e-15, rtol=0)
    xm = o.grid("xm")[:n]
    assert np.allclose(xm, 0.5 * (xc[:-1] + xc[1:]), atol=1e-15)


def test_clipped_chebyshev_grid_closed_form():
    """grid.F90:55-94 with nclip = 0: x_k = (1 - eta_k/eta_1)/2, eta = cos(pi (k-1/2)/(n+1))."""
    n = 20
    o = po.Oracle(dict(nxm=n, nym=4, nzm=4, istr3=6, str3=0.0))
    xc = o.grid("xc")
    i = np.arange(1, n + 2)
    eta = np.cos(np.pi * (i - 0.5) / (n + 1))
    ref = 0.5 * (1 - 2 * eta / (eta[0] - eta[-1]))
    ref[0], ref[-1] = 0.0, 1.0
    assert np.allclose(xc, ref, atol=1e-15)


def test_metrics_invariants():
    n = 24
    o = po.Oracle(dict(nxm=n, nym=4, nzm=4, istr3=4, str3=1.5, alx3=2.0))
    s = o.scalars()
    d3xc = o.grid("d3xc")
    assert abs(d3xc.sum() / s["dx"] - 2.0) < 1e-14          # sum of cell widths = alx3
    assert np.allclose(o.grid("udx3m")[:n], s["dx"] / d3xc, rtol=1e-15)


@pytest.mark.parametrize("which", ["ck", "sk", "ssk"])
def test_second_derivative_exact_on_quadratics(which):
    """Interior rows of the three-point second derivative are exact for x^2."""
    n = 16
    o = po.Oracle(dict(nxm=n, nym=4, nzm=4, istr3=4, str3=2.0))
    x = o.grid("xc") if which == "ck" else o.grid("xm")[:n]
    ap, ac, am = (o.grid(p + "3" + which) for p in ("ap", "ac", "am"))
    m = len(x)
    for f, d2 in ((x ** 2, 2.0), (x, 0.0), (np.ones_like(x), 0.0)):
        v = ap[1:m - 1] * f[2:] + ac[1:m - 1] * f[1:m - 1] + am[1:m - 1] * f[:m - 2]
        assert np.allclose(v, d2, atol=1e-9 * np.abs(ap).max())


def test_dirichlet_wall_row_folds_ghost():
    """Wall row of the centred operator: ghost value 2 T_w - T_1 at the mirror
    point (grid.F90:275-281) -> a linear profile through the wall value has
    zero second derivative once the wall term is added back."""
    n = 12
    o = po.Oracle(dict(nxm=n, nym=4, nzm=4, istr3=4, str3=2.0))
    x = o.grid("xm")[:n]
    ap, ac = o.grid("ap3ssk"), o.grid("ac3ssk")
    T = 0.3 + 0.7 * x          # linear, wall value 0.3 at x = 0
    d2 = ap[0] * T[1] + ac[0] * T[0] - (ap[0] + ac[0]) * 0.3
    assert abs(d2) < 1e-9 * abs(ap[0])


def test_modified_wavenumbers():
    o = po.Oracle(dict(nxm=8, nym=12, nzm=10, ylen=1.5, zlen=2.0))
    ak1, ak2 = o.grid("ak1"), o.grid("ak2")
    k = np.arange(10)
    assert np.allclose(ak1, 4 * np.sin(np.pi * k / 10) ** 2 * (10 / 2.0) ** 2, rtol=1e-13, atol=1e-12)
    k = np.arange(12)
    assert np.allclose(ak2, 4 * np.sin(np.pi * k / 12) ** 2 * (12 / 1.5) ** 2, rtol=1e-13, atol=1e-12)


def test_rk3_coefficients():
    # Rai & Moin low-storage RK3: gamma + rho sums to one over the three substeps
    g = np.array([8 / 15, 5 / 12, 3 / 4]); r = np.array([0, -17 / 60, -5 / 12])
    assert abs((g + r).sum() - 1.0) < 1e-15


# ------------------------------------------------------------ explicit terms
def test_temperature_advection_discrete_symbol():
    """T = cos(k y), vy = U, vz = W, vx = 0: the central flux difference and
    the second difference have the symbols i sin(k h)/h and -4 sin^2(k h/2)/h^2."""
    ny, U = 16, 0.7
    cfg = dict(nxm=8, nym=ny, nzm=4, ylen=1.0, zlen=1.0, rayt=1e4, prat=1.0)
    o = po.Oracle(cfg)
    y = (np.arange(ny) + 0.5) / ny
    kk = 2 * np.pi * 3
    T = np.cos(kk * y)[None, :, None] * np.ones((4, ny, 9))
    T[..., 8] = 0.0
    o.field("temp")[...] = T
    o.field("vy")[...] = U
    o.field("vz")[...] = 0.25
    o.field("tempbp")[...] = T[..., :1]
    o.field("temptp")[...] = T[..., :1]
    o.set_substep(1e-3, 1)
    o.call("explicit_terms_temp")
    h = 1.0 / ny
    pect = o.scalars()["pect"]
    expect = U * np.sin(kk * h) / h * np.sin(kk * y) - 4 * np.sin(kk * h / 2) ** 2 / h ** 2 / pect * np.cos(kk * y)
    got = o.field("hro")[0, :, 3]
    assert np.allclose(got, expect, atol=1e-11 * np.abs(expect).max())


def _divergence_free_state(cfg, seed):
    o = po.Oracle(cfg)
    st = random_state(cfg, seed)
    load_state(o, st, upload=False)
    o.set_substep(1e-3, 1)
    for r in ("calc_local_divergence", "solve_pressure_correction", "correct_velocity"):
        o.call(r)
    return o


def test_projection_makes_velocity_divergence_free():
    cfg = dict(nxm=16, nym=12, nzm=8, istr3=4, str3=1.2)
    o = po.Oracle(cfg)
    load_state(o, random_state(cfg, 2), upload=False)
    d0, _ = o.check_divergence()
    o.set_substep(1e-3, 2)
    for r in ("calc_local_divergence", "solve_pressure_correction", "correct_velocity"):
        o.call(r)
    d1, _ = o.check_divergence()
    assert d0 > 1.0 and d1 < 1e-12 * d0


def test_poisson_against_dense_operator():
    """The FFT + tridiagonal solve inverts the discrete operator
    L = [amphk, acphk, apphk]_x + dyq D2_y + dzq D2_z (pressure.F90:103-114,
    :209-220) for a consistent right-hand side; checked with a dense matrix."""
    nxm, nym, nzm = 6, 4, 6
    cfg = dict(nxm=nxm, nym=nym, nzm=nzm, istr3=4, str3=1.5, ylen=1.3, zlen=0.7)
    o = po.Oracle(cfg)
    s = o.scalars()
    am, ac, ap = o.grid("amphk"), o.grid("acphk"), o.grid("apphk")
    N = nxm * nym * nzm
    Lm = np.zeros((N, N))
    idx = lambda k, j, i: (i % nzm) * nym * nxm + (j % nym) * nxm + k
    for i in range(nzm):
        for j in range(nym):
            for k in range(nxm):
                r = idx(k, j, i)
                Lm[r, r] += ac[k] - 2 * s["dy"] ** 2 - 2 * s["dz"] ** 2
                if k > 0: Lm[r, idx(k - 1, j, i)] += am[k]
                if k < nxm - 1: Lm[r, idx(k + 1, j, i)] += ap[k]
                Lm[r, idx(k, j + 1, i)] += s["dy"] ** 2
                Lm[r, idx(k, j - 1, i)] += s["dy"] ** 2
                Lm[r, idx(k, j, i + 1)] += s["dz"] ** 2
                Lm[r, idx(k, j, i - 1)] += s["dz"] ** 2
    # consistent rhs: the operator applied to a random field
    phi = uniform(8, (nzm, nym, nxm))
    rhs = (Lm @ phi.reshape(-1)).reshape(nzm, nym, nxm)
    d = o.field("dph")
    d[...] = 0.0
    d[..., :nxm] = rhs
    o.set_substep(1e-3, 1)
    o.call("solve_pressure_correction")
    sol = o.field("dph")[..., :nxm]
    diff = sol - phi
    diff -= diff.mean()                 # pressure is defined up to a constant
    assert np.abs(diff).max() < 1e-10 * np.abs(phi).max()


def test_momentum_advection_conserves_energy():
    """With diffusion switched off (huge Ra) and a divergence-free velocity,
    the conservative advection terms exchange no kinetic energy:
    sum of u.H weighted by the cell volumes of each staggered component is
    zero to round-off."""
    cfg = dict(nxm=12, nym=8, nzm=10, istr3=4, str3=1.5, rayt=1e40, prat=1.0, active_t=0)
    o = _divergence_free_state(cfg, 5)
    for r in ("explicit_terms_vx", "explicit_terms_vy", "explicit_terms_vz"):
        o.call(r)
    s = o.scalars()
    d3xc = o.grid("d3xc")            # centred cell widths (times dx)
    g3rc = o.grid("g3rc")[:12]       # face-centred cell widths (times dx)
    vx, vy, vz = o.field("vx"), o.field("vy"), o.field("vz")
    qc, dp, dq = o.field("qcap"), o.field("dph"), o.field("dq")
    ex = (vx[..., 1:12] * qc[..., 1:12] * g3rc[1:12]).sum()
    ey = (vy[..., :12] * dp[..., :12] * d3xc).sum()
    ez = (vz[..., :12] * dq[..., :12] * d3xc).sum()
    scale = (np.abs(vx[..., 1:12] * qc[..., 1:12] * g3rc[1:12]).sum() + np.abs(vy[..., :12] * dp[..., :12] * d3xc).sum())
    assert abs(ex + ey + ez) < 1e-12 * scale


def test_temperature_advection_is_conservative():
    cfg = dict(nxm=12, nym=8, nzm=10, istr3=4, str3=1.5, rayt=1e40, prat=1.0)
    o = _divergence_free_state(cfg, 6)
    o.call("explicit_terms_temp")
    d3xc = o.grid("d3xc")
    hro = o.field("hro")[..., :12]
    assert abs((hro * d3xc).sum()) < 1e-12 * np.abs(hro * d3xc).sum()


# --------------------------------------------------------------- implicit
def test_conduction_and_couette_are_steady():
    """Linear T between its wall values and linear vy between the wall
    velocities are steady solutions: one full step leaves them unchanged."""
    nxm = 16
    cfg = dict(nxm=nxm, nym=4, nzm=4, istr3=4, str3=1.5, xplusu=0.5, xminusu=-0.25, active_t=0)
    o = po.Oracle(cfg)
    xm = o.grid("xm")[:nxm]
    T = np.zeros((4, 4, nxm + 1)); T[..., :nxm] = 1.0 - xm
    V = np.zeros((4, 4, nxm + 1)); V[..., :nxm] = -0.25 + 0.75 * xm
    o.field("temp")[...] = T
    o.field("vy")[...] = V
    o.field("tempbp")[...] = 1.0
    o.field("temptp")[...] = 0.0
    for _ in range(2):
        o.time_marcher(1e-2)
    assert np.abs(o.field("temp")[..., :nxm] - T[..., :nxm]).max() < 1e-13
    assert np.abs(o.field("vy")[..., :nxm] - V[..., :nxm]).max() < 1e-13


def test_one_d_diffusion_known_answer():
    """docs/convergence.md:3-36: kappa = 0.01 (Ra = 1e4, Pr = 1), dt = 1e-5,
    t in [0,1], 64 uniform cells, T = 0.5 - x + sin(2 pi x) - sin(6 pi x):
    published L_inf = 2.71e-3 and L2 = 1.28e-3 (digitised from
    docs/figures/spatial_convergence.svg).  L2 is time-averaged over output
    samples every 0.1 including t = 0."""
    n, dt = 64, 1e-5
    o = po.Oracle(dict(nxm=n, nym=1, nzm=1, istr3=0, rayt=1e4, prat=1.0, gaxis=3, active_t=0,
                       ylen=0.01, zlen=0.01))
    po.set_threads(1)
    xm = o.grid("xm")[:n]
    o.field("tempbp")[...] = 0.5
    o.field("temptp")[...] = -0.5
    T = o.field("temp")
    T[0, 0, :n] = 0.5 - xm + np.sin(2 * np.pi * xm) - np.sin(6 * np.pi * xm)
    kap = 0.01
    linf, l2 = 0.0, [0.0]
    for step in range(1, 100001):
        o.time_marcher(dt)
        t = step * dt
        if step % 100 == 0:
            Ta = (0.5 - xm + np.exp(-kap * (2 * np.pi) ** 2 * t) * np.sin(2 * np.pi * xm)
                  - np.exp(-kap * (6 * np.pi) ** 2 * t) * np.sin(6 * np.pi * xm))
            e = T[0, 0, :n] - Ta
            linf = max(linf, np.abs(e).max())
            if step % 10000 == 0:
                l2.append(np.mean(e ** 2))
    po.set_threads(2)
    L2 = np.sqrt(np.mean(l2))
    assert abs(linf - 2.71e-3) < 0.01 * 2.71e-3
    assert abs(L2 - 1.28e-3) < 0.05 * 1.28e-3


# ----------------------------------------------------------- multires
MR = dict(nxm=12, nym=8, nzm=6, istr3=4, str3=1.5, multires=1, salinity=1, nxmr=24, nymr=16, nzmr=12,
          istr3r=4, rayt=1e6, prat=1.0, rays=-1e6, pras=4.0)


@pytest.mark.parametrize("axis,coarse,fine,L", [("x", "xm", "xmr", 1.0)])
def test_hermite_stencil_exact_for_quadratics_x(axis, coarse, fine, L):
    """Weights of construct_stencil (docs/interpolation.md: cubic Hermite with
    second-order derivative estimates) reproduce quadratics exactly away from
    the walls and linears everywhere."""
    o = po.Oracle(MR)
    xm = o.grid("xm")[:MR["nxm"]]
    xmr = o.grid("xmr")[:MR["nxmr"]]
    w = o.grid("cxrs").reshape(-1, 4)
    irangs = o.int_grid(0)
    # coarse values with mirror ghosts at -xm(1) and 2L - xm(n)
    xo = np.concatenate([[-xm[1], -xm[0]], xm, [2 * L - xm[-1], 2 * L - xm[-2]]])   # index kc = -1..nxm+2
    for kr in range(MR["nxmr"]):
        kc = max(i for i in range(MR["nxm"] + 1) if irangs[i] <= kr + 1)   # owning interval
        pts = xo[kc - 1 + 1: kc + 3 + 1]
        for f, wall_ok in ((lambda x: 1 + 0 * x, True), (lambda x: x, False), (lambda x: x * x, False)):
            val = (w[kr] * f(pts)).sum()
            if 0 < kc < MR["nxm"] or wall_ok:
                assert abs(val - f(xmr[kr])) < 1e-12, (kr, kc)
        assert abs(w[kr].sum() - 1.0) < 1e-14


def test_refined_velocity_divergence_free():
    """InterpVelMgrd integrates Hermite-transferred gradients, so a discretely
    divergence-free coarse velocity gives a discretely divergence-free
    refined velocity (docs/interpolation.md:41-63), including the periodic seam."""
    o = _divergence_free_state(MR, 7)
    d, _ = o.check_divergence()
    o.call("interp_vel_mgrd")
    _, dr = o.check_divergence()
    g = np.abs(o.field("vxr")).max() * 2 * MR["nxmr"]
    assert d < 1e-12 * g and dr < 1e-11 * g


def test_salc_linear_profile_exact():
    """A linear S profile through its wall values transfers to the coarse
    faces exactly (gAxis = 1: salc lives at xc, CreateMgrdStencil.F90:25-28)."""
    o = po.Oracle(MR)
    xmr = o.grid("xmr")[:MR["nxmr"]]
    S = np.zeros((MR["nzmr"], MR["nymr"], MR["nxmr"] + 1))
    S[..., :MR["nxmr"]] = 0.5 - xmr
    o.field("sal")[...] = S
    o.field("salbp")[...] = 0.5
    o.field("saltp")[...] = -0.5
    o.call("interp_sal_multigrid")
    xc = o.grid("xc")[:MR["nxm"]]
    assert np.abs(o.field("salc")[..., :MR["nxm"]] - (0.5 - xc)).max() < 1e-13


def test_cfl_uniform_flow():
    o = po.Oracle(dict(nxm=8, nym=6, nzm=5, ylen=1.0, zlen=2.0))
    o.field("vz")[...] = 0.3
    o.field("vy")[...] = -0.2
    cfl, _ = o.calc_max_cfl()
    assert abs(cfl - (0.3 * 5 / 2.0 + 0.2 * 6)) < 1e-14


def test_nusselt_conduction_is_one():
    nxm = 10
    o = po.Oracle(dict(nxm=nxm, nym=4, nzm=4, istr3=4, str3=1.0))
    xm = o.grid("xm")[:nxm]
    o.field("temp")[..., :nxm] = 1 - xm
    q = o.global_quantities()
    assert abs(q["nu"] - 1.0) < 1e-15 and q["ke"] == 0.0
