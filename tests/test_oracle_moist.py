"""Pins of the oracle's moisture routines (src/moisture.F90: SetHumidityBCs,
ExplicitHumidity, AddCondensation, UpdateSaturation, ImplicitHumidity with
SolveImpEqnUpdate_Humidity), each by a closed form or an invariant that
shares no code with the oracle, with planted mistakes (made in the oracle,
on the CPU) that must fail:

  * the parameters of ReadMoistParameters (:99-110): pecq = pect / Sm,
    tau_q scaled by pect when ifdiff, gamma_q < 0 replaced by the value that
    makes the moist-static jump 1;
  * saturated walls: humbp / humtp are qsat = exp(alpha (T - beta x)) at the
    wall values x = 0 / 1, and 0 at the top with qfixN = 2;
  * the discrete symbols of the y advection and diffusion of q (i sin kh / h,
    -4 sin^2(kh/2) / h^2 times kappa_q = Sm / pect);
  * condensation moves moisture into buoyancy: hro + gamma_q rkhumid does
    not depend on tau_q, while rkhumid does where q > qsat and does not
    elsewhere;
  * the x-diffusion and its Crank-Nicolson solve: q = sin(m pi x) with zero
    walls is an eigenvector of the discrete operator, so one RK3 step
    multiplies it by prod_s (1 + a_s b lam) / (1 - a_s b lam), lam the
    eigenvalue over pecq.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import random_state, load_state, uniform

GAM = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
ROM = (0.0, -17.0 / 60.0, -5.0 / 12.0)
MOIST = dict(moist=1, q_alpha=3.0, q_beta=1.2, q_gamma=0.19, q_tau=5e-5, q_ifdiff=1, q_fixn=1, q_fixs=1, q_sm=2.5)


def test_parameters():
    cfg = dict(nxm=8, nym=4, nzm=4, rayt=2e6, prat=0.7, **MOIST)
    o = po.Oracle(cfg)
    pect = o.scalars()["pect"]
    m = o.moist_scalars()
    assert abs(pect - np.sqrt(2e6 / 0.7) * 0.7) < 1e-9 * pect
    assert abs(m["pecq"] - pect / 2.5) < 1e-12 * pect
    assert abs(m["tau_q"] - 5e-5 * pect) < 1e-15 * pect
    assert m["gamma_q"] == 0.19
    o.close()
    o = po.Oracle(dict(cfg, q_gamma=-1.0, q_ifdiff=0))
    m = o.moist_scalars()
    assert m["tau_q"] == 5e-5
    assert abs(m["gamma_q"] - 1.2 / (1.0 - np.exp(-3.0))) < 1e-15
    o.close()
    o = po.Oracle(dict(cfg, q_gamma=-1.0, q_fixn=2))      # unsaturated top: the jump is beta
    assert o.moist_scalars()["gamma_q"] == 1.2
    o.close()


@pytest.mark.parametrize("fixn", [1, 2])
def test_saturated_walls(fixn):
    cfg = dict(nxm=8, nym=6, nzm=4, **dict(MOIST, q_fixn=fixn))
    o = po.Oracle(cfg)
    tb = 0.2 + 0.1 * uniform(3, (4, 6, 1))
    tt = -0.3 + 0.1 * uniform(4, (4, 6, 1))
    o.field("tempbp")[...] = tb
    o.field("temptp")[...] = tt
    o.call("set_humidity_bcs")
    a, b = 3.0, 1.2
    assert np.allclose(o.field("humbp"), np.exp(a * (tb - b * 0.0)), rtol=1e-15, atol=0)
    top = np.exp(a * (tt - b * 1.0)) if fixn == 1 else np.zeros_like(tt)
    assert np.allclose(o.field("humtp"), top, rtol=1e-15, atol=0)
    o.close()


def test_humidity_advection_diffusion_symbol():
    """q = cos(k y), vy = U, vz = W, vx = 0, qsat far above q: rkhumid =
    U sin(kh)/h sin(ky) - kappa_q 4 sin^2(kh/2)/h^2 cos(ky), kappa_q = Sm/pect."""
    ny, U = 16, 0.7
    cfg = dict(nxm=8, nym=ny, nzm=4, ylen=1.0, zlen=1.0, rayt=1e4, prat=1.0, **MOIST)
    o = po.Oracle(cfg)
    y = (np.arange(ny) + 0.5) / ny
    kk = 2 * np.pi * 3
    q = np.cos(kk * y)[None, :, None] * np.ones((4, ny, 9))
    q[..., 8] = 0.0
    o.field("humid")[...] = q
    o.field("qsat")[...] = 10.0
    o.field("vy")[...] = U
    o.field("vz")[...] = 0.25
    o.field("humbp")[...] = q[..., :1]
    o.field("humtp")[...] = q[..., :1]
    o.set_substep(1e-3, 1)
    o.call("explicit_humidity")
    h = 1.0 / ny
    kappa = 2.5 / o.scalars()["pect"]
    expect = U * np.sin(kk * h) / h * np.sin(kk * y) - kappa * 4 * np.sin(kk * h / 2) ** 2 / h ** 2 * np.cos(kk * y)
    got = o.field("hhumid")[0, :, 3]
    assert np.allclose(got, expect, atol=1e-11 * np.abs(expect).max())
    o.close()


def condensation_terms(tau, seed=5):
    cfg = dict(nxm=12, nym=8, nzm=6, istr3=4, str3=1.5, rayt=1e5, prat=1.0, **dict(MOIST, q_tau=tau))
    o = po.Oracle(cfg)
    load_state(o, random_state(cfg, seed), upload=False)
    shape = o.field("humid").shape
    o.field("humid")[...] = 1.0 + 0.5 * uniform(seed + 20, shape)
    o.field("qsat")[...] = 1.0 + 0.5 * uniform(seed + 21, shape)       # about half the cells supersaturated
    o.field("humbp")[...] = 1.1
    o.field("humtp")[...] = 0.9
    o.set_substep(1e-3, 2)
    o.call("explicit_terms_temp")
    o.call("explicit_humidity")
    o.call("add_condensation")
    nxm = cfg["nxm"]
    sup = (o.field("humid") > o.field("qsat"))[..., :nxm]
    out = (o.field("hro")[..., :nxm].copy(), o.field("hhumid")[..., :nxm].copy(), o.moist_scalars()["gamma_q"], sup)
    o.close()
    return out


def condensation_invariant_error():
    h1, r1, g, sup = condensation_terms(5e-5)
    h2, r2, _, _ = condensation_terms(2e-4)
    inv = np.abs((h1 + g * r1) - (h2 + g * r2)).max() / np.abs(h1 + g * r1).max()
    return inv, sup, r1, r2


def test_condensation_moves_moisture_into_buoyancy():
    inv, sup, r1, r2 = condensation_invariant_error()
    assert 0.2 < sup.mean() < 0.8
    assert inv < 1e-13, inv
    assert np.abs(r1 - r2)[sup].min() > 0.0            # the sink acts where q > qsat ...
    assert np.array_equal(r1[~sup], r2[~sup])          # ... and nowhere else


def test_condensation_pin_catches_planted_mistake():
    with po.planted(7):                                # heating with the wrong sign
        inv, _, _, _ = condensation_invariant_error()
    assert inv > 1e-3


def humidity_decay_error(m=3, nxm=24, dt=2e-3, sm=2.5):
    ra, pr = 900.0, 1.0
    cfg = dict(nxm=nxm, nym=4, nzm=4, istr3=0, rayt=ra, prat=pr, gaxis=1, active_t=0, **dict(MOIST, q_sm=sm))
    o = po.Oracle(cfg)
    for n in ("vx", "vy", "vz", "pr", "rux", "ruy", "ruz", "rutemp", "ruhumid"):
        o.field(n)[...] = 0.0
    xm = (np.arange(nxm) + 0.5) / nxm
    beta = 1.2
    o.field("tempbp")[...] = 1.0
    o.field("temptp")[...] = 1.0 + beta
    o.field("temp")[..., :nxm] = 1.0 + beta * xm          # conductive, steady: qsat = exp(alpha) everywhere
    o.field("humbp")[...] = 0.0
    o.field("humtp")[...] = 0.0
    q0 = 0.1 * np.sin(m * np.pi * xm)
    o.field("humid")[..., :nxm] = q0
    o.call("update_saturation")
    o.time_marcher(dt)
    pecq = o.moist_scalars()["pecq"]
    lam = -4.0 * nxm * nxm * np.sin(m * np.pi / (2 * nxm)) ** 2 / pecq
    g = 1.0
    for ga, ro in zip(GAM, ROM):
        a = ga + ro
        g *= (1 + a * 0.5 * dt * lam) / (1 - a * 0.5 * dt * lam)
    err = np.abs(o.field("humid")[..., :nxm] - g * q0).max()
    o.close()
    return err / (0.1 * (1.0 - g))        # relative to the decay itself


def test_humidity_diffusion_eigenmode():
    assert humidity_decay_error() < 1e-9


def test_humidity_pin_catches_planted_mistake():
    with po.planted(8):                               # x-diffusion with pect instead of pecq
        assert humidity_decay_error() > 1e-3


def test_moist_stats_closed_forms():
    """CalcMoistStats (moisture.F90:422-505) on fields with closed-form
    averages: q = f(x) + A sin(2 pi y / Ly), qsat = c(x) uniform in y and z,
    vx = U on the interior faces (0 on the walls), vy = V, vz = W uniform:
    qbar = f, qsbar = c, qrms = sqrt(f^2 + A^2 / 2), qrel = f / c, vxq =
    U f (U / 2 f in the two wall cells), vyq = V f, vzq = W f."""
    nxm, ny, nz = 10, 8, 6
    o = po.Oracle(dict(nxm=nxm, nym=ny, nzm=nz, istr3=4, str3=1.5, ylen=1.5, zlen=0.75, **MOIST))
    xm = o.grid("xm")[:nxm]
    f, c = 1.0 + 0.5 * xm, 2.0 - xm
    A, U, V, W = 0.3, 0.2, -0.7, 0.45
    y = np.arange(ny)
    q = np.zeros_like(o.field("humid"))
    q[..., :nxm] = f[None, None, :] + A * np.sin(2 * np.pi * y / ny)[None, :, None]
    o.field("humid")[...] = q
    o.field("qsat")[...] = 1.0
    o.field("qsat")[..., :nxm] = c
    o.field("vx")[...] = 0.0
    o.field("vx")[..., 1:nxm] = U
    o.field("vy")[...] = V
    o.field("vz")[...] = W
    m = o.moist_stats()
    wall = np.ones(nxm); wall[0] = wall[-1] = 0.5
    expect = dict(qbar=f, qsbar=c, qrms=np.sqrt(f * f + A * A / 2), qrel=f / c, vxq=U * wall * f, vyq=V * f, vzq=W * f)
    for k, v in expect.items():
        assert np.allclose(m[k], v, rtol=1e-13, atol=1e-14), k
    o.close()
