"""Pins of the oracle's phase-field restatement (src/phasefield.F90) by
closed forms and invariants, independent of the restated formulas:

* ExplicitPhase: the Allen-Cahn source of a uniform field in closed form;
  pure lateral diffusion of a 0/1 pattern (source term vanishes) against a
  numpy stencil.
* ImplicitPhase: zero-gradient Laplacian conserves the grid-weighted sum;
  a Crank-Nicolson step multiplies an eigenvector of the x-operator by
  (1 + b lambda)/(1 - b lambda), b = al dt pf_D / 2.
* AddVolumePenalty: closed form of the Brinkman term.
* AddLatentHeat / InterpPhiMultigrid: partition of unity with zero-gradient
  ghosts; InterpTempMultigrid: exact on a linear profile whose wall values
  match the fixed-temperature ghosts.
* masked x-solves: all-liquid masks reproduce the unmasked solve, all-solid
  masks force the velocity to zero.
"""
import numpy as np
import pytest

import pyoracle
from helpers import uniform

PF = dict(nxm=16, nym=12, nzm=10, istr3=8, multires=1, phasefield=1, nxmr=32, nymr=24, nzmr=20, istr3r=8,
          rayt=1e6, prat=1.0, pf_a=1.0, pf_s=1.0, pf_tm=0.0)


def make(cfg=PF):
    return pyoracle.Oracle(cfg)


def test_pf_scalars():
    o = make(dict(PF, pf_a=2.0, pf_s=0.5))
    s = o.scalars()
    pf = o.pf_scalars()
    assert pf["pf_d"] == pytest.approx(1.2 / s["pect"] / 0.5 / 2.0, rel=1e-15)   # ReadInputFile.F90:125
    assert pf["pf_eps"] == pytest.approx(1.0 / 32, rel=1e-15)                    # :126
    assert pf["pf_eta"] == pytest.approx(s["ren"] * (1.51044285 / 32) ** 2, rel=1e-14)
    o2 = make(dict(PF, pf_eps=1.5 / 32))
    assert o2.pf_scalars()["pf_eps"] == 1.5 / 32


def neumann_weights(o, L=1.0):
    """1/a33 of second_derivative_coeff (grid.F90:264-307) for the refined
    cell centres: the weights in which the zero-gradient operator is
    conservative."""
    x = o.grid("xmr")[: PF["nxmr"]]
    n = len(x)
    w = np.empty(n)
    w[0] = 0.5 * (x[0] + x[1])
    w[1:-1] = 0.5 * (x[2:] - x[:-2])
    w[-1] = 0.5 * (2 * L - x[-1] - x[-2])
    return w


def x_operator(o):
    ap, ac, am = o.grid("ap3spkr"), o.grid("ac3spkr"), o.grid("am3spkr")
    n = len(ap)
    M = np.diag(ac) + np.diag(ap[:-1], 1) + np.diag(am[1:], -1)
    return M


def test_neumann_operator_conservative():
    o = make()
    M = x_operator(o)
    w = neumann_weights(o)
    assert np.abs(M.sum(axis=1)).max() < 1e-9 * np.abs(M).max()        # constants are in the kernel
    assert np.abs(w @ M).max() < 1e-11 * np.abs(M).max()                 # flux form, no wall flux


def test_explicit_phase_uniform():
    o = make(dict(PF, pf_a=0.7, pf_tm=0.2))
    c, d = 0.3, 0.45
    o.field("phi")[...] = c
    o.field("tempr")[...] = 0.2 + d
    o.call("explicit_phase")
    pfb = o.pf_scalars()["pf_b"]
    expect = -pfb * c * (1 - c) * (1 - 2 * c + 0.7 * d)
    h = o.field("hphi")[..., : PF["nxmr"]]
    assert np.abs(h - expect).max() < 1e-13 * abs(expect)


def test_explicit_phase_lateral_diffusion():
    o = make()
    shape = o.field("phi").shape
    pat = (uniform(7, shape) > 0.5).astype(float)          # 0/1: the Allen-Cahn source vanishes
    o.field("phi")[...] = pat
    o.field("tempr")[...] = uniform(8, shape)
    o.call("explicit_phase")
    D = o.pf_scalars()["pf_d"]
    dy, dz = PF["nymr"] / 1.0, PF["nzmr"] / 1.0
    lap = (np.roll(pat, -1, 1) - 2 * pat + np.roll(pat, 1, 1)) * dy ** 2 + \
          (np.roll(pat, -1, 0) - 2 * pat + np.roll(pat, 1, 0)) * dz ** 2
    n = PF["nxmr"]
    assert np.abs(o.field("hphi")[..., :n] - D * lap[..., :n]).max() < 1e-10


@pytest.mark.parametrize("ns", [1, 2, 3])
def test_implicit_phase_eigen_decay(ns):
    """A column holding an eigenvector v of the x-operator (L v = lam v) is
    multiplied by the Crank-Nicolson factor (1 + b lam)/(1 - b lam)."""
    o = make()
    M = x_operator(o)
    lam, V = np.linalg.eig(M)
    order = np.argsort(lam.real)
    k = order[-4]                     # a slowly decaying, non-constant mode
    lam_k, v = lam[k].real, V[:, k].real
    n = PF["nxmr"]
    phi = o.field("phi")
    phi[...] = 0.0
    phi[..., :n] = v
    o.field("hphi")[...] = 0.0
    o.field("ruphi")[...] = 0.0
    dt = 4e-2
    o.set_substep(dt, ns)
    o.call("implicit_phase")
    gam, rom = [(8 / 15, 0.0), (5 / 12, -17 / 60), (3 / 4, -5 / 12)][ns - 1]
    b = 0.5 * (gam + rom) * dt * o.pf_scalars()["pf_d"]
    g = (1 + b * lam_k) / (1 - b * lam_k)
    assert np.abs(o.field("phi")[..., :n] - g * v).max() < 1e-12 * np.abs(v).max()
    # rhsr holds the increment (used by AddLatentHeat)
    assert np.abs(o.field("rhsr")[..., :n] - (g - 1) * v).max() < 1e-12 * np.abs(v).max()


def test_implicit_phase_conserves_weighted_sum():
    o = make()
    n = PF["nxmr"]
    phi = o.field("phi")
    phi[...] = uniform(3, phi.shape)
    before = phi[..., :n].copy()
    o.field("hphi")[...] = 0.0
    o.field("ruphi")[...] = 0.0
    o.set_substep(5e-2, 2)
    o.call("implicit_phase")
    w = neumann_weights(o)
    diff = (o.field("phi")[..., :n] - before) @ w
    assert np.abs(diff).max() < 1e-13
    assert np.abs(o.field("phi")[..., :n] - before).max() > 1e-3       # it did diffuse


def test_volume_penalty_closed_form():
    o = make()
    pf = o.pf_scalars()
    nxm = PF["nxm"]
    o.field("phic")[...] = 0.0
    o.field("phic")[..., :nxm] = 1.0
    for f, v in (("vx", 0.3), ("vy", -0.2), ("vz", 0.1)):
        o.field(f)[...] = v
    for f in ("qcap", "dph", "dq"):
        o.field(f)[...] = 0.0
    o.call("add_volume_penalty")
    rate = 1.0 / pf["pf_eta"]              # (phic + phic') heta = 2 * 0.5/eta
    assert np.allclose(o.field("qcap")[..., 1:nxm], -rate * 0.3, rtol=1e-14)
    assert np.all(o.field("qcap")[..., 0] == 0.0)
    assert np.allclose(o.field("dph")[..., :nxm], rate * 0.2, rtol=1e-14)
    assert np.allclose(o.field("dq")[..., :nxm], -rate * 0.1, rtol=1e-14)
    # the y-average takes the neighbour j-1 (periodic): a solid row j0 drags dph at j0 and j0+1
    o.field("phic")[...] = 0.0
    o.field("phic")[:, 5, :nxm] = 1.0
    o.field("dph")[...] = 0.0
    o.call("add_volume_penalty")
    dph = o.field("dph")[..., :nxm]
    assert np.allclose(dph[:, 5], 0.5 * rate * 0.2) and np.allclose(dph[:, 6], 0.5 * rate * 0.2)
    assert np.all(dph[:, [4, 7]] == 0.0)


def test_latent_heat_and_phic_partition_of_unity():
    o = make()
    nxm = PF["nxm"]
    o.set_substep(1e-2, 1)
    o.field("rhsr")[...] = 0.25
    o.field("hro")[...] = 1.0
    o.call("add_latent_heat")
    al_dt = (8 / 15) * 1e-2
    assert np.allclose(o.field("hro")[..., :nxm], 1.0 + 1.0 * 0.25 / al_dt, rtol=1e-13)
    o.field("phi")[...] = 0.7
    o.call("interp_phi_multigrid")
    phic = o.field("phic")
    assert np.allclose(phic[..., :nxm], 0.7, rtol=1e-13) and np.all(phic[..., nxm] == 0.0)


def test_interp_temp_linear_exact():
    o = make(dict(PF, tfixs=1, tfixn=1))
    xm = o.grid("xm")[: PF["nxm"]]
    xmr = o.grid("xmr")[: PF["nxmr"]]
    a, b = 0.8, -0.6
    T = o.field("temp")
    T[...] = 0.0
    T[..., : PF["nxm"]] = a + b * xm
    o.field("tempbp")[...] = a            # wall values of the same line: ghosts continue it
    o.field("temptp")[...] = a + b
    o.call("interp_temp_multigrid")
    tr = o.field("tempr")[..., : PF["nxmr"]]
    assert np.abs(tr - (a + b * xmr)).max() < 1e-12


@pytest.mark.parametrize("routine,field", [("implicit_and_update_vx", "vx"), ("implicit_and_update_vy", "vy"),
                                           ("implicit_and_update_vz", "vz")])
def test_masked_solves(routine, field):
    base = dict(PF, phasefield=0, multires=0)
    st = {n: uniform(20 + q, make(base).field(n).shape) - 0.5
          for q, n in enumerate(["vx", "vy", "vz", "pr", "qcap", "dph", "dq", "rux", "ruy", "ruz"])}
    st["vx"][..., 0] = 0.0
    st["vx"][..., -1] = 0.0
    ref = make(base)
    for n, a in st.items():
        ref.field(n)[...] = a
    ref.set_substep(1e-3, 2)
    ref.call(routine)
    for val, liquid in ((0.0, True), (1.0, False)):
        o = make()
        for n, a in st.items():
            o.field(n)[...] = a
        o.field("phic")[...] = val
        o.set_substep(1e-3, 2)
        o.call(routine)
        got = o.field(field)
        if liquid:
            assert np.abs(got - ref.field(field)).max() < 1e-14
        else:
            nxm = PF["nxm"]
            rows = np.s_[..., 1:nxm] if field == "vx" else np.s_[..., :nxm]
            assert np.abs(got[rows]).max() == 0.0
