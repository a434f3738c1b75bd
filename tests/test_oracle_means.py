"""Pins of the oracle's mean profiles (CalcMeanProfiles.F90:11-229) and
refined salinity profiles (CalcSalStats, salinity.F90:510-690), CPU only.

Each check compares the oracle with something it shares no code with:
closed forms of the discrete operators on fields whose profiles are known
exactly (a conductive temperature profile, a linear shear, single Fourier
modes in y and z), and two invariances of the wall-parallel averages on
random fields (exchange of y and z; periodic shifts).  A wrong neighbour
index, a dropped term, a wrong wall form or a wrong normalisation fails at
least one of them.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import random_state, load_state, uniform

po.set_threads(2)

NAMES = po.Oracle.MEAN_PROFILES


def consts(cfg):
    c = dict(po.DEFAULTS, **cfg)
    ren = np.sqrt(abs(c["rayt"]) / c["prat"])
    return ren, ren * c["prat"]          # ren, pect


def fresh(cfg):
    o = po.Oracle(cfg)
    for n in ("vx", "vy", "vz", "temp"):
        o.field(n)[...] = 0.0
    o.field("tempbp")[...] = 1.0
    o.field("temptp")[...] = 0.0
    return o


@pytest.mark.parametrize("istr3", [0, 8])
def test_conductive_profile_dissipation_is_one_over_pect(istr3):
    """T = 1 - x at the cell centres with T = 1, 0 on the walls: every
    difference quotient of chiT (interior and both wall forms) is the slope
    -1 exactly, on any x grid, so chiT = 1/pect; Tbar = T, Trms = |T|."""
    cfg = dict(nxm=12, nym=6, nzm=10, istr3=istr3, str3=0.0 if istr3 == 0 else 0.0, rayt=3e5, prat=2.0)
    o = fresh(cfg)
    xm = o.grid("xm")[:12]
    T = 1.0 - xm
    o.field("temp")[..., :12] = T
    m = o.mean_profiles()
    _, pect = consts(cfg)
    assert np.allclose(m["chiT"], 1.0 / pect, rtol=1e-13, atol=0)
    assert np.allclose(m["Tbar"], T, rtol=1e-14, atol=0)
    assert np.allclose(m["Trms"], np.abs(T), rtol=1e-14, atol=0)
    for n in ("vybar", "vzbar", "vxT", "vyT", "vzT", "vxrms", "vyrms", "vzrms", "vxvy", "vxvz", "vyvz", "epsilon"):
        assert np.all(m[n] == 0.0), n


def test_linear_shear_dissipation():
    """vx = c x on the x faces: (du/dx)^2 = c^2 exactly (udx3m = 1/(xc(k+1)-xc(k)))
    and no other term of epsilon is nonzero, so epsilon = c^2/ren; vxrms is the
    face average of vx^2."""
    cfg = dict(nxm=10, nym=4, nzm=6, istr3=8, rayt=1e4, prat=0.7)
    o = fresh(cfg)
    xc = o.grid("xc")[:11]
    c = 1.7
    o.field("vx")[..., :11] = c * xc
    m = o.mean_profiles()
    ren, _ = consts(cfg)
    assert np.allclose(m["epsilon"], c * c / ren, rtol=1e-13, atol=0)
    assert np.allclose(m["vxrms"], np.sqrt(0.5 * ((c * xc[:-1]) ** 2 + (c * xc[1:]) ** 2)), rtol=1e-14, atol=0)
    assert np.all(m["vxT"] == 0.0) and np.all(m["vxvy"] == 0.0)


@pytest.mark.parametrize("axis,mode", [("z", 1), ("z", 3), ("y", 2)])
def test_wave_dissipation_discrete_symbol(axis, mode):
    """A single mode w = A sin(2 pi m i/N) in the velocity component along its
    own axis: the forward difference has mean square 2 A^2 sin^2(pi m/N), so
    epsilon = 2 A^2 sin^2(pi m/N) d^2 / ren; its rms is A/sqrt(2), its mean 0.
    Free-slip walls (inslws = inslwn = 0) leave the wall x-terms zero."""
    N = 12
    cfg = dict(nxm=8, nym=N if axis == "y" else 4, nzm=N if axis == "z" else 4, istr3=8, rayt=1e5, prat=1.0,
               inslws=0, inslwn=0, ylen=1.5, zlen=2.0)
    o = fresh(cfg)
    A = 0.3
    th = 2 * np.pi * mode * np.arange(N) / N
    if axis == "z":
        o.field("vz")[..., :8] = (A * np.sin(th))[:, None, None]
        d = N / cfg["zlen"]
        name = "vz"
    else:
        o.field("vy")[..., :8] = (A * np.sin(th))[None, :, None]
        d = N / cfg["ylen"]
        name = "vy"
    m = o.mean_profiles()
    ren, _ = consts(cfg)
    assert np.allclose(m["epsilon"], 2 * A * A * np.sin(np.pi * mode / N) ** 2 * d * d / ren, rtol=1e-12, atol=0)
    assert np.allclose(m[name + "rms"], A / np.sqrt(2), rtol=1e-13, atol=0)
    assert np.all(np.abs(m[name + "bar"]) < 1e-15)


def test_cross_derivative_symbol():
    """vy = A sin(2 pi m i/N) varying in z: the centred (dv/dz)^2 term of
    epsilon, 0.125 dz^2 [(vy(jp,ip)-vy(jp,im))^2 + (vy(j,ip)-vy(j,im))^2], has
    mean 0.5 A^2 sin^2(2 pi m/N) dz^2; nothing else contributes."""
    N, mode, A = 10, 2, 0.4
    cfg = dict(nxm=6, nym=4, nzm=N, istr3=8, rayt=1e5, prat=1.0, inslws=0, inslwn=0, zlen=1.3)
    o = fresh(cfg)
    o.field("vy")[..., :6] = (A * np.sin(2 * np.pi * mode * np.arange(N) / N))[:, None, None]
    m = o.mean_profiles()
    ren, _ = consts(cfg)
    dz = N / cfg["zlen"]
    assert np.allclose(m["epsilon"], 0.5 * A * A * np.sin(2 * np.pi * mode / N) ** 2 * dz * dz / ren, rtol=1e-12, atol=0)


def test_correlations_of_separable_fields():
    """vx = a, vy = b + A cos(y mode), vz = d + B sin(z mode), T = T(x): the
    face averages of the modes vanish, so vxvy = a b, vxvz = a d, vyvz = b d,
    vyT = b T, vzT = d T, vxT = a T, vybar = b, vzbar = d."""
    cfg = dict(nxm=6, nym=8, nzm=10, istr3=8)
    o = fresh(cfg)
    a, b, d, A, B = 0.3, -0.7, 0.45, 0.2, 0.6
    o.field("vx")[...] = a
    o.field("vy")[..., :6] = (b + A * np.cos(2 * np.pi * 3 * np.arange(8) / 8))[None, :, None]
    o.field("vz")[..., :6] = (d + B * np.sin(2 * np.pi * 2 * np.arange(10) / 10))[:, None, None]
    T = 0.5 + 0.1 * np.arange(6)
    o.field("temp")[..., :6] = T
    m = o.mean_profiles()
    for n, v in (("vxvy", a * b), ("vxvz", a * d), ("vyvz", b * d), ("vybar", b), ("vzbar", d)):
        assert np.allclose(m[n], v, rtol=1e-13, atol=1e-15), n
    assert np.allclose(m["vyT"], b * T, rtol=1e-13) and np.allclose(m["vzT"], d * T, rtol=1e-13)
    assert np.allclose(m["vxT"], a * T, rtol=1e-14)


def _random(cfg, seed):
    o = po.Oracle(cfg)
    st = random_state(cfg, seed)
    load_state(o, st, upload=False)
    o.field("tempbp")[...] = 1.0 + 0.1 * uniform(seed + 7, o.field("tempbp").shape)
    o.field("temptp")[...] = 0.1 * uniform(seed + 8, o.field("temptp").shape)
    return o


def test_y_z_exchange_symmetry():
    """Exchanging y and z (and vy with vz, with dy = dz and zero wall velocity
    offsets) exchanges vybar/vzbar, vyT/vzT, vyrms/vzrms, vxvy/vxvz and leaves
    the other profiles unchanged: each derivative term of epsilon and chiT
    maps onto its partner."""
    cfg = dict(nxm=10, nym=8, nzm=8, istr3=8, ylen=1.0, zlen=1.0)
    o = _random(cfg, 5)
    m = o.mean_profiles()
    q = po.Oracle(cfg)
    sw = lambda a: np.ascontiguousarray(np.swapaxes(a, 0, 1))
    q.field("vx")[...] = sw(o.field("vx"))
    q.field("vy")[...] = sw(o.field("vz"))
    q.field("vz")[...] = sw(o.field("vy"))
    q.field("temp")[...] = sw(o.field("temp"))
    q.field("tempbp")[...] = sw(o.field("tempbp"))
    q.field("temptp")[...] = sw(o.field("temptp"))
    n = q.mean_profiles()
    pairs = {"vybar": "vzbar", "vyT": "vzT", "vyrms": "vzrms", "vxvy": "vxvz"}
    pairs.update({v: k for k, v in pairs.items()})
    for k in NAMES:
        assert np.allclose(n[pairs.get(k, k)], m[k], rtol=1e-12, atol=1e-15), k


def test_periodic_shift_invariance():
    """Rolling every field (and the wall planes) in y and z leaves the
    wall-parallel means unchanged: the neighbours wrap periodically."""
    cfg = dict(nxm=8, nym=6, nzm=10, istr3=8)
    o = _random(cfg, 9)
    m = o.mean_profiles()
    q = po.Oracle(cfg)
    for n in ("vx", "vy", "vz", "temp", "tempbp", "temptp"):
        q.field(n)[...] = np.roll(np.roll(o.field(n), 3, axis=0), -2, axis=1)
    r = q.mean_profiles()
    for k in NAMES:
        assert np.allclose(r[k], m[k], rtol=1e-12, atol=1e-15), k


# ------------------------------------------------- CalcSalStats (refined grid)
SAL_NAMES = po.Oracle.SAL_PROFILES


def mr_cfg(**kw):
    c = dict(nxm=8, nym=4, nzm=6, istr3=8, multires=1, salinity=1, nxmr=16, nymr=8, nzmr=12, istr3r=8,
             rayt=1e6, prat=1.0, rays=-1e6, pras=4.0)
    c.update(kw)
    return c


def pecs_of(cfg):
    ren, _ = consts(cfg)
    return ren * cfg["pras"]


def fresh_mr(cfg):
    o = po.Oracle(cfg)
    for n in ("sal", "vxr", "vyr", "vzr"):
        o.field(n)[...] = 0.0
    o.field("salbp")[...] = 1.0
    o.field("saltp")[...] = 0.0
    return o


def test_sal_conductive_profile_dissipation_is_one_over_pecs():
    """S = 1 - x on the refined cell centres with S = 1, 0 on the walls:
    chiS = 1/pecs exactly (interior and both wall forms), Sbar = S, Srms = |S|."""
    cfg = mr_cfg()
    o = fresh_mr(cfg)
    xmr = o.grid("xmr")[:16]
    S = 1.0 - xmr
    o.field("sal")[..., :16] = S
    m = o.sal_stats()
    assert np.allclose(m["chiS"], 1.0 / pecs_of(cfg), rtol=1e-13, atol=0)
    assert np.allclose(m["Sbar"], S, rtol=1e-14, atol=0)
    assert np.allclose(m["Srms"], np.abs(S), rtol=1e-14, atol=0)
    for n in ("vxS", "vyS", "vzS"):
        assert np.all(m[n] == 0.0), n


def test_sal_fluxes_of_separable_fields():
    """vxr = a, vyr = b + A cos(y mode), vzr = d + B sin(z mode), S = S(x): the
    face averages of the modes vanish, so vxS = a S, vyS = b S, vzS = d S."""
    cfg = mr_cfg()
    o = fresh_mr(cfg)
    a, b, d = 0.4, -0.3, 0.25
    o.field("vxr")[...] = a
    o.field("vyr")[..., :16] = (b + 0.2 * np.cos(2 * np.pi * 3 * np.arange(8) / 8))[None, :, None]
    o.field("vzr")[..., :16] = (d + 0.5 * np.sin(2 * np.pi * 5 * np.arange(12) / 12))[:, None, None]
    S = 0.2 + 0.05 * np.arange(16)
    o.field("sal")[..., :16] = S
    m = o.sal_stats()
    assert np.allclose(m["vxS"], a * S, rtol=1e-14)
    assert np.allclose(m["vyS"], b * S, rtol=1e-13)
    assert np.allclose(m["vzS"], d * S, rtol=1e-13)


def test_sal_dissipation_discrete_symbol():
    """S = A sin(2 pi m j/N) in y with no wall flux terms (sfixs = sfixn = 0):
    the centred y difference has mean square 2 A^2 sin^2(2 pi m/N), so chiS =
    0.5 A^2 sin^2(2 pi m/N) dyr^2 / pecs."""
    cfg = mr_cfg(sfixs=0, sfixn=0, ylen=1.7)
    o = fresh_mr(cfg)
    A, mode, N = 0.6, 3, 8
    o.field("sal")[..., :16] = (A * np.sin(2 * np.pi * mode * np.arange(N) / N))[None, :, None]
    m = o.sal_stats()
    dyr = N / cfg["ylen"]
    assert np.allclose(m["chiS"], 0.5 * A * A * np.sin(2 * np.pi * mode / N) ** 2 * dyr * dyr / pecs_of(cfg), rtol=1e-12, atol=0)


def _random_mr(cfg, seed):
    o = po.Oracle(cfg)
    for q, n in enumerate(("sal", "vxr", "vyr", "vzr")):
        o.field(n)[...] = uniform(seed * 10 + q, o.field(n).shape) - 0.5
    o.field("salbp")[...] = 1.0 + 0.1 * uniform(seed + 7, o.field("salbp").shape)
    o.field("saltp")[...] = 0.1 * uniform(seed + 8, o.field("saltp").shape)
    return o


def test_sal_y_z_exchange_and_shift():
    """Exchanging y and z (vyr with vzr) swaps vyS and vzS and leaves the rest;
    a periodic roll of every field leaves all six profiles unchanged."""
    cfg = mr_cfg(nym=4, nzm=4, nymr=8, nzmr=8, ylen=1.0, zlen=1.0)
    o = _random_mr(cfg, 4)
    m = o.sal_stats()
    q = po.Oracle(cfg)
    sw = lambda a: np.ascontiguousarray(np.swapaxes(a, 0, 1))
    for src, dst in (("sal", "sal"), ("vxr", "vxr"), ("vyr", "vzr"), ("vzr", "vyr"), ("salbp", "salbp"), ("saltp", "saltp")):
        q.field(dst)[...] = sw(o.field(src))
    n = q.sal_stats()
    pairs = {"vyS": "vzS", "vzS": "vyS"}
    for k in SAL_NAMES:
        assert np.allclose(n[pairs.get(k, k)], m[k], rtol=1e-12, atol=1e-15), k
    r = po.Oracle(cfg)
    for f in ("sal", "vxr", "vyr", "vzr", "salbp", "saltp"):
        r.field(f)[...] = np.roll(np.roll(o.field(f), 2, axis=0), -3, axis=1)
    t = r.sal_stats()
    for k in SAL_NAMES:
        assert np.allclose(t[k], m[k], rtol=1e-12, atol=1e-15), k


# ------------------------------------------------- CalcPhiStats (refined grid)
def test_phi_stats_closed_forms():
    """phi = f(x) + c + A sin(2 pi m j/N): the y mode averages out, so phibar =
    f + c and phirms^2 = (f + c)^2 + A^2/2; a periodic roll changes nothing."""
    cfg = dict(nxm=8, nym=4, nzm=6, istr3=8, multires=1, phasefield=1, nxmr=16, nymr=12, nzmr=10, istr3r=8)
    o = po.Oracle(cfg)
    f = 0.1 * np.arange(16) - 0.4
    A, c = 0.3, 0.25
    phi = np.zeros(o.field("phi").shape)
    phi[..., :16] = f + c + (A * np.sin(2 * np.pi * 5 * np.arange(12) / 12))[None, :, None]
    o.field("phi")[...] = phi
    m = o.phi_stats()
    assert np.allclose(m["phibar"], f + c, rtol=1e-13, atol=1e-15)
    assert np.allclose(m["phirms"], np.sqrt((f + c) ** 2 + A * A / 2), rtol=1e-13, atol=0)
    o.field("phi")[...] = np.roll(np.roll(phi, 4, axis=0), -5, axis=1)
    r = o.phi_stats()
    assert np.allclose(r["phibar"], m["phibar"], rtol=1e-13, atol=1e-15)
    assert np.allclose(r["phirms"], m["phirms"], rtol=1e-13, atol=0)
