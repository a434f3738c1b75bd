"""GPU parity of the moisture routines (src/moisture.F90) against the oracle
on the same seeded random states: SetHumidityBCs, UpdateSaturation,
ExplicitHumidity, AddCondensation, ImplicitHumidity with its solve, and the
fused step with the humidity column kernel.

Tolerances: FP64 GPU against the FP64 oracle, the same arithmetic in
another order (wavefront scans for the Thomas sweeps): 1e-12 relative to each
field's max norm per routine, 1e-10 after whole steps (pressure modulo its
mean).  The mixed form keeps humidity in FP64 but its velocity comes from
FP32 explicit terms: relative L2 <= 1e-5 against the oracle after 3 steps,
Nu within 1e-6.
"""
import numpy as np
import pytest

import pyafid
import pyoracle
from helpers import load_state, random_state, relerr, uniform

pytestmark = pytest.mark.gpu

MOIST = dict(moist=1, q_alpha=3.0, q_beta=1.2, q_gamma=0.19, q_tau=5e-5, q_ifdiff=1, q_fixn=1, q_fixs=1, q_sm=2.0)
CFG = dict(nxm=32, nym=24, nzm=16, istr3=8, rayt=1e5, prat=1.0, **MOIST)
ODD = dict(nxm=70, nym=12, nzm=18, istr3=4, str3=1.5, rayt=2e6, prat=0.7, ylen=1.5, zlen=1.0,
           **dict(MOIST, q_gamma=-1.0, q_fixn=2, q_fixs=0, q_sm=0.7, q_ifdiff=0, q_tau=0.02))


def moist_pair(cfg, seed, form="fp64"):
    """A solver and an oracle holding the same random state, humidity
    between 0.5 and 1.5, and the humidity walls of SetHumidityBCs."""
    st = random_state(cfg, seed)
    nz, ny, nx = st["vx"].shape
    st["humid"] = 1.0 + (uniform(seed + 40, (nz, ny, nx)) - 0.5)
    st["humid"][..., nx - 1] = 0.0
    st["ruhumid"] = uniform(seed + 41, (nz, ny, nx)) - 0.5
    st["tempbp"] = 0.2 * uniform(seed + 42, (nz, ny, 1))
    st["temptp"] = -0.2 * uniform(seed + 43, (nz, ny, 1))
    g = pyafid.Solver(dict(cfg, precision=pyafid.MIXED if form == "mixed" else pyafid.FP64))
    o = pyoracle.Oracle(cfg)
    load_state(g, st, upload=True)
    load_state(o, st, upload=False)
    g.call("set_humidity_bcs")
    o.call("set_humidity_bcs")
    g.call("update_saturation")
    o.call("update_saturation")
    return g, o


def interior(cfg, a):
    return a[..., :cfg["nxm"]]


@pytest.mark.parametrize("cfg", [CFG, ODD], ids=["rbc", "odd"])
def test_bcs_and_saturation(cfg):
    g, o = moist_pair(cfg, 3)
    for n in ("humbp", "humtp"):
        assert relerr(g.download(n), o.field(n)) <= 1e-15, n
    e = relerr(interior(cfg, g.download("qsat")), interior(cfg, o.field("qsat")))
    assert e <= 1e-14, e
    q, qs = interior(cfg, o.field("humid")), interior(cfg, o.field("qsat"))
    assert 0.1 < np.mean(q > qs) < 0.9            # both branches of the condensation switch occur
    g.close(); o.close()


@pytest.mark.parametrize("ns", [1, 2, 3])
@pytest.mark.parametrize("cfg", [CFG, ODD], ids=["rbc", "odd"])
def test_humidity_routines(cfg, ns):
    """ExplicitHumidity -> rkhumid; ExplicitTermsTemp + AddCondensation ->
    hro; ImplicitHumidity -> humid, hkhumid and the increment in rhs."""
    g, o = moist_pair(cfg, 5 + ns)
    for obj in (g, o):
        obj.set_substep(2e-3, ns)
        for r in ("explicit_terms_temp", "explicit_humidity", "add_condensation"):
            obj.call(r)
    for n in ("hhumid", "hro"):
        e = relerr(interior(cfg, g.download(n)), interior(cfg, o.field(n)))
        assert e <= 1e-12, (n, e)
    for obj in (g, o):
        obj.call("implicit_humidity")
    for n in ("humid", "ruhumid", "rhs"):
        e = relerr(interior(cfg, g.download(n)), interior(cfg, o.field(n)))
        assert e <= 1e-12, (n, e)
    g.close(); o.close()


@pytest.mark.parametrize("cfg", [CFG, ODD], ids=["rbc", "odd"])
def test_moist_time_marcher(cfg):
    """Three fused steps against the oracle's TimeMarcher with moisture."""
    g, o = moist_pair(cfg, 11)
    for _ in range(3):
        g.time_marcher(1e-3)
        o.time_marcher(1e-3)
    for n in ("vx", "vy", "vz", "temp", "humid", "qsat", "ruhumid", "pr"):
        a, b = g.download(n), o.field(n)
        if n != "vx":
            a, b = interior(cfg, a), interior(cfg, b)
        if n == "pr":
            a, b = a - a.mean(), b - b.mean()
        e = relerr(a, b)
        assert e <= 1e-10, (n, e)
    qg, qo = g.global_quantities(), o.global_quantities()
    assert abs(qg["nu"] - qo["nu"]) <= 1e-10 * abs(qo["nu"])
    g.close(); o.close()


def test_moist_mixed_vs_oracle():
    g, o = moist_pair(CFG, 13, form="mixed")
    for _ in range(3):
        g.time_marcher(1e-3)
        o.time_marcher(1e-3)
    d, _ = g.check_divergence()
    assert d <= 1e-10
    for n in ("vx", "vy", "vz", "temp", "humid"):
        a, b = interior(CFG, g.download(n)), interior(CFG, o.field(n))
        e = float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
        assert e <= 1e-5, (n, e)
    qg, qo = g.global_quantities(), o.global_quantities()
    assert abs(qg["nu"] - qo["nu"]) <= 1e-6 * abs(qo["nu"])
    g.close(); o.close()


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_moist_graph_replay_bit_identical(form):
    a, _ = moist_pair(CFG, 17, form)
    b, _ = moist_pair(CFG, 17, form)
    for rep in range(2):
        for _ in range(2):
            a.time_marcher(1e-3)
        b.time_marcher_graph(1e-3, 2)
        assert a.check_divergence() == b.check_divergence()
        for n in ("vx", "temp", "humid", "qsat"):
            assert np.array_equal(a.download(n), b.download(n)), (rep, n)
    a.close(); b.close()


def test_moisture_calls_need_moist():
    g = pyafid.Solver(dict(nxm=16, nym=8, nzm=8))
    g.set_substep(1e-3, 1)
    for r in ("set_humidity_bcs", "explicit_humidity", "add_condensation", "implicit_humidity", "update_saturation"):
        with pytest.raises(pyafid.AfidError, match="moisture is not enabled"):
            g.call(r)
    g.close()


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_moist_stats(form):
    """CalcMoistStats against the oracle on the same state (FP64 sums in a
    fixed order on both sides: 1e-12 of each profile's max norm)."""
    g, o = moist_pair(ODD, 19, form)
    a, b = g.moist_stats(), o.moist_stats()
    for n in pyafid.MOIST_PROFILES:
        e = relerr(a[n], b[n])
        assert e <= 1e-12, (n, e)
    g.close(); o.close()


def test_humidity_restart_onto_new_grid():
    """InterpInputHum (moisture.F90:539-582) onto a new grid: the library
    against the oracle from the same old field and wall values."""
    new = dict(CFG, nxm=40, nym=12, nzm=24, istr3=4, str3=1.5)
    old = dict(nx=CFG["nxm"] + 1, ny=CFG["nym"] + 1, nz=CFG["nzm"] + 1, istr3=8, str3=0.0)
    Q = 1.0 + uniform(23, (CFG["nzm"], CFG["nym"], CFG["nxm"] + 1))
    zero = np.zeros_like(Q)
    g = pyafid.Solver(new)
    o = pyoracle.Oracle(new)
    for obj, up in ((g, True), (o, False)):
        planes = {"tempbp": np.full((24, 12, 1), 0.1), "temptp": np.full((24, 12, 1), 0.4)}
        load_state(obj, planes, upload=up)
        obj.call("set_humidity_bcs")
    g.interp_input(old, dict(vx=zero, vy=zero, vz=zero, temp=zero, humid=Q))
    o.interp_input_humid(pyoracle.old_grid(dict(nxm=CFG["nxm"], nym=CFG["nym"], nzm=CFG["nzm"], istr3=8, str3=0.0)), Q)
    e = relerr(g.download("humid")[..., :40], o.field("humid")[..., :40])
    assert e <= 1e-13, e
    assert not np.any(g.download("ruhumid"))
    g.close(); o.close()
