"""GPU parity: every entry point of the C ABI against the FP64 CPU oracle on
the same seeded inputs (small grids the oracle finishes in milliseconds).

Tolerance: FP64 in both; the kernels reorder sums, fuse multiply-adds and use
a different FFT factorisation, so agreement is to round-off, stated per check
(relative to the field's max norm).
"""
import numpy as np
import pytest

import pyafid
import pyoracle
from helpers import load_state, random_state, relerr, uniform

pytestmark = pytest.mark.gpu

BASE = dict(nxm=16, nym=12, nzm=8, istr3=8, rayt=1e5, prat=1.0)
MR = dict(nxm=16, nym=8, nzm=12, istr3=8, multires=1, salinity=1, nxmr=32, nymr=16, nzmr=24, istr3r=8,
          rayt=1e6, prat=1.0, rays=-1e6, pras=4.0)
TOL = 1e-12
TOL_POISSON = 1e-11


def make_pair(cfg, seed=3):
    st = random_state(cfg, seed)
    o = pyoracle.Oracle(cfg)
    g = pyafid.Solver(cfg)
    load_state(o, st, upload=False)
    load_state(g, st, upload=True)
    return o, g, st


def put(o, g, name, a):
    o.field(name)[...] = a
    g.upload(name, a)


def cmp(o, g, name, sl=np.s_[...], tol=TOL):
    e = relerr(g.download(name)[sl], o.field(name)[sl])
    assert e <= tol, f"{name}: relative error {e:.3e} > {tol:.1e}"
    return e


@pytest.mark.parametrize("routine,field,xs", [
    ("explicit_terms_vx", "qcap", np.s_[..., 1:16]),
    ("explicit_terms_vy", "dph", np.s_[..., :16]),
    ("explicit_terms_vz", "dq", np.s_[..., :16]),
    ("explicit_terms_temp", "hro", np.s_[..., :16]),
])
@pytest.mark.parametrize("gaxis", [1, 2, 3])
def test_explicit_terms(routine, field, xs, gaxis):
    cfg = dict(BASE, gaxis=gaxis, dpdy=0.3, inslwn=1, inslws=0)
    o, g, _ = make_pair(cfg)
    for ns in (1, 2):
        o.set_substep(1e-3, ns); g.set_substep(1e-3, ns)
        o.call(routine); g.call(routine)
        cmp(o, g, field, xs)


@pytest.mark.parametrize("routine,inp,outs", [
    ("implicit_and_update_vx", "qcap", ["vx", "rux"]),
    ("implicit_and_update_vy", "dph", ["vy", "ruy"]),
    ("implicit_and_update_vz", "dq", ["vz", "ruz"]),
    ("implicit_and_update_temp", "hro", ["temp", "rutemp"]),
])
@pytest.mark.parametrize("ns", [1, 2, 3])
def test_implicit_update(routine, inp, outs, ns):
    cfg = dict(BASE, xplusu=0.2, xminusu=-0.1)
    o, g, st = make_pair(cfg)
    put(o, g, inp, uniform(77, st["vx"].shape) - 0.5)
    o.set_substep(2e-3, ns); g.set_substep(2e-3, ns)
    o.call(routine); g.call(routine)
    for f in outs:
        cmp(o, g, f, np.s_[..., :16] if f != "vx" else np.s_[...])


def test_vx_walls_zero():
    o, g, st = make_pair(BASE)
    put(o, g, "qcap", uniform(5, st["vx"].shape))
    g.set_substep(1e-3, 1)
    g.call("implicit_and_update_vx")
    vx = g.download("vx")
    assert np.all(vx[..., 0] == 0.0) and np.all(vx[..., -1] == 0.0)


def test_local_divergence():
    o, g, _ = make_pair(BASE)
    o.set_substep(1e-3, 2); g.set_substep(1e-3, 2)
    o.call("calc_local_divergence"); g.call("calc_local_divergence")
    cmp(o, g, "dph", np.s_[..., :16])


@pytest.mark.parametrize("shape", [(16, 12, 8), (24, 16, 16), (16, 1, 32), (32, 6, 9)])
def test_pressure_correction(shape):
    nxm, nym, nzm = shape
    cfg = dict(BASE, nxm=nxm, nym=nym, nzm=nzm, istr3=8 if nxm != 32 else 6, str3=4.0)
    o, g, st = make_pair(cfg)
    put(o, g, "dph", uniform(11, st["vx"].shape) - 0.5)
    o.set_substep(1e-3, 1); g.set_substep(1e-3, 1)
    o.call("solve_pressure_correction"); g.call("solve_pressure_correction")
    cmp(o, g, "dph", np.s_[..., :nxm], tol=TOL_POISSON)


def test_projection_is_divergence_free():
    """After local divergence -> Poisson -> projection the discrete divergence
    vanishes to round-off (the defining property of the projection)."""
    o, g, _ = make_pair(dict(BASE, nxm=24, nym=16, nzm=12))
    g.set_substep(1e-3, 1)
    d0, _ = g.check_divergence()
    for r in ("calc_local_divergence", "solve_pressure_correction", "correct_velocity"):
        g.call(r)
    d1, _ = g.check_divergence()
    assert d0 > 1.0 and d1 < 1e-11 * d0


@pytest.mark.parametrize("routine,fields", [
    ("correct_velocity", ["vx", "vy", "vz"]),
    ("correct_pressure", ["pr"]),
])
def test_projection(routine, fields):
    o, g, st = make_pair(BASE)
    put(o, g, "dph", uniform(12, st["vx"].shape) - 0.5)
    o.set_substep(1e-3, 3); g.set_substep(1e-3, 3)
    o.call(routine); g.call(routine)
    for f in fields:
        cmp(o, g, f, np.s_[..., :16] if f != "vx" else np.s_[...])


def test_reductions():
    o, g, _ = make_pair(BASE)
    a = np.array(g.calc_max_cfl()); b = np.array(o.calc_max_cfl())
    assert relerr(a, b) < 1e-14
    a = np.array(g.check_divergence()); b = np.array(o.check_divergence())
    assert relerr(a, b) < 1e-14
    gq, oq = g.global_quantities(), o.global_quantities()
    for k in ("nu", "ke", "vrms", "tmean"):
        assert abs(gq[k] - oq[k]) <= 1e-13 * max(1.0, abs(oq[k])), k


@pytest.mark.parametrize("cfg", [BASE, dict(BASE, nym=1, nzm=16, gaxis=3), dict(BASE, nsst=1)])
def test_time_marcher(cfg):
    o, g, _ = make_pair(cfg)
    for _ in range(3):
        o.time_marcher(2e-3)
        g.time_marcher(2e-3)
    nxm = cfg["nxm"]
    for f in ("vx", "vy", "vz", "temp", "pr", "rux", "ruy", "ruz", "rutemp"):
        cmp(o, g, f, np.s_[..., :nxm] if f != "vx" else np.s_[...], tol=1e-10)


# ----------------------------------------------------------------- multires
def make_mr_pair(cfg=MR, seed=4):
    o, g, st = make_pair(cfg, seed)
    shr = (cfg["nzmr"], cfg["nymr"], cfg["nxmr"] + 1)
    for q, f in enumerate(("vxr", "vyr", "vzr")):
        put(o, g, f, uniform(900 + q, shr) - 0.5)
    put(o, g, "salc", uniform(950, st["vx"].shape) - 0.5)
    return o, g, st


def test_mgrd_setup_tables():
    o, g, _ = make_mr_pair()
    for name in ("xcr", "xmr", "udx3mr", "ap3sskr", "ac3sskr", "am3sskr", "cxrs", "cyrs", "czrs",
                 "cysalc", "czsalc"):
        a, b = g.grid_array(name), o.grid(name)
        n = min(len(a), len(b))
        assert relerr(a[:n], b[:n]) < 1e-15, name
    a, b = g.grid_array("cxsalc"), o.grid("cxsalc")
    assert relerr(a[:len(b)], b) < 1e-15


MR_ODD = dict(MR, nxm=12, nym=8, nzm=6, nxmr=30, nymr=20, nzmr=18, istr3=4, str3=1.4, istr3r=6)
MR_BIG = dict(MR, nxm=40, nym=24, nzm=36, nxmr=80, nymr=72, nzmr=72)


@pytest.mark.parametrize("cfg", [MR, MR_ODD, MR_BIG])
def test_interp_vel_mgrd(cfg):
    o, g, _ = make_mr_pair(cfg)
    o.call("interp_vel_mgrd"); g.call("interp_vel_mgrd")
    for f in ("vxr", "vyr", "vzr"):
        cmp(o, g, f, tol=1e-12)


# refined x barely finer than the base (84 over 80 cells): a chunk of 64
# refined x points spans more than 64 coarse points, so vyr / vzr take the
# second per-lane register of their wavefront kernel.  In FP64 the vxr tile
# of such a grid exceeds 64 KiB of LDS and set-up refuses it; the mixed form
# (FP32 transfers) runs it.
MR_FLATX = dict(MR, nxm=80, nym=8, nzm=12, nxmr=84, nymr=16, nzmr=24)


def test_interp_vel_mgrd_refined_x_near_base(monkeypatch):
    """FP32 transfers of FP64 velocities against the FP64 oracle: relative
    max-norm error <= 1e-5 (FP32 rounding of a 4 x 4 x 4 contraction and a
    16-step running sum).  The LDS tile form of vyr / vzr (AFID_VYZ_TILE=1)
    forms the same sums in the same order: within 1e-6."""
    st = random_state(MR_FLATX, 5)
    o = pyoracle.Oracle(MR_FLATX)
    load_state(o, st, upload=False)
    o.call("interp_vel_mgrd")
    got = {}
    for tile in ("0", "1"):
        monkeypatch.setenv("AFID_VYZ_TILE", tile)
        g = pyafid.Solver(dict(MR_FLATX, precision=pyafid.MIXED))
        load_state(g, st, upload=True)
        g.call("interp_vel_mgrd")
        got[tile] = {f: g.download(f) for f in ("vxr", "vyr", "vzr")}
        g.close()
    for f in ("vxr", "vyr", "vzr"):
        e = relerr(got["0"][f], o.field(f))
        assert e <= 1e-5, (f, e)
        assert relerr(got["0"][f], got["1"][f]) <= 1e-6, f
    o.close()


def test_interp_vel_mgrd_refined_divergence():
    """InterpVelMgrd integrates transferred gradients, so a discretely
    divergence-free coarse velocity (projected on the GPU) gives a refined
    velocity whose discrete divergence is round-off (docs/interpolation.md:41-63)."""
    _, g, _ = make_mr_pair()
    g.set_substep(1e-3, 1)
    for r in ("calc_local_divergence", "solve_pressure_correction", "correct_velocity", "interp_vel_mgrd"):
        g.call(r)
    d, dr = g.check_divergence()
    scale = np.abs(g.download("vxr")).max() * 2 * MR["nxmr"]
    assert d < 1e-12 * scale and dr < 1e-11 * scale


@pytest.mark.parametrize("cfg", [MR, MR_ODD, MR_BIG])
def test_interp_sal_multigrid(cfg):
    o, g, _ = make_mr_pair(cfg)
    o.call("interp_sal_multigrid"); g.call("interp_sal_multigrid")
    cmp(o, g, "salc", tol=1e-13)


@pytest.mark.parametrize("ns", [1, 3])
def test_salinity_substep(ns):
    o, g, st = make_mr_pair()
    o.set_substep(1e-3, ns); g.set_substep(1e-3, ns)
    o.call("explicit_salinity"); g.call("explicit_salinity")
    cmp(o, g, "hsal", np.s_[..., :MR["nxmr"]])
    o.call("implicit_salinity"); g.call("implicit_salinity")
    cmp(o, g, "sal", np.s_[..., :MR["nxmr"]])
    cmp(o, g, "rusal", np.s_[..., :MR["nxmr"]])


def test_add_sal_buoyancy():
    cfg = dict(MR, active_s=1)
    o, g, st = make_mr_pair(cfg)
    put(o, g, "qcap", uniform(31, st["vx"].shape))
    o.set_substep(1e-3, 1); g.set_substep(1e-3, 1)
    o.call("add_sal_buoyancy"); g.call("add_sal_buoyancy")
    cmp(o, g, "qcap", np.s_[..., 1:16])


@pytest.mark.parametrize("active_s,base", [(0, "MR"), (1, "MR"), (0, "MR_ODD"), (1, "MR_BIG")])
def test_time_marcher_multires(active_s, base):
    cfg = dict({"MR": MR, "MR_ODD": MR_ODD, "MR_BIG": MR_BIG}[base], active_s=active_s)
    o, g, _ = make_mr_pair(cfg)
    for _ in range(2):
        o.time_marcher(1e-3)
        g.time_marcher(1e-3)
    for f in ("vx", "vy", "vz", "temp", "pr", "salc"):
        cmp(o, g, f, np.s_[..., :cfg["nxm"]] if f != "vx" else np.s_[...], tol=1e-10)
    for f in ("sal", "vxr", "vyr", "vzr"):
        cmp(o, g, f, tol=1e-10)
    a = np.array(g.calc_max_cfl()); b = np.array(o.calc_max_cfl())
    assert relerr(a, b) < 1e-12
    a = np.array(g.check_divergence()); b = np.array(o.check_divergence())
    assert np.all(a < 1e-9) and np.all(b < 1e-9)


@pytest.mark.parametrize("istr3,str3", [(0, 0.0), (1, 2.0), (2, 0.0), (3, 0.0), (4, 1.4), (6, 3.0), (7, 2.0), (8, 0.0)])
def test_setup_arrays(istr3, str3):
    cfg = dict(BASE, nxm=20, istr3=istr3, str3=str3, tfixs=1, tfixn=0, inslws=0, inslwn=1)
    o, g, _ = make_pair(cfg)
    for name in ("xc", "xm", "g3rc", "udx3c", "udx3m", "ap3ck", "ac3ck", "am3ck", "ap3sk", "ac3sk", "am3sk",
                 "ap3ssk", "ac3ssk", "am3ssk", "ak1", "ak2", "amphk", "acphk", "apphk"):
        a, b = g.grid_array(name), o.grid(name)
        n = min(len(a), len(b))
        assert relerr(a[:n], b[:n]) < 1e-14, name


@pytest.mark.parametrize("cfg", [dict(BASE, nxm=12, nym=8, nzm=6, istr3=4, str3=1.4),
                                 dict(BASE, nxm=12, nym=8, nzm=6),
                                 dict(BASE, nym=8, nzm=6, istr3=4, str3=1.4)])
def test_time_marcher_shapes(cfg):
    o, g, _ = make_pair(cfg)
    for _ in range(2):
        o.time_marcher(2e-3)
        g.time_marcher(2e-3)
    nxm = cfg["nxm"]
    for f in ("vx", "vy", "vz", "temp", "pr"):
        cmp(o, g, f, np.s_[..., :nxm] if f != "vx" else np.s_[...], tol=1e-10)


# refined grid that does not bracket the base one (uniform refined grid over
# a wall-clustered base grid): the general refined -> coarse path
# (HermiteInterpolations.F90:218-250, salinity.F90:485-489)
MR_GEN = dict(MR, nxm=24, nym=8, nzm=12, nxmr=40, nymr=16, nzmr=24, istr3=8, istr3r=0)


def test_general_path_grid():
    o = pyoracle.Oracle(MR_GEN)
    assert o.grid("xmr")[0] >= o.grid("xm")[0]
    o.close()


@pytest.mark.parametrize("gaxis", [1, 2, 3])
@pytest.mark.parametrize("base", ["MR", "MR_GEN"])
def test_interp_sal_multigrid_paths(base, gaxis):
    """Both refined -> coarse paths for every gravity axis: the fast path puts
    salc on the faces along gAxis (cxsalc / cysalc / czsalc), the general path
    on the cell centres."""
    cfg = dict({"MR": MR, "MR_GEN": MR_GEN}[base], gaxis=gaxis)
    o, g, _ = make_mr_pair(cfg)
    o.call("interp_sal_multigrid"); g.call("interp_sal_multigrid")
    cmp(o, g, "salc", np.s_[..., :cfg["nxm"]], tol=1e-13)


@pytest.mark.parametrize("gaxis", [1, 2])
def test_time_marcher_general_path(gaxis):
    cfg = dict(MR_GEN, active_s=1, gaxis=gaxis)
    o, g, _ = make_mr_pair(cfg)
    for _ in range(2):
        o.time_marcher(1e-3)
        g.time_marcher(1e-3)
    for f in ("vx", "vy", "vz", "temp", "pr", "salc"):
        cmp(o, g, f, np.s_[..., :cfg["nxm"]] if f != "vx" else np.s_[...], tol=1e-10)
    for f in ("sal", "vxr", "vyr", "vzr"):
        cmp(o, g, f, tol=1e-10)
