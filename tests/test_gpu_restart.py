"""GPU parity of the restart onto a new grid (ReadFlowInterp.F90:90-141,
IC_interpolation/InterpInputVel.F90, InterpInputSal.F90, InterpInputPhi.F90)
against the oracle, on the same old fields: finer, coarser and differently
clustered base grids, the refined S and phi grids, and the whole file path
(a checkpoint written on one grid, read by a context on another, stepped).

Tolerance: the same Hermite weights and summation order in FP64, fused
multiply-adds on the GPU: 1e-12 relative to each field's max norm; the
integrated velocity components accumulate along a line, 1e-11.
"""
import os

import numpy as np
import pytest

import pyafid
import pyoracle
from helpers import relerr, uniform

pytestmark = pytest.mark.gpu

OLD = dict(nxm=16, nym=8, nzm=12, istr3=4, str3=1.5, rayt=1e5, prat=1.0, ylen=1.0, zlen=1.5)
NEW = {
    "finer": dict(OLD, nxm=24, nym=12, nzm=18, istr3=8),
    "coarser": dict(OLD, nxm=12, nym=6, nzm=9, istr3=0),
    "cheb": dict(OLD, nxm=20, nym=16, nzm=12, istr3=6, str3=0.0),
    "same": OLD,
}


def old_fields(cfg, seed):
    nx, ny, nz = cfg["nxm"] + 1, cfg["nym"], cfg["nzm"]
    f = {n: uniform(seed * 10 + q, (nz, ny, nx)) - 0.5 for q, n in enumerate(("vx", "vy", "vz", "temp"))}
    f["vx"][..., 0] = 0.0
    f["vx"][..., -1] = 0.0
    return f


def og_dict(cfg):
    d = dict(nx=cfg["nxm"] + 1, ny=cfg["nym"] + 1, nz=cfg["nzm"] + 1, istr3=cfg.get("istr3", 0),
             str3=cfg.get("str3", 0.0), ylen=cfg["ylen"], zlen=cfg["zlen"])
    if cfg.get("multires"):
        d.update(nxr=cfg["nxmr"] + 1, nyr=cfg["nymr"] + 1, nzr=cfg["nzmr"] + 1, istr3r=cfg.get("istr3r", 0))
    return d


@pytest.mark.parametrize("case", sorted(NEW))
def test_interp_input_vel(case):
    new = NEW[case]
    f = old_fields(OLD, 3)
    g = pyafid.Solver(new)
    o = pyoracle.Oracle(new)
    g.interp_input(og_dict(OLD), f)
    o.interp_input_vel(pyoracle.old_grid(OLD), f["vx"], f["vy"], f["vz"], f["temp"])
    nxm = new["nxm"]
    for n in ("vx", "vy", "vz", "temp"):
        sl = np.s_[...] if n == "vx" else np.s_[..., :nxm]
        e = relerr(g.download(n)[sl], o.field(n)[sl])
        assert e <= (1e-12 if n == "temp" else 1e-11), f"{case} {n}: {e:.3e}"
    for n in ("pr", "rux", "ruy", "ruz", "rutemp"):
        assert not np.any(g.download(n)), n
    g.close(); o.close()


OLD_MR = dict(OLD, multires=1, salinity=1, nxmr=24, nymr=16, nzmr=24, istr3r=4, rays=-1e5, pras=4.0)
NEW_MR = dict(OLD_MR, nxm=20, nym=12, nzm=12, nxmr=36, nymr=24, nzmr=24, istr3=8, istr3r=8)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
def test_interp_input_refined(form):
    """S and phi onto the new refined grids (FP32 storage in the mixed form)."""
    nr = (OLD_MR["nzmr"], OLD_MR["nymr"], OLD_MR["nxmr"] + 1)
    S = uniform(17, nr) - 0.5
    prec = pyafid.MIXED if form == "mixed" else pyafid.FP64
    tol = 1e-7 if form == "mixed" else 1e-12          # FP32 storage rounds the result
    g = pyafid.Solver(dict(NEW_MR, precision=prec))
    o = pyoracle.Oracle(NEW_MR)
    f = old_fields(OLD_MR, 5)
    g.interp_input(og_dict(OLD_MR), dict(f, sal=S))
    o.interp_input_ref(pyoracle.old_grid(OLD_MR), "sal", S)
    n = NEW_MR["nxmr"]
    assert relerr(g.download("sal")[..., :n], o.field("sal")[..., :n]) <= tol
    g.close(); o.close()
    pf_old = dict(OLD_MR, salinity=0, phasefield=1, pf_a=1.0, pf_s=1.0, pf_tm=0.0)
    pf_new = dict(NEW_MR, salinity=0, phasefield=1, pf_a=1.0, pf_s=1.0, pf_tm=0.0)
    P = uniform(19, nr)
    g = pyafid.Solver(dict(pf_new, precision=prec))
    o = pyoracle.Oracle(pf_new)
    g.interp_input(og_dict(pf_old), dict(f, phi=P))
    o.interp_input_ref(pyoracle.old_grid(pf_old), "phi", P)
    assert relerr(g.download("phi")[..., :n], o.field("phi")[..., :n]) <= tol
    g.close(); o.close()


def test_old_grid_errors():
    g = pyafid.Solver(NEW["finer"])
    f = old_fields(OLD, 3)
    with pytest.raises(pyafid.AfidError, match="domain size"):
        g.interp_input(dict(og_dict(OLD), ylen=2.0), f)
    with pytest.raises(pyafid.AfidError, match="salinity is not enabled"):
        g.interp_input(og_dict(OLD_MR), dict(f, sal=np.zeros((24, 16, 25))))
    g.close()


def test_checkpoint_onto_new_grid_and_step(tmp_path):
    """A multires run on the old grids writes its checkpoint; a context on the
    new grids reads it (ReadFlowInterp with the interpolation, then the
    start-up transfers of main.F90:205-213) and steps: every field matches the
    oracle doing the same from the same files."""
    a = pyafid.Solver(OLD_MR)
    nz, ny = OLD_MR["nzm"], OLD_MR["nym"]
    a.upload("tempbp", np.full((nz, ny, 1), 0.5)); a.upload("temptp", np.full((nz, ny, 1), -0.5))
    a.upload("salbp", np.full((OLD_MR["nzmr"], OLD_MR["nymr"], 1), 0.5))
    a.upload("saltp", np.full((OLD_MR["nzmr"], OLD_MR["nymr"], 1), -0.5))
    a.create_initial_conditions(7, 0.05)
    for _ in range(3):
        a.time_marcher(2e-3)
    a.write_flow_field(str(tmp_path), 0.006, final=True)
    a.close()
    b = pyafid.Solver(NEW_MR)
    nz, ny = NEW_MR["nzm"], NEW_MR["nym"]
    b.upload("tempbp", np.full((nz, ny, 1), 0.5)); b.upload("temptp", np.full((nz, ny, 1), -0.5))
    b.upload("salbp", np.full((NEW_MR["nzmr"], NEW_MR["nymr"], 1), 0.5))
    b.upload("saltp", np.full((NEW_MR["nzmr"], NEW_MR["nymr"], 1), -0.5))
    t = b.read_flow_field(str(tmp_path))
    assert abs(t - 0.006) < 1e-15
    o = pyoracle.Oracle(NEW_MR)
    o.field("tempbp")[...] = 0.5; o.field("temptp")[...] = -0.5
    o.field("salbp")[...] = 0.5; o.field("saltp")[...] = -0.5
    old = {n: pyafid.h5_read_field(os.path.join(tmp_path, f"continua_{n}.h5"),
                                   (OLD_MR["nzm"], OLD_MR["nym"], OLD_MR["nxm"] + 1)) for n in ("vx", "vy", "vz", "temp")}
    S = pyafid.h5_read_field(os.path.join(tmp_path, "continua_sal.h5"), (OLD_MR["nzmr"], OLD_MR["nymr"], OLD_MR["nxmr"] + 1))
    o.interp_input_vel(pyoracle.old_grid(OLD_MR), old["vx"], old["vy"], old["vz"], old["temp"])
    o.interp_input_ref(pyoracle.old_grid(OLD_MR), "sal", S)
    o.call("interp_vel_mgrd"); o.call("interp_sal_multigrid")
    for _ in range(2):
        b.time_marcher(2e-3)
        o.time_marcher(2e-3)
    nxm, nxmr = NEW_MR["nxm"], NEW_MR["nxmr"]
    for n in ("vx", "vy", "vz", "temp", "sal"):
        sl = np.s_[...] if n == "vx" else np.s_[..., :(nxmr if n == "sal" else nxm)]
        e = relerr(b.download(n)[sl], o.field(n)[sl])
        assert e <= 1e-10, f"{n}: {e:.3e}"
    pg, po = b.download("pr")[..., :nxm], o.field("pr")[..., :nxm]
    assert relerr(pg - pg.mean(), po - po.mean()) <= 1e-10
    b.close(); o.close()
