"""Checkpoint and restart on the GPU (WriteFlowField.F90:11-86,
ReadFlowInterp.F90:143-171 and the start-up transfers of main.F90:205-213).

A run of 4 RK3 steps and a run of 2 steps, a checkpoint, a fresh context
that reads it and 2 more steps must end in the same bits: every array the
step carries over is in the checkpoint (the RK3 history arrays are not, as
the first substep's rho = 0 does not read them) or is rebuilt by the start-up
transfers.  Both forms; the mixed form stores S / phi in FP32, which the FP64
files hold exactly.  The means.h5 writer is checked against
afid_calc_mean_profiles on the same state.
"""
import numpy as np
import pytest

import pyafid

pytestmark = pytest.mark.gpu

MR = dict(nxm=16, nym=12, nzm=12, istr3=8, multires=1, salinity=1, nxmr=32, nymr=24, nzmr=24, istr3r=8,
          rayt=1e6, prat=1.0, rays=-1e6, pras=4.0)
PF = dict(nxm=16, nym=12, nzm=12, istr3=8, multires=1, phasefield=1, nxmr=32, nymr=24, nzmr=24, istr3r=8,
          rayt=1e6, prat=1.0, pf_a=1.0, pf_s=1.0, pf_tm=0.0)
RBC = dict(nxm=24, nym=16, nzm=16, istr3=8, rayt=1e6, prat=1.0)
CASES = {"rbc": RBC, "multires": MR, "phasefield": PF}
DT = 2e-3


def state(g):
    c = g.cfgd
    names = ["vx", "vy", "vz", "temp", "pr"] + (["sal"] if c["salinity"] else []) + (["phi"] if c["phasefield"] else [])
    return {n: g.download(n) for n in names}


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_restart_is_exact(tmp_path, case, form):
    cfg = dict(CASES[case], precision=pyafid.MIXED if form == "mixed" else pyafid.FP64)
    a = pyafid.Solver(cfg)
    a.create_initial_conditions(7, eps=1e-2)
    for _ in range(2):
        a.time_marcher(DT)
    a.write_flow_field(tmp_path, time=2 * DT, final=True)
    saved = state(a)
    for _ in range(2):
        a.time_marcher(DT)
    ref = state(a)
    a.close()

    b = pyafid.Solver(cfg)
    t = b.read_flow_field(tmp_path)
    assert t == 2 * DT
    back = state(b)
    for n, v in saved.items():
        assert np.array_equal(back[n], v), f"{n} after reading the checkpoint"
    for _ in range(2):
        b.time_marcher(DT)
    got = state(b)
    b.close()
    for n, v in ref.items():
        assert np.isfinite(v).all(), n
        assert np.array_equal(got[n], v), f"{n}: max diff {np.abs(got[n] - v).max():.3e}"


def test_restart_on_another_grid_interpolates(tmp_path):
    """A checkpoint of another refined grid is read through the restart
    transfer (ReadFlowInterp.F90:90-141, tests/test_gpu_restart.py checks its
    values): S lands on the new grid and the base fields, whose grid did not
    change, come back unchanged.  Another domain length is refused
    (ReadFlowInterp.F90:76-82)."""
    a = pyafid.Solver(MR)
    a.create_initial_conditions(3)
    a.write_flow_field(tmp_path, time=0.0, final=True)
    base = {n: a.download(n) for n in ("temp", "vy")}
    a.close()
    b = pyafid.Solver(dict(MR, nxmr=48))
    b.read_flow_field(tmp_path)
    s = b.download("sal")
    assert s.shape[-1] == 49 and np.isfinite(s).all()
    assert np.abs(s[..., :48]).max() > 0.0
    nxm = MR["nxm"]
    for n, v in base.items():
        assert np.allclose(b.download(n)[..., :nxm], v[..., :nxm], rtol=0, atol=1e-12), n
    b.close()
    c = pyafid.Solver(dict(MR, ylen=2.0 * MR.get("ylen", 1.0)))
    with pytest.raises(pyafid.AfidError, match="domain size"):
        c.read_flow_field(tmp_path)
    c.close()


def test_means_file_matches_profiles(tmp_path):
    g = pyafid.Solver(MR)
    g.create_initial_conditions(5, eps=1e-2)
    g.time_marcher(DT)
    prof = g.write_means(tmp_path, time=0.5, tout=0.25)
    again = g.mean_profiles()
    again.update(g.sal_stats())
    f = tmp_path / "means.h5"
    for name in pyafid.MEAN_PROFILES + pyafid.SAL_PROFILES:
        n = MR["nxm"] if name in pyafid.MEAN_PROFILES else MR["nxmr"]
        v = pyafid.h5_read_1d(f, f"{name}/00002", n)
        assert np.array_equal(v, prof[name]), name
        assert np.array_equal(v, again[name]), name
    g.close()
