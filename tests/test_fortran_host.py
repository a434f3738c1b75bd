"""The Fortran host (driver loop of main.F90:281-422 and the TimeMarcher call
tree, bound with iso_c_binding) against the FP64 CPU oracle driven the same
way: same bou.in, same seeded initial state, same number of steps."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle
from helpers import uniform

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "afid-murphfi_amd", "afid_host")
EXAMPLE = os.path.join(ROOT, "afid-murphfi_amd", "fortran", "examples", "rbc_64", "bou.in")


def write_bou(path, nxm, nym, nzm, ntst, dt):
    txt = open(EXAMPLE).read()
    txt = txt.replace("64     64     64     3", f"{nxm}     {nym}     {nzm}     3", 1)
    txt = txt.replace("199     7000", f"{ntst}     7000", 1)
    txt = txt.replace("0       1e-3    1.e-3", f"0       {dt}    1.e-3", 1)
    open(path, "w").write(txt)


def run_host(tmp_path, mode):
    out = subprocess.run([EXE] + (["fused"] if mode == "fused" else []) + ["seed=7"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = dict(re.findall(r"(\w+)=\s*([-+0-9.eE]+)", out.stdout))
    return {k: float(v) for k, v in vals.items()}


def oracle_run(nxm, nym, nzm, nsteps, dt, seed=7):
    o = pyoracle.Oracle(dict(nxm=nxm, nym=nym, nzm=nzm, istr3=8, rayt=1e5, prat=1.0))
    # SetTempBCs.F90:27-33 (RBC, inslwN = 1): T = +0.5 bottom, -0.5 top
    o.field("tempbp")[...] = 0.5
    o.field("temptp")[...] = -0.5
    xm = o.grid("xm")[:nxm]
    eps = 1e-3
    amp = np.where(np.abs(xm - 0.5) + eps > 0.5, 0.5 - np.abs(xm - 0.5), eps)
    T = o.field("temp")
    T[..., :nxm] = 0.5 - xm + amp * (2 * uniform(seed, T.shape)[..., :nxm] - 1)
    for _ in range(nsteps):
        o.time_marcher(dt)
    d, _ = o.check_divergence()
    return o.global_quantities()["nu"], d


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")
@pytest.mark.parametrize("mode", ["calltree", "fused"])
def test_fortran_host_matches_oracle(tmp_path, mode):
    nxm, nym, nzm, ntst, dt = 32, 24, 16, 19, 2e-3
    write_bou(os.path.join(tmp_path, "bou.in"), nxm, nym, nzm, ntst, dt)
    r = run_host(tmp_path, mode)
    assert int(r["steps"]) == ntst + 1 and int(r["errorcode"]) == 555      # normal end: ntst reached
    nu, d = oracle_run(nxm, nym, nzm, ntst + 1, dt)
    assert abs(r["nu"] - nu) <= 1e-10 * abs(nu)
    assert r["maxdiv"] < 1e-12 and d < 1e-12


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")
def test_fortran_host_time_limit(tmp_path):
    """A step that carries time past TMAX ends the run with errorcode 333
    (main.F90:361), after one step, as the reference's loop does."""
    write_bou(os.path.join(tmp_path, "bou.in"), 16, 8, 8, 5, 1e6)
    r = run_host(tmp_path, "fused")
    assert int(r["errorcode"]) == 333 and int(r["steps"]) == 1


def with_salinity(path):
    """Turn the bou.in at `path` into a multires case with salinity on a 2x refined grid."""
    txt = open(path).read()
    nxm, nym, nzm = (int(v) for v in txt.split("NSST(>1=RK,else AB)\n", 1)[1].split()[:3])
    old = "MULTIRES    NXMR    NYMR    NZMR\n0           64     64     64"
    assert txt.count(old) == 1 and txt.count("FLAGSAL     FLAGPF\n0           0") == 1
    txt = txt.replace(old, f"MULTIRES    NXMR    NYMR    NZMR\n1           {2 * nxm}     {2 * nym}     {2 * nzm}")
    txt = txt.replace("FLAGSAL     FLAGPF\n0           0", "FLAGSAL     FLAGPF\n1           0")
    open(path, "w").write(txt)


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")
@pytest.mark.parametrize("case", ["rbc", "multires"])
def test_fortran_host_checkpoint_and_restart(tmp_path, case):
    """NREAD = 1 restarts from outputdir/continua_*.h5 (ReadFlowInterp.F90:143-171):
    5 steps, the checkpoint a normal end writes (QuitRoutine.F90:34-38), then
    5 more steps from it give the bits of 10 uninterrupted steps.  The run also
    writes means.h5 every TOUT and the start-up snapshot of main.F90:241.  The
    multires case restarts S on the refined grid and reruns the start-up
    transfers (main.F90:205-209)."""
    import pyafid
    nxm, nym, nzm, dt = (24, 16, 16, 2e-3) if case == "rbc" else (16, 12, 12, 2e-3)
    one, two = tmp_path / "one", tmp_path / "two"
    os.makedirs(one)
    os.makedirs(two)
    write_bou(os.path.join(one, "bou.in"), nxm, nym, nzm, 9, dt)
    if case == "multires":
        with_salinity(os.path.join(one, "bou.in"))
    r10 = run_host(one, "fused")
    assert int(r10["steps"]) == 10 and int(r10["errorcode"]) == 555
    write_bou(os.path.join(two, "bou.in"), nxm, nym, nzm, 4, dt)
    if case == "multires":
        with_salinity(os.path.join(two, "bou.in"))
    r5 = run_host(two, "fused")
    assert int(r5["steps"]) == 5
    out = two / "outputdir"
    assert pyafid.h5_dims(out / "continua_vx.h5") == (nxm + 1, nym, nzm)
    assert pyafid.h5_read_int(out / "continua_master.h5", "nx") == nxm + 1
    assert abs(pyafid.h5_read_real(out / "continua_master.h5", "time") - 5 * dt) < 1e-15
    assert pyafid.h5_dims(out / "fields" / "00000_temp.h5") == (nxm + 1, nym, nzm)
    assert pyafid.h5_dims(out / "means.h5", "Tbar/00000") == (nxm,)
    txt = open(os.path.join(two, "bou.in")).read()
    assert txt.count("NREAD       IRESET\n0           0") == 1
    open(os.path.join(two, "bou.in"), "w").write(txt.replace("NREAD       IRESET\n0           0",
                                                             "NREAD       IRESET\n1           0"))
    r = run_host(two, "fused")
    assert int(r["steps"]) == 5
    assert abs(r["time"] - 10 * dt) < 1e-14
    assert r["nu"] == r10["nu"] and r["ke"] == r10["ke"] and r["maxdiv"] == r10["maxdiv"]
    a = pyafid.h5_read_field(one / "outputdir" / "continua_temp.h5", (nzm, nym, nxm + 1))
    b = pyafid.h5_read_field(out / "continua_temp.h5", (nzm, nym, nxm + 1))
    assert np.array_equal(a, b)
    if case == "multires":
        shape = (2 * nzm, 2 * nym, 2 * nxm + 1)
        a = pyafid.h5_read_field(one / "outputdir" / "continua_sal.h5", shape)
        b = pyafid.h5_read_field(out / "continua_sal.h5", shape)
        assert np.isfinite(a).all() and np.array_equal(a, b)
        assert pyafid.h5_read_int(out / "continua_master.h5", "nxr") == 2 * nxm + 1
        assert pyafid.h5_dims(out / "means.h5", "Sbar/00000") == (2 * nxm,)
