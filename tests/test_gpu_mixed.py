"""GPU tests of the mixed FP32/FP64 form against the FP64 GPU form and the
FP64 CPU oracle.

The mixed form keeps u, T, p in FP64 and the wall-normal solves in FP64; the
explicit terms, the refined scalar and velocities, the Hermite transfers and
the Poisson transforms run in FP32, the Poisson solve inside FP64-residual
iterative refinement to max|div u| <= div_tol.  Bounds, per check:
  - max|div u| on the base grid <= div_tol (1e-10 unless stated);
  - Nu_vol and KE relative error <= 1e-6 against the FP64 form;
  - relative L2 error of the state after the run: base fields <= 1e-5,
    refined scalar <= 1e-5 (FP32 storage, rounding 6e-8 per update).
"""
import numpy as np
import pytest

import pyafid
import pyoracle
from helpers import load_state, random_state

pytestmark = pytest.mark.gpu

MR = dict(nxm=32, nym=24, nzm=24, istr3=8, multires=1, salinity=1, nxmr=64, nymr=48, nzmr=48, istr3r=8,
          rayt=1e6, prat=1.0, rays=-1e6, pras=4.0, ylen=1.0, zlen=1.0)
RBC = dict(nxm=32, nym=32, nzm=32, istr3=8, rayt=1e6, prat=1.0)
BASE_FIELDS = ["vx", "vy", "vz", "temp", "pr"]


def l2rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def interior(cfg, name, a):
    """Cells the step defines: x 1..nxm (vx: 2..nxm), pr minus its mean."""
    if name in ("sal",):
        return a[..., :cfg["nxmr"]]
    if name == "vx":
        return a[..., 1:cfg["nxm"]]
    a = a[..., :cfg["nxm"]]
    return a - a.mean() if name == "pr" else a


def start(cfg, precision, seed=5, eps=0.05, **extra):
    g = pyafid.Solver(dict(cfg, precision=precision, **extra))
    nz, ny = cfg["nzm"], cfg["nym"]
    g.upload("tempbp", np.ones((nz, ny, 1)))
    g.upload("temptp", np.zeros((nz, ny, 1)))
    if cfg.get("salinity"):
        g.upload("salbp", np.ones((cfg["nzmr"], cfg["nymr"], 1)))
        g.upload("saltp", np.zeros((cfg["nzmr"], cfg["nymr"], 1)))
    g.create_initial_conditions(seed, eps)
    return g


def run(g, steps, dt):
    for _ in range(steps):
        g.time_marcher(dt)
    g.synchronize()


@pytest.mark.parametrize("cfg,steps,dt", [(MR, 12, 2e-3), (RBC, 12, 2e-3)], ids=["multires", "rbc"])
def test_mixed_vs_fp64(cfg, steps, dt):
    a = start(cfg, pyafid.FP64)
    b = start(cfg, pyafid.MIXED)
    run(a, steps, dt)
    run(b, steps, dt)
    div_a, divr_a = a.check_divergence()
    div_b, divr_b = b.check_divergence()
    assert div_a < 1e-12
    assert div_b <= 1e-10, f"mixed max|div u| {div_b:.3e}"
    qa, qb = a.global_quantities(), b.global_quantities()
    print(f"max|div|: fp64 {div_a:.3e} mixed {div_b:.3e}; residuals {b.refinement_residuals()}")
    assert qa["ke"] > 1e-8                            # the flow has developed
    for k in ("nu", "ke"):
        e = abs(qb[k] - qa[k]) / abs(qa[k])
        print(f"{k}: fp64 {qa[k]:.12e} mixed {qb[k]:.12e} rel {e:.2e}")
        assert e <= 1e-6, f"{k}: FP64 {qa[k]:.12e} mixed {qb[k]:.12e} rel {e:.2e}"
    for f in BASE_FIELDS:
        e = l2rel(interior(cfg, f, b.download(f)), interior(cfg, f, a.download(f)))
        print(f"{f}: relative L2 {e:.3e}")
        assert e <= 1e-5, f"{f}: relative L2 {e:.3e}"
    if cfg.get("salinity"):
        e = l2rel(interior(cfg, "sal", b.download("sal")), interior(cfg, "sal", a.download("sal")))
        print(f"sal: relative L2 {e:.3e}")
        assert e <= 1e-5, f"sal: relative L2 {e:.3e}"
        # refined divergence: FP32 storage of the refined velocity bounds it by
        # the FP32 unit round-off (6e-8) of max|u_r| over the smallest refined
        # spacing, times the few terms of the divergence and of the
        # x-integration (constant 100 here)
        umax = max(np.abs(a.download(n)).max() for n in ("vxr", "vyr", "vzr"))
        invh = max(a.grid_array("udx3mr").max(), cfg["nymr"] / cfg["ylen"], cfg["nzmr"] / cfg["zlen"])
        bound = 100 * 6e-8 * umax * invh
        print(f"refined max|div|: fp64 {divr_a:.3e} mixed {divr_b:.3e} (bound {bound:.3e})")
        assert divr_b <= bound, f"refined div {divr_b:.3e} > {bound:.3e}"
    a.close()
    b.close()


def test_mixed_vs_oracle():
    """Mixed GPU step against the FP64 CPU oracle from the same state."""
    cfg = MR
    g = start(cfg, pyafid.MIXED, seed=9)
    st = {n: g.download(n) for n in ["vx", "vy", "vz", "temp", "pr", "tempbp", "temptp", "sal", "salbp", "saltp"]}
    o = pyoracle.Oracle(cfg)
    load_state(o, st, upload=False)
    o.call("interp_sal_multigrid")
    for _ in range(6):
        g.time_marcher(2e-3)
        o.time_marcher(2e-3)
    g.synchronize()
    for f in BASE_FIELDS:
        e = l2rel(interior(cfg, f, g.download(f)), interior(cfg, f, o.field(f)))
        assert e <= 1e-5, f"{f}: relative L2 {e:.3e}"
    e = l2rel(interior(cfg, "sal", g.download("sal")), interior(cfg, "sal", o.field("sal")))
    assert e <= 1e-5, f"sal: relative L2 {e:.3e}"
    nu_o = o.global_quantities()["nu"]
    nu_g = g.global_quantities()["nu"]
    assert abs(nu_g - nu_o) / abs(nu_o) <= 1e-6
    o.close()
    g.close()


def test_refinement_contracts():
    """On a rough random state each refinement sweep cuts the FP64 residual
    divergence by a large factor, and the gates skip the sweeps that are
    not needed."""
    cfg = dict(MR, salinity=0, multires=0, nxm=32, nym=32, nzm=32)
    st = random_state(cfg, 3)
    g = pyafid.Solver(dict(cfg, precision=pyafid.MIXED, div_tol=1e-300, max_sweeps=4))
    load_state(g, st, upload=True)
    g.time_marcher(1e-3)
    r = g.refinement_residuals()
    assert len(r) == 3 and r[0] > 0.0
    assert r[1] < 1e-2 * r[0] and r[2] < 1e-2 * r[1], r
    d, _ = g.check_divergence()
    assert d < 1e-2 * r[2] + 1e-13, (d, r)
    g.close()
    # a loose tolerance stops after the first check
    g = pyafid.Solver(dict(cfg, precision=pyafid.MIXED, div_tol=1e30, max_sweeps=3))
    load_state(g, st, upload=True)
    g.time_marcher(1e-3)
    r = g.refinement_residuals()
    assert r[0] > 0.0 and r[1] == 0.0, r
    g.close()


def test_mixed_per_routine_calls_refused():
    g = pyafid.Solver(dict(RBC, precision=pyafid.MIXED))
    g.set_substep(1e-3, 1)
    for r in ("explicit_terms_vx", "implicit_and_update_temp", "solve_pressure_correction", "correct_pressure"):
        with pytest.raises(pyafid.AfidError):
            g.call(r)
    with pytest.raises(pyafid.AfidError):
        g.download("qcap")
    g.close()


def test_fp32_field_roundtrip():
    g = pyafid.Solver(dict(MR, precision=pyafid.MIXED))
    st = random_state(MR, 2)
    for name in ("sal", "rusal", "rux", "rutemp"):
        g.upload(name, st[name])
        back = g.download(name)
        assert np.array_equal(back, st[name].astype(np.float32).astype(np.float64)), name
    for name in ("vx", "temp", "pr"):       # FP64 state is exact
        g.upload(name, st[name])
        assert np.array_equal(g.download(name), st[name]), name
    g.close()
