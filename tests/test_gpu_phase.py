"""GPU parity of the phase-field routines (src/phasefield.F90) against the
FP64 CPU oracle on the same seeded inputs.

Tolerance: FP64 in both; kernels reorder sums and fuse multiply-adds, so
agreement is to round-off relative to each field's max norm (1e-12; 1e-11
for whole steps, which include the Poisson solve).
"""
import numpy as np
import pytest

import pyafid
import pyoracle
from helpers import load_state, random_state, relerr, uniform

pytestmark = pytest.mark.gpu

PF = dict(nxm=16, nym=12, nzm=12, istr3=8, multires=1, phasefield=1, nxmr=32, nymr=24, nzmr=24, istr3r=8,
          rayt=1e6, prat=1.0, pf_a=1.0, pf_s=1.0, pf_tm=0.0)
PF_ODD = dict(PF, nxm=12, nym=8, nzm=6, nxmr=30, nymr=20, nzmr=18, istr3=4, str3=1.4, istr3r=6, pf_a=0.7, pf_s=1.3,
              pf_tm=0.1, pf_eps=0.05)
PF_BIG = dict(PF, nxm=40, nym=24, nzm=36, nxmr=80, nymr=72, nzmr=72)
# uniform refined grid over a wall-clustered base grid: the refined cell
# centres do not bracket the coarse ones, so phic and the latent heat take
# the general path (phasefield.F90:489-492, :701-723)
PF_GEN = dict(PF, nxm=24, nxmr=40, istr3=8, istr3r=0)
TOL = 1e-12


def pf_state(cfg, seed=4):
    """Random base state plus a phase field in [0, 1] whose coarse image has
    solid (> 0.9), liquid and interface rows."""
    st = random_state(cfg, seed)
    for f in ("rux", "ruy", "ruz", "rutemp"):
        st[f] *= 0.1
    nxr, nyr, nzr = cfg["nxmr"] + 1, cfg["nymr"], cfg["nzmr"]
    nx, ny, nz = cfg["nxm"] + 1, cfg["nym"], cfg["nzm"]
    st["phi"] = uniform(seed * 10 + 1, (nzr, nyr, nxr))
    st["ruphi"] = uniform(seed * 10 + 2, (nzr, nyr, nxr)) - 0.5
    st["tempr"] = uniform(seed * 10 + 3, (nzr, nyr, nxr))
    phic = uniform(seed * 10 + 4, (nz, ny, nx))
    phic[phic > 0.6] = 1.0           # solid patches
    phic[..., nx - 1] = 0.0
    st["phic"] = phic
    return st


def make_pair(cfg, seed=4):
    st = pf_state(cfg, seed)
    o = pyoracle.Oracle(cfg)
    g = pyafid.Solver(cfg)
    load_state(o, st, upload=False)
    load_state(g, st, upload=True)
    return o, g, st


def cmp(o, g, name, sl=np.s_[...], tol=TOL):
    e = relerr(g.download(name)[sl], o.field(name)[sl])
    assert e <= tol, f"{name}: relative error {e:.3e} > {tol:.1e}"
    return e


@pytest.mark.parametrize("cfg", [PF, PF_ODD], ids=["pf", "pf_odd"])
def test_explicit_phase(cfg):
    o, g, _ = make_pair(cfg)
    for ns in (1, 2):
        o.set_substep(2e-3, ns); g.set_substep(2e-3, ns)
        o.call("explicit_phase"); g.call("explicit_phase")
        cmp(o, g, "hphi", np.s_[..., : cfg["nxmr"]])


@pytest.mark.parametrize("cfg", [PF, PF_ODD], ids=["pf", "pf_odd"])
@pytest.mark.parametrize("ns", [1, 2, 3])
def test_implicit_phase(cfg, ns):
    o, g, st = make_pair(cfg)
    h = uniform(91, st["phi"].shape) - 0.5
    o.field("hphi")[...] = h; g.upload("hphi", h)
    o.set_substep(2e-3, ns); g.set_substep(2e-3, ns)
    o.call("implicit_phase"); g.call("implicit_phase")
    n = cfg["nxmr"]
    for f in ("phi", "ruphi", "rhsr"):
        cmp(o, g, f, np.s_[..., :n])


def test_volume_penalty():
    o, g, st = make_pair(PF)
    for f in ("qcap", "dph", "dq"):
        a = uniform(50 + len(f), st["vx"].shape) - 0.5
        o.field(f)[...] = a; g.upload(f, a)
    o.set_substep(1e-3, 1); g.set_substep(1e-3, 1)
    o.call("add_volume_penalty"); g.call("add_volume_penalty")
    nxm = PF["nxm"]
    cmp(o, g, "qcap", np.s_[..., 1:nxm])
    cmp(o, g, "dph", np.s_[..., :nxm])
    cmp(o, g, "dq", np.s_[..., :nxm])


@pytest.mark.parametrize("cfg", [PF, PF_ODD, PF_BIG, PF_GEN], ids=["pf", "pf_odd", "pf_big", "pf_general"])
def test_latent_heat(cfg):
    o, g, st = make_pair(cfg)
    d = uniform(61, st["phi"].shape) - 0.5
    o.field("rhsr")[...] = d; g.upload("rhsr", d)
    hro = uniform(62, st["vx"].shape)
    o.field("hro")[...] = hro; g.upload("hro", hro)
    o.set_substep(2e-3, 2); g.set_substep(2e-3, 2)
    o.call("add_latent_heat"); g.call("add_latent_heat")
    cmp(o, g, "hro", np.s_[..., : cfg["nxm"]])


@pytest.mark.parametrize("cfg", [PF, PF_ODD, PF_BIG, PF_GEN], ids=["pf", "pf_odd", "pf_big", "pf_general"])
def test_interp_phi_and_temp(cfg):
    o, g, _ = make_pair(cfg)
    o.call("interp_phi_multigrid"); g.call("interp_phi_multigrid")
    cmp(o, g, "phic")
    o.call("interp_temp_multigrid"); g.call("interp_temp_multigrid")
    cmp(o, g, "tempr")


@pytest.mark.parametrize("routine,field,h", [("implicit_and_update_vx", "vx", "qcap"),
                                             ("implicit_and_update_vy", "vy", "dph"),
                                             ("implicit_and_update_vz", "vz", "dq")])
@pytest.mark.parametrize("cfg", [PF, PF_ODD], ids=["pf", "pf_odd"])
def test_masked_velocity_solves(routine, field, h, cfg):
    o, g, st = make_pair(cfg)
    a = uniform(70, st["vx"].shape) - 0.5
    o.field(h)[...] = a; g.upload(h, a)
    o.set_substep(2e-3, 2); g.set_substep(2e-3, 2)
    o.call(routine); g.call(routine)
    cmp(o, g, field)
    # rows under solid phic are forced to zero
    v = g.download(field)
    phic = st["phic"]
    nxm = cfg["nxm"]
    if field == "vy":
        solid = 0.5 * (phic[..., :nxm] + np.roll(phic, 1, axis=1)[..., :nxm]) > 0.9
        assert solid.any() and np.all(v[..., :nxm][solid] == 0.0)


@pytest.mark.parametrize("cfg", [PF, PF_ODD, PF_GEN], ids=["pf", "pf_odd", "pf_general"])
def test_time_marcher_phase(cfg):
    """The fused step (phi column, latent heat, masked base column, Poisson,
    transfers) against the oracle's TimeMarcher call tree."""
    o, g, st = make_pair(cfg)
    o.call("interp_temp_multigrid"); o.call("interp_phi_multigrid")
    g.call("interp_temp_multigrid"); g.call("interp_phi_multigrid")
    for _ in range(2):
        o.time_marcher(1e-3)
        g.time_marcher(1e-3)
    nxm, nxmr = cfg["nxm"], cfg["nxmr"]
    for f in ("vx", "vy", "vz", "temp"):
        cmp(o, g, f, np.s_[..., :nxm] if f != "vx" else np.s_[...], tol=1e-11)
    pr_o = o.field("pr")[..., :nxm]
    pr_g = g.download("pr")[..., :nxm]
    assert relerr(pr_g - pr_g.mean(), pr_o - pr_o.mean()) <= 1e-10
    for f in ("phi", "ruphi", "tempr"):
        cmp(o, g, f, np.s_[..., :nxmr], tol=1e-11)
    cmp(o, g, "phic", tol=1e-11)


def test_sphere_initial_condition():
    cfg = dict(PF, nxm=24, nym=16, nzm=16, nxmr=48, nymr=32, nzmr=32, pf_r0=0.2, pf_eps=0.03)
    g = pyafid.Solver(cfg)
    nz, ny = cfg["nzm"], cfg["nym"]
    g.upload("tempbp", np.ones((nz, ny, 1)))
    g.upload("temptp", np.zeros((nz, ny, 1)))
    g.create_initial_conditions(3, 1e-3)
    xmr = g.grid_array("xmr")[: cfg["nxmr"]]
    ymr = (np.arange(cfg["nymr"]) + 0.5) / cfg["nymr"]
    zmr = (np.arange(cfg["nzmr"]) + 0.5) / cfg["nzmr"]
    r = np.sqrt((xmr[None, None, :] - 0.5) ** 2 + (ymr[None, :, None] - 0.5) ** 2 + (zmr[:, None, None] - 0.5) ** 2)
    phi = g.download("phi")[..., : cfg["nxmr"]]
    assert np.abs(phi - 0.5 * (1 - np.tanh(0.5 * (r - 0.2) / 0.03))).max() < 1e-14
    xm = g.grid_array("xm")[: cfg["nxm"]]
    ym = (np.arange(ny) + 0.5) / ny
    zm = (np.arange(nz) + 0.5) / nz
    rb = np.sqrt((xm[None, None, :] - 0.5) ** 2 + (ym[None, :, None] - 0.5) ** 2 + (zm[:, None, None] - 0.5) ** 2)
    T = g.download("temp")[..., : cfg["nxm"]]
    assert np.all(T[rb <= 0.2] == 0.0) and np.all(T[rb > 0.25] > 0.0)
    # phic / tempr follow from phi / T as at start-up
    o = pyoracle.Oracle(cfg)
    for f in ("temp", "tempbp", "temptp", "phi"):
        o.field(f)[...] = g.download(f)
    o.call("interp_phi_multigrid"); o.call("interp_temp_multigrid")
    cmp(o, g, "phic")
    cmp(o, g, "tempr")


# ---- mixed FP32/FP64 form of the phase-field step ----
PF_SPHERE = dict(PF, nxm=24, nym=16, nzm=16, nxmr=48, nymr=32, nzmr=32, pf_r0=0.2, pf_eps=0.03, rayt=1e6)


def sphere_start(cfg, precision, seed=3):
    g = pyafid.Solver(dict(cfg, precision=precision))
    nz, ny = cfg["nzm"], cfg["nym"]
    g.upload("tempbp", np.ones((nz, ny, 1)))
    g.upload("temptp", np.zeros((nz, ny, 1)))
    g.create_initial_conditions(seed, 1e-2)
    return g


def l2rel(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / max(np.linalg.norm(np.asarray(b).ravel()), 1e-300))


def test_mixed_phase_vs_fp64():
    """Mixed form (FP32 phi, tempr, explicit terms and transfers; FP64
    state and solves; refined Poisson) against the FP64 form from the same
    seeded sphere state.  Bounds as for the mixed salinity step: max|div u|
    <= 1e-10, Nu_vol and KE within 1e-6 relative, relative L2 of the fields
    <= 1e-5 (FP32 rounding is 6e-8 per update)."""
    cfg = PF_SPHERE
    a = sphere_start(cfg, pyafid.FP64)
    b = sphere_start(cfg, pyafid.MIXED)
    for _ in range(8):
        a.time_marcher(2e-3)
        b.time_marcher(2e-3)
    nxm, nxmr = cfg["nxm"], cfg["nxmr"]
    assert a.check_divergence()[0] < 1e-12
    assert b.check_divergence()[0] <= 1e-10
    qa, qb = a.global_quantities(), b.global_quantities()
    assert qa["ke"] > 1e-10
    for k in ("nu", "ke"):
        assert abs(qb[k] - qa[k]) <= 1e-6 * abs(qa[k]), (k, qa[k], qb[k])
    for f, sl in (("vx", np.s_[..., 1:nxm]), ("vz", np.s_[..., :nxm]), ("vy", np.s_[..., :nxm]),
                  ("temp", np.s_[..., :nxm]), ("phi", np.s_[..., :nxmr]), ("tempr", np.s_[..., :nxmr]),
                  ("phic", np.s_[..., :nxm])):
        e = l2rel(b.download(f)[sl], a.download(f)[sl])
        assert e <= 1e-5, f"{f}: relative L2 {e:.3e}"
    pa, pb = a.download("pr")[..., :nxm], b.download("pr")[..., :nxm]
    assert l2rel(pb - pb.mean(), pa - pa.mean()) <= 1e-5
    # the sphere has started to melt in both
    assert abs(a.download("phi").sum() - b.download("phi").sum()) <= 1e-5 * a.download("phi").sum()


def test_mixed_phase_vs_oracle():
    cfg = PF_SPHERE
    b = sphere_start(cfg, pyafid.MIXED)
    o = pyoracle.Oracle(cfg)
    for f in ("vx", "vy", "vz", "temp", "pr", "tempbp", "temptp", "phi", "phic", "tempr"):
        o.field(f)[...] = b.download(f)
    for _ in range(3):
        b.time_marcher(2e-3)
        o.time_marcher(2e-3)
    nxm, nxmr = cfg["nxm"], cfg["nxmr"]
    for f, sl in (("vz", np.s_[..., :nxm]), ("temp", np.s_[..., :nxm]), ("phi", np.s_[..., :nxmr])):
        e = l2rel(b.download(f)[sl], o.field(f)[sl])
        assert e <= 1e-5, f"{f}: relative L2 {e:.3e}"


def test_mixed_phase_refuses_per_routine_calls():
    g = pyafid.Solver(dict(PF, precision=pyafid.MIXED))
    g.set_substep(1e-3, 1)
    for r in ("explicit_phase", "implicit_phase", "add_volume_penalty", "add_latent_heat"):
        with pytest.raises(pyafid.AfidError, match="FP64 form"):
            g.call(r)
    g.call("interp_phi_multigrid")          # the transfers run in both forms
    g.call("interp_temp_multigrid")
