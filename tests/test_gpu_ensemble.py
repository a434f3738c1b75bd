"""GPU parity of the batched 2-D ensemble (include/afid_ensemble.h) against
the FP64 CPU oracle run by run, with the reference's driver loop
(main.F90:281-422: CFL-based dt, exit codes) restated in Python around the
oracle's TimeMarcher.

Tolerance: FP64 both sides; kernels reorder sums and fuse multiply-adds, so
fields agree to round-off relative to their max norm (1e-11 after a few
steps, the bound of the single-run whole-step tests); pressure modulo its
mean.  Launch shapes and graph replay change no arithmetic: those results
must be bit-identical.
"""
import numpy as np
import pytest

import pyafid
import pyoracle
from helpers import relerr, uniform

pytestmark = pytest.mark.gpu

# uneven runs: wall-normal sizes on both sides of a 64-row lane chunk,
# periodic sizes with factors 2 and 3, different Ra / Pr / aspect ratio
RUNS = [
    dict(nxm=16, nzm=32, rayt=1e5, prat=1.0, zlen=2.0),
    dict(nxm=24, nzm=48, rayt=3e5, prat=7.0, zlen=2.0),
    dict(nxm=70, nzm=36, rayt=1e6, prat=1.0, zlen=1.0, istr3=8),
    dict(nxm=33, nzm=64, rayt=2e6, prat=7.0, zlen=3.0),
    dict(nxm=130, nzm=24, rayt=5e5, prat=1.0, zlen=0.5),
]
BASE = dict(nym=1, ylen=1.0, alx3=1.0, istr3=8, nsst=3, gaxis=1, active_t=1)
TOL = 1e-11


def cfgs(runs=RUNS):
    return [dict(BASE, **r) for r in runs]


def random_run_state(cfg, seed):
    """Admissible random state of a 2-D run (v = 0, vx = 0 on the walls)."""
    nx, nz = cfg["nxm"] + 1, cfg["nzm"]
    st = {}
    for q, name in enumerate(["vx", "vz", "temp", "pr", "rux", "ruz", "rutemp"]):
        a = 0.1 * (uniform(seed * 100 + q, (nz, 1, nx)) - 0.5)
        a[..., nx - 1] = 0.0
        st[name] = a
    st["vx"][..., 0] = 0.0
    st["temp"][..., : nx - 1] += 1.0 - np.linspace(0, 1, nx - 1)
    st["tempbp"] = np.ones((nz, 1, 1))
    st["temptp"] = np.zeros((nz, 1, 1))
    return st


def load_both(ens, run, o, st):
    for name, a in st.items():
        ens.upload(run, name, a)
        o.field(name)[...] = a


def oracle_driver(o, ctl, nsteps):
    """main.F90:281-385 for one run around the oracle's TimeMarcher; returns
    (time, dt of the next step, ntime, status, steps done)."""
    dt, time, status, steps = ctl["dt"], 0.0, 0, 0
    dtmax = ctl.get("dtmax", 0.0) or dt
    ntst = ctl.get("ntst", 1 << 30)

    def top(ntime, dt):
        cfl = max(o.calc_max_cfl()[0], 1e-8)
        st = 0
        if ctl.get("variable_dt"):
            if ntime > 1:
                dt = dtmax if cfl < 1e-8 else min(ctl["limit_cfl"] / cfl, dtmax)
            if dt < ctl.get("dtmin", 1e-8):
                st = 166
        elif cfl * dt > ctl["limit_cfl"]:
            st = 165
        return dt, st

    ntime = 0
    dt, status = top(ntime, dt)
    while not status and steps < nsteps:
        o.time_marcher(dt)
        steps += 1
        time += dt
        dmax = o.check_divergence()[0]
        if not dmax <= ctl.get("resid", 1e-3):
            status = 169
        elif ntime == ntst:
            status = 555
        ntime += 1
        if not status:
            dt, status = top(ntime, dt)
    return dict(time=time, dt=dt, ntime=ntime, status=status, steps=steps)


def compare_run(ens, run, o, tol=TOL):
    cfg = ens.cfgds[run]
    nxm = cfg["nxm"]
    for f in ("vx", "vz", "temp"):
        g = ens.download(run, f)
        ref = o.field(f)
        sl = np.s_[...] if f == "vx" else np.s_[..., :nxm]
        e = relerr(g[sl], ref[sl])
        assert e <= tol, f"run {run} {f}: relative error {e:.3e}"
    pg = ens.download(run, "pr")[..., :nxm]
    po = o.field("pr")[..., :nxm]
    assert relerr(pg - pg.mean(), po - po.mean()) <= 1e-10, f"run {run} pr"
    assert np.all(ens.download(run, "vy") == 0.0)


def test_fixed_dt_matches_oracle():
    cs = cfgs()
    ctl = dict(variable_dt=0, dt=2e-3, limit_cfl=10.0)
    ens = pyafid.Ensemble(cs, [ctl] * len(cs))
    orcs = []
    for r, c in enumerate(cs):
        o = pyoracle.Oracle(c)
        load_both(ens, r, o, random_run_state(c, 7 + r))
        orcs.append(o)
    ens.reset()
    ens.step(4, graph=False)
    for r, o in enumerate(orcs):
        ref = oracle_driver(o, ctl, 4)
        s = ens.run_state(r)
        assert s["status"] == 0 and s["steps"] == 4 and s["ntime"] == 4
        assert abs(s["time"] - ref["time"]) <= 1e-15
        compare_run(ens, r, o)
        cfl_o = o.calc_max_cfl()[0]
        assert abs(s["cfl"] - cfl_o) <= 1e-12 * cfl_o
        assert s["max_div"] <= 1e-10
    ens.close()


def test_variable_dt_and_initial_conditions_match_oracle():
    """Seeded initial state (CreateInitialConditions, RBC branch) and the
    CFL-controlled dt of IDTV = 1 over 6 steps: dt switches to CFLMAX/CFL
    from the third step on."""
    cs = cfgs(RUNS[:4])
    ctl = dict(variable_dt=1, dt=1e-3, limit_cfl=0.5, dtmax=2e-2, dtmin=1e-8)
    ens = pyafid.Ensemble(cs, [ctl] * len(cs))
    for r, c in enumerate(cs):
        ens.upload(r, "tempbp", np.ones((c["nzm"], 1, 1)))
        ens.upload(r, "temptp", np.zeros((c["nzm"], 1, 1)))
    ens.create_initial_conditions(3000, 0.05)
    orcs = []
    for r, c in enumerate(cs):
        o = pyoracle.Oracle(c)
        for f in ("vx", "vz", "temp", "pr", "rux", "ruz", "rutemp", "tempbp", "temptp"):
            o.field(f)[...] = ens.download(r, f)
        orcs.append(o)
    # the noise differs between runs (seed 3000 + r)
    assert not np.array_equal(ens.download(0, "temp")[:4, :, :8], ens.download(1, "temp")[:4, :, :8])
    ens.step(6, graph=False)
    for r, o in enumerate(orcs):
        ref = oracle_driver(o, ctl, 6)
        s = ens.run_state(r)
        assert s["status"] == 0 and s["steps"] == 6
        assert abs(s["time"] - ref["time"]) <= 1e-12 * ref["time"]
        assert abs(s["dt"] - ref["dt"]) <= 1e-10 * ref["dt"]
        compare_run(ens, r, o)
        gq, go = ens.global_quantities(r), o.global_quantities()
        assert abs(gq["nu"] - go["nu"]) <= 1e-11 * abs(go["nu"])
    ens.close()


def _fields(ens):
    return [{f: ens.download(r, f) for f in ("vx", "vz", "temp", "pr", "rux", "ruz", "rutemp")} for r in range(ens.n)]


def _fresh(cs, ctl, seed=11):
    ens = pyafid.Ensemble(cs, [ctl] * len(cs))
    for r, c in enumerate(cs):
        for name, a in random_run_state(c, seed + r).items():
            ens.upload(r, name, a)
    ens.reset()
    return ens


def test_graph_replay_and_launch_shapes_are_bit_identical():
    """The persistent form under three launch shapes and the phased form
    under two line caps, replayed from a captured graph, against eager
    launches of the default shape."""
    cs = cfgs()
    ctl = dict(variable_dt=1, dt=1e-3, limit_cfl=0.5, dtmax=1e-2)
    a = _fresh(cs, ctl)
    a.step(13, graph=False)
    ref = _fields(a)
    a.close()
    shapes = [dict(threads=1024, workgroups=0, lines=8), dict(threads=256, workgroups=2, lines=2),
              dict(threads=512, workgroups=3, lines=4), dict(phased=1, lines=1), dict(phased=1, lines=4)]
    for sh in shapes:
        b = _fresh(cs, ctl)
        b.set_tuning(**sh)
        b.step(13, graph=True)       # one captured 10-step graph + 3 launched steps
        got = _fields(b)
        for r in range(len(cs)):
            for f in ref[r]:
                assert np.array_equal(got[r][f], ref[r][f]), f"{sh} run {r} {f}"
        b.close()


def test_exit_codes_freeze_runs():
    cs = cfgs(RUNS[:4])
    ctls = [dict(variable_dt=1, dt=1e-3, limit_cfl=0.5, dtmax=1e-2, ntst=2),     # 555 after steps 0, 1, 2
            dict(variable_dt=1, dt=1e-3, limit_cfl=0.5, dtmin=1.0),               # 166 before the first step
            dict(variable_dt=0, dt=1e-3, limit_cfl=1e-9),                          # 165 before the first step
            dict(variable_dt=1, dt=1e-3, limit_cfl=0.5, dtmax=1e-2, tmax=2.5e-3)]  # 333 after the third step
    ens = pyafid.Ensemble(cs, ctls)
    orcs = []
    for r, c in enumerate(cs):
        o = pyoracle.Oracle(c)
        load_both(ens, r, o, random_run_state(c, 21 + r))
        orcs.append(o)
    ens.reset()
    ens.step(12, graph=True)
    expect = [(555, 3), (166, 0), (165, 0), (333, 3)]
    for r, (code, steps) in enumerate(expect):
        s = ens.run_state(r)
        assert (s["status"], s["steps"]) == (code, steps), (r, s)
    # frozen runs hold the state they stopped with
    for r in (0, 1, 2):
        oracle_driver(orcs[r], ctls[r], 12)
        compare_run(ens, r, orcs[r])


def test_autotune_keeps_state_and_results():
    cs = cfgs()
    ctl = dict(variable_dt=1, dt=1e-3, limit_cfl=0.5, dtmax=1e-2)
    a = _fresh(cs, ctl, seed=31)
    before = _fields(a)
    best, log = a.autotune(steps=1)
    assert log and best["threads"] in (256, 512, 1024) and best["ms_per_step"] > 0
    assert {c["phased"] for c in log} == {0, 1}
    after = _fields(a)
    for r in range(len(cs)):
        for f in before[r]:
            assert np.array_equal(after[r][f], before[r][f])
    assert all(a.run_state(r)["steps"] == 0 for r in range(len(cs)))
    a.step(5)
    tuned = _fields(a)
    a.close()
    b = _fresh(cs, ctl, seed=31)
    b.step(5)
    plain = _fields(b)
    b.close()
    for r in range(len(cs)):
        for f in plain[r]:
            assert np.array_equal(tuned[r][f], plain[r][f])


def test_rejects_unsupported_runs():
    with pytest.raises(pyafid.AfidError, match="nym = 1"):
        pyafid.Ensemble([dict(BASE, nxm=16, nzm=16, nym=4)], [dict(dt=1e-3)])
    with pytest.raises(pyafid.AfidError, match="nsst"):
        pyafid.Ensemble([dict(BASE, nxm=16, nzm=16), dict(BASE, nxm=16, nzm=16, nsst=1)], [dict(dt=1e-3)] * 2)
    ens = pyafid.Ensemble([dict(BASE, nxm=16, nzm=16)], [dict(dt=1e-3)])
    with pytest.raises(pyafid.AfidError, match="identically zero"):
        ens.upload(0, "vy", np.ones((16, 1, 17)))
    ens.close()


# ---- mixed FP32/FP64 form of the ensemble ----
def test_mixed_ensemble_vs_fp64():
    """Mixed form (FP32 explicit terms and Poisson sweeps, FP64 state,
    solves and residual; further sweeps while max|div u| > 1e-10) against
    the FP64 ensemble from the same states: max|div u| <= 1e-10 in every
    run, Nu within 1e-6, relative L2 of the fields <= 1e-5."""
    cs = cfgs()
    ctl = dict(variable_dt=1, dt=1e-3, limit_cfl=0.5, dtmax=1e-2)
    a = _fresh(cs, ctl, seed=41)
    b = pyafid.Ensemble([dict(c, precision=pyafid.MIXED) for c in cs], [ctl] * len(cs))
    for r, c in enumerate(cs):
        for name, arr in random_run_state(c, 41 + r).items():
            b.upload(r, name, arr)
    b.reset()
    a.step(6, graph=False)
    b.step(6, graph=True)
    for r, c in enumerate(cs):
        sa, sb = a.run_state(r), b.run_state(r)
        assert sb["status"] == 0 and sb["steps"] == 6
        assert sb["max_div"] <= 1e-10, (r, sb["max_div"])
        nxm = c["nxm"]
        for f, sl in (("vx", np.s_[..., 1:nxm]), ("vz", np.s_[..., :nxm]), ("temp", np.s_[..., :nxm])):
            x, y = b.download(r, f)[sl], a.download(r, f)[sl]
            e = float(np.linalg.norm((x - y).ravel()) / np.linalg.norm(y.ravel()))
            assert e <= 1e-5, f"run {r} {f}: relative L2 {e:.3e}"
        na, nb = a.global_quantities(r)["nu"], b.global_quantities(r)["nu"]
        assert abs(nb - na) <= 1e-6 * abs(na)
    a.close()
    b.close()


def test_mixed_ensemble_shapes_bit_identical_and_phased_only():
    cs = cfgs()
    ctl = dict(variable_dt=1, dt=1e-3, limit_cfl=0.5, dtmax=1e-2)
    mk = lambda: pyafid.Ensemble([dict(c, precision=pyafid.MIXED) for c in cs], [ctl] * len(cs))
    outs = []
    for lines, graph in ((2, False), (1, True), (4, True)):
        e = mk()
        for r, c in enumerate(cs):
            for name, arr in random_run_state(c, 51 + r).items():
                e.upload(r, name, arr)
        e.reset()
        e.set_tuning(phased=1, lines=lines)
        e.step(12, graph=graph)
        outs.append(_fields(e))
        if lines == 2:
            with pytest.raises(pyafid.AfidError, match="phased"):
                e.set_tuning(threads=1024, phased=0)
        e.close()
    for o in outs[1:]:
        for r in range(len(cs)):
            for f in o[r]:
                assert np.array_equal(o[r][f], outs[0][r][f])
