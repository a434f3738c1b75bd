"""afid_time_marcher_graph: RK3 steps replayed from a captured HIP graph are
bit-identical to the eager afid_time_marcher steps, replay after replay,
with eager work (reductions, downloads) between the replays; a new dt
recaptures."""
import numpy as np
import pytest

import pyafid

pytestmark = pytest.mark.gpu

CASES = {
    "rbc": dict(nxm=32, nym=24, nzm=16, istr3=8, rayt=1e5, prat=1.0),
    "multires": dict(nxm=16, nym=12, nzm=12, istr3=8, multires=1, salinity=1, nxmr=32, nymr=24, nzmr=24, istr3r=8,
                     rayt=1e6, prat=1.0, rays=-1e6, pras=4.0),
    "phasefield": dict(nxm=16, nym=12, nzm=12, istr3=8, multires=1, phasefield=1, nxmr=32, nymr=24, nzmr=24, istr3r=8,
                       rayt=1e6, prat=1.0, pf_a=1.0, pf_s=1.0, pf_tm=0.0, pf_r0=0.2, pf_eps=0.03),
}
NAMES = {"rbc": ["vx", "vy", "vz", "temp", "pr"], "multires": ["vx", "vy", "vz", "temp", "pr", "sal", "vxr"],
         "phasefield": ["vx", "vy", "vz", "temp", "pr", "phi", "phic", "tempr"]}


def start(cfg, form):
    g = pyafid.Solver(dict(cfg, precision=pyafid.MIXED if form == "mixed" else pyafid.FP64))
    nz, ny = cfg["nzm"], cfg["nym"]
    g.upload("tempbp", np.ones((nz, ny, 1)))
    g.upload("temptp", np.zeros((nz, ny, 1)))
    if cfg.get("salinity"):
        g.upload("salbp", np.ones((cfg["nzmr"], cfg["nymr"], 1)))
        g.upload("saltp", np.zeros((cfg["nzmr"], cfg["nymr"], 1)))
    g.create_initial_conditions(5, 0.05)
    return g


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_graph_replay_bit_identical(case, form):
    cfg = CASES[case]
    a, b = start(cfg, form), start(cfg, form)
    for rep in range(3):                        # three replays of the 2-step graph
        for _ in range(2):
            a.time_marcher(2e-3)
        b.time_marcher_graph(2e-3, 2)
        da, db = a.check_divergence(), b.check_divergence()     # eager work between replays
        assert da == db, (rep, da, db)
        for n in NAMES[case]:
            assert np.array_equal(a.download(n), b.download(n)), f"replay {rep}: {n}"
    # an odd count: one replay and one eager step; then a new dt recaptures
    for _ in range(3):
        a.time_marcher(2e-3)
    b.time_marcher_graph(2e-3, 3)
    for _ in range(2):
        a.time_marcher(1e-3)
    b.time_marcher_graph(1e-3, 2)
    for n in NAMES[case]:
        assert np.array_equal(a.download(n), b.download(n)), n
    a.close(); b.close()
