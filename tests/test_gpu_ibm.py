"""GPU parity of the immersed-boundary solves (afid_set_ibm + the implicit
updates; src/ibm/SolveImpEqnUpdate_*_ibm.F90) against the oracle on the same
state, masks and distance lists: liquid, solid and boundary rows in one
column set, every solve kind.  FP64 in both, the same rows solved by Thomas
(oracle) and by the Moebius-scan wavefront solve (GPU): 1e-12 relative."""
import numpy as np
import pytest

import pyafid
import pyoracle
from helpers import load_state, random_state, relerr, uniform

pytestmark = pytest.mark.gpu

CFG = dict(nxm=40, nym=8, nzm=6, istr3=4, str3=1.5, rayt=1e5, prat=1.0)
MRC = dict(CFG, nxm=20, multires=1, salinity=1, nxmr=72, nymr=16, nzmr=12, istr3r=4, rays=-1e5, pras=4.0)
ROUT = {"vx": "implicit_and_update_vx", "vy": "implicit_and_update_vy", "vz": "implicit_and_update_vz",
        "temp": "implicit_and_update_temp", "sal": "implicit_salinity"}
H = {"vx": "qcap", "vy": "dph", "vz": "dq", "temp": "hro", "sal": "hsal"}


def random_mask(shape, n, seed):
    """Liquid everywhere but: solid slabs in x, with an upper boundary point
    just below and a lower one just above each (as a body across the
    column), in a random half of the columns."""
    r = np.random.default_rng(seed)
    m = np.full(shape, 2, dtype=np.int32)
    for i in range(shape[0]):
        for j in range(shape[1]):
            if r.random() < 0.5:
                k0 = int(r.integers(2, n // 2))
                k1 = int(r.integers(k0 + 1, n - 2))
                m[i, j, k0:k1] = 0
                m[i, j, k0 - 1] = 1 if r.random() < 0.7 else 2
                m[i, j, k1] = -1 if r.random() < 0.7 else 2
    return m


@pytest.mark.parametrize("field", sorted(ROUT))
def test_ibm_solve_matches_oracle(field):
    cfg = MRC if field == "sal" else CFG
    st = random_state(cfg, 7)
    o, g = pyoracle.Oracle(cfg), pyafid.Solver(cfg)
    load_state(o, st, upload=False)
    load_state(g, st, upload=True)
    hs = o.field(H[field]).shape
    h = uniform(77, hs) - 0.5
    o.field(H[field])[...] = h
    g.upload(H[field], h)
    n = cfg["nxmr"] if field == "sal" else cfg["nxm"]
    m = random_mask(hs, n, 11)
    nb = int(np.sum((m[..., :n] == 1) | (m[..., :n] == -1)))
    dist = None if field == "sal" else 0.2 + 0.6 * uniform(88, (nb,))
    o.set_ibm(field, m, dist)
    g.set_ibm(field, m, dist)
    for ns in (1, 2):
        o.set_substep(2e-3, ns); g.set_substep(2e-3, ns)
        o.call(ROUT[field]); g.call(ROUT[field])
        sl = np.s_[...] if field == "vx" else np.s_[..., :n]
        e = relerr(g.download(field)[sl], o.field(field)[sl])
        assert e <= 1e-12, f"{field} substep {ns}: {e:.3e}"
    # the fused step refuses while a mask is set; clearing it restores the plain solve
    with pytest.raises(pyafid.AfidError, match="immersed-boundary"):
        g.time_marcher(1e-3)
    g.set_ibm(field, None)
    g.time_marcher(1e-3)
    o.close(); g.close()
