"""Pins of the oracle's immersed-boundary solves (src/ibm/
SolveImpEqnUpdate_X_ibm.F90, _YZ_ibm, _Temp_ibm, _Sal_ibm):

  * an all-liquid mask gives the plain solve bit for bit;
  * solid rows force the new value: u = 0, T = Tb = 1;
  * an upper / lower boundary row ties the new value to its neighbour,
    q(k) = d q(k+1) / q(k) = d q(k-1) (T: q - Tb = d (q(k+-1) - Tb)), with d
    taken from the list in the reference's loop order (z, y, x), which the
    check reads back from the result;
  * S treats solid rows as liquid and ties boundary points with d = 1.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import random_state, load_state, uniform

CFG = dict(nxm=12, nym=6, nzm=5, istr3=4, str3=1.5, rayt=1e5, prat=1.0)
MRC = dict(CFG, multires=1, salinity=1, nxmr=24, nymr=12, nzmr=10, istr3r=4, rays=-1e5, pras=4.0)
ROUT = {"vx": "implicit_and_update_vx", "vy": "implicit_and_update_vy", "vz": "implicit_and_update_vz",
        "temp": "implicit_and_update_temp", "sal": "implicit_salinity"}
H = {"vx": "qcap", "vy": "dph", "vz": "dq", "temp": "hro", "sal": "hsal"}


def pair(field, seed=3):
    cfg = MRC if field == "sal" else CFG
    st = random_state(cfg, seed)
    a, b = po.Oracle(cfg), po.Oracle(cfg)
    for o in (a, b):
        load_state(o, st, upload=False)
        o.field(H[field])[...] = uniform(seed + 50, o.field(H[field]).shape) - 0.5
        o.set_substep(2e-3, 2)
    return cfg, a, b


def shape(cfg, field):
    return (cfg["nzmr"], cfg["nymr"], cfg["nxmr"] + 1) if field == "sal" else (cfg["nzm"], cfg["nym"], cfg["nxm"] + 1)


@pytest.mark.parametrize("field", sorted(ROUT))
def test_all_liquid_is_the_plain_solve(field):
    cfg, a, b = pair(field)
    b.set_ibm(field, np.full(shape(cfg, field), 2, dtype=np.int32), np.zeros(1))
    a.call(ROUT[field]); b.call(ROUT[field])
    assert np.array_equal(a.field(field), b.field(field))
    a.close(); b.close()


@pytest.mark.parametrize("field", ["vx", "vy", "vz", "temp"])
def test_solid_rows_force_the_value(field):
    cfg, a, b = pair(field)
    m = np.full(shape(cfg, field), 2, dtype=np.int32)
    m[1, 2, :] = 0                       # one solid column
    m[3, 4, 5] = 0                       # one solid row inside a liquid column
    b.set_ibm(field, m, np.zeros(1))
    b.call(ROUT[field])
    q = b.field(field)
    tb = 1.0 if field == "temp" else 0.0
    n = cfg["nxm"]
    lo = 1 if field == "vx" else 0
    assert np.abs(q[1, 2, lo:n] - tb).max() < 1e-15
    assert abs(q[3, 4, 5] - tb) < 1e-15
    a.close(); b.close()


@pytest.mark.parametrize("field", ["vx", "vy", "vz", "temp", "sal"])
def test_boundary_rows_and_list_order(field):
    cfg, a, b = pair(field)
    shp = shape(cfg, field)
    n = cfg["nxmr"] if field == "sal" else cfg["nxm"]
    m = np.full(shp, 2, dtype=np.int32)
    pts = [(0, 1, 4, 1), (2, 3, 6, -1), (4, 0, n - 2, 1), (1, 5 % shp[1], 3, -1)]   # (i, j, k, type)
    for i, j, k, t in pts:
        m[i, j, k] = t
    order = sorted(pts, key=lambda p: (p[0], p[1], p[2]))          # the reference's loop order
    dist = [0.25 + 0.1 * q for q in range(len(order))]
    b.set_ibm(field, m, None if field == "sal" else np.array(dist))
    b.call(ROUT[field])
    q = b.field(field)
    tb = 1.0 if field == "temp" else 0.0
    for q_i, (i, j, k, t) in enumerate(order):
        d = 1.0 if field == "sal" else dist[q_i]
        nb = k + 1 if t == 1 else k - 1
        assert abs((q[i, j, k] - tb) - d * (q[i, j, nb] - tb)) < 1e-13, (field, i, j, k, t)
    # columns without a boundary point are the plain solve
    a.call(ROUT[field])
    untouched = np.ones(shp[:2], dtype=bool)
    for i, j, _, _ in pts:
        untouched[i, j] = False
    assert np.array_equal(a.field(field)[untouched], q[untouched])
    a.close(); b.close()


def test_short_distance_list_is_refused():
    cfg, a, b = pair("vy")
    m = np.full(shape(cfg, "vy"), 2, dtype=np.int32)
    m[0, 0, 3] = 1
    m[0, 1, 3] = -1
    with pytest.raises(RuntimeError):
        b.set_ibm("vy", m, np.array([0.5]))
    a.close(); b.close()
