"""The two forms of the base column kernel (ExplicitTerms*, ImplicitAndUpdate*
and SolveImpEqnUpdate_{X,YZ,Temp} for vx, vy, vz, T in the fused step):

* the single pass (the default up to 512 wall-normal cells), which forms the
  four right-hand sides in one sweep and sizes its solves by the nxm rows;
* the four passes, one field each (the form above 512 cells, and the one
  AFID_COLUMN_FOUR_PASS=1 selects), with solves sized by the nx faces.

Both against the FP64 oracle after whole steps (1e-10 of each field's max
norm, as the other step tests), on grids where nxm is a multiple of 64 (the
single pass leaves the face nx out of its solves), odd, with the phase field
(masked solves), and with nxm = 520, where the four passes are the default.
The forms against each other: the same terms in the same order, the scan
association aside, so within 1e-12 in FP64; in the mixed form the FP32
explicit terms of the second step see inputs that differ by that round-off,
so within 1e-7.
"""
import numpy as np
import pytest

import pyafid
import pyoracle
from helpers import load_state, random_state, relerr
from test_gpu_phase import PF, pf_state

pytestmark = pytest.mark.gpu

FIELDS = ("vx", "vy", "vz", "temp", "pr")
RBC64 = dict(nxm=64, nym=8, nzm=12, istr3=8, rayt=1e6, prat=0.7)
RBC_ODD = dict(nxm=70, nym=6, nzm=8, istr3=4, str3=1.5, rayt=1e5, prat=1.0, gaxis=2)
RBC_TALL = dict(nxm=520, nym=4, nzm=6, istr3=6, str3=1.2, rayt=1e7, prat=1.0)
PF64 = dict(PF, nxm=64, nxmr=128)


def run(cfg, st, four, monkeypatch, form="fp64", steps=2, dt=1e-3):
    monkeypatch.setenv("AFID_COLUMN_FOUR_PASS", "1" if four else "0")
    g = pyafid.Solver(dict(cfg, precision=pyafid.MIXED if form == "mixed" else pyafid.FP64))
    load_state(g, st, upload=True)
    for _ in range(steps):
        g.time_marcher(dt)
    out = {f: g.download(f) for f in FIELDS}
    g.close()
    return out


def oracle(cfg, st, steps=2, dt=1e-3):
    o = pyoracle.Oracle(cfg)
    load_state(o, st, upload=False)
    for _ in range(steps):
        o.time_marcher(dt)
    out = {f: o.field(f).copy() for f in FIELDS}
    o.close()
    return out


def interior(cfg, name, a):
    a = a[..., :cfg["nxm"]] if name != "vx" else a
    return a - a.mean() if name == "pr" else a


@pytest.mark.parametrize("cfg,pf", [(RBC64, False), (RBC_ODD, False), (RBC_TALL, False), (PF64, True)],
                         ids=["nxm64", "odd", "nxm520", "pf"])
def test_forms_match_the_oracle(cfg, pf, monkeypatch):
    st = pf_state(cfg, 7) if pf else random_state(cfg, 7)
    ref = oracle(cfg, st)
    for four in (False, True):
        got = run(cfg, st, four, monkeypatch)
        for f in FIELDS:
            e = relerr(interior(cfg, f, got[f]), interior(cfg, f, ref[f]))
            assert e <= 1e-10, (four, f, e)


@pytest.mark.parametrize("form,tol", [("fp64", 1e-12), ("mixed", 1e-7)])
@pytest.mark.parametrize("cfg,pf", [(RBC64, False), (PF64, True)], ids=["nxm64", "pf"])
def test_forms_agree(cfg, pf, form, tol, monkeypatch):
    st = pf_state(cfg, 9) if pf else random_state(cfg, 9)
    a = run(cfg, st, False, monkeypatch, form)
    b = run(cfg, st, True, monkeypatch, form)
    for f in FIELDS:
        e = relerr(interior(cfg, f, a[f]), interior(cfg, f, b[f]))
        assert e <= tol, (f, e)
    assert np.isfinite(a["vx"]).all()
