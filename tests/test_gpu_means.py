"""GPU parity of afid_calc_mean_profiles (CalcMeanProfiles.F90:11-229)
against the oracle on the same seeded random states, in both forms.

Tolerance: FP64 in both; the GPU sums each profile over 4 wavefronts and
up to 512 column groups in a fixed order, the oracle in the reference's
serial loop order, so agreement is to round-off (1e-12 relative to each
profile's max norm).
"""
import numpy as np
import pytest

import pyafid
import pyoracle
from helpers import load_state, random_state, relerr, uniform

pytestmark = pytest.mark.gpu

CFGS = [
    dict(nxm=16, nym=12, nzm=8, istr3=8, rayt=1e5, prat=1.0),
    dict(nxm=70, nym=6, nzm=12, istr3=6, str3=0.0, rayt=2e7, prat=7.0, inslws=0, tfixn=0, xplusu=0.3, xminusu=-0.2),
    dict(nxm=130, nym=4, nzm=3, istr3=8, rayt=1e6, prat=0.7),
]


def pair(cfg, seed):
    st = random_state(cfg, seed)
    st["tempbp"] = 1.0 + 0.1 * uniform(seed + 7, st["tempbp"].shape)
    st["temptp"] = 0.1 * uniform(seed + 8, st["temptp"].shape)
    o = pyoracle.Oracle(cfg)
    load_state(o, st, upload=False)
    return o, st


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("ci", range(len(CFGS)))
def test_mean_profiles_match_oracle(ci, form):
    cfg = CFGS[ci]
    o, st = pair(cfg, 11 + ci)
    g = pyafid.Solver(dict(cfg, precision=pyafid.MIXED if form == "mixed" else pyafid.FP64))
    load_state(g, st, upload=True)
    a, b = g.mean_profiles(), o.mean_profiles()
    for n in pyafid.MEAN_PROFILES:
        e = relerr(a[n], b[n])
        assert e <= 1e-12, f"{n}: relative error {e:.2e}"
    # a second call on the same state gives the same bits (fixed summation order)
    c = g.mean_profiles()
    for n in pyafid.MEAN_PROFILES:
        assert np.array_equal(a[n], c[n]), n
    g.close()
    o.close()


def test_mean_profiles_after_steps_and_closed_form():
    """After RK3 steps from the conductive state the profiles still match the
    oracle's, and in the conductive state itself chiT = 1/pect."""
    cfg = dict(nxm=24, nym=8, nzm=8, istr3=8, rayt=1e5, prat=1.0)
    g = pyafid.Solver(cfg)
    o = pyoracle.Oracle(cfg)
    xm = o.grid("xm")[:24]
    st = {n: np.zeros((8, 8, 25)) for n in ("vx", "vy", "vz", "temp", "pr", "rux", "ruy", "ruz", "rutemp")}
    st["temp"][..., :24] = 1.0 - xm
    st["tempbp"] = np.ones((8, 8, 1)); st["temptp"] = np.zeros((8, 8, 1))
    load_state(g, st, upload=True); load_state(o, st, upload=False)
    m = g.mean_profiles()
    assert np.allclose(m["chiT"], 1.0 / np.sqrt(1e5), rtol=1e-13, atol=0)
    o.field("temp")[..., :24] += 1e-3 * (uniform(3, (8, 8, 24)) - 0.5)
    g.upload("temp", o.field("temp").copy())
    for _ in range(2):
        o.time_marcher(1e-3)
        g.time_marcher(1e-3)
    a, b = g.mean_profiles(), o.mean_profiles()
    # profiles that average to (nearly) zero are compared on the scale of
    # what they average: U for the velocity means and fluxes, U^2 for the
    # velocity correlations (U the largest rms velocity)
    U = max(b["vxrms"].max(), b["vyrms"].max(), b["vzrms"].max())
    scale = {"vybar": U, "vzbar": U, "vxT": U, "vyT": U, "vzT": U, "vxvy": U * U, "vxvz": U * U, "vyvz": U * U}
    for n in pyafid.MEAN_PROFILES:
        den = max(np.abs(b[n]).max(), scale.get(n, 0.0))
        e = np.abs(a[n] - b[n]).max() / den
        assert e <= 1e-11, f"{n}: {e:.2e}"



MR = dict(nxm=16, nym=8, nzm=12, istr3=8, multires=1, salinity=1, nxmr=32, nymr=16, nzmr=24, istr3r=8,
          rayt=1e6, prat=1.0, rays=-1e6, pras=4.0)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("cfg", [MR, dict(MR, nxmr=70, nymr=24, nzmr=12, sfixn=0, istr3r=6, str3=0.0)])
def test_sal_stats_match_oracle(cfg, form):
    """CalcSalStats on random refined fields.  The mixed form stores the
    refined fields in FP32: the oracle gets the values the GPU holds
    (downloaded), so both sides sum the same numbers in FP64."""
    g = pyafid.Solver(dict(cfg, precision=pyafid.MIXED if form == "mixed" else pyafid.FP64))
    o = pyoracle.Oracle(cfg)
    shp = o.field("sal").shape
    for q, n in enumerate(("sal", "vxr", "vyr", "vzr")):
        g.upload(n, uniform(40 + q, shp) - 0.5)
    g.upload("salbp", 1.0 + 0.1 * uniform(50, o.field("salbp").shape))
    g.upload("saltp", 0.1 * uniform(51, o.field("saltp").shape))
    for n in ("sal", "vxr", "vyr", "vzr", "salbp", "saltp"):
        o.field(n)[...] = g.download(n)
    a, b = g.sal_stats(), o.sal_stats()
    for n in pyafid.SAL_PROFILES:
        e = relerr(a[n], b[n])
        assert e <= 1e-12, f"{n}: relative error {e:.2e}"
    c = g.sal_stats()
    for n in pyafid.SAL_PROFILES:
        assert np.array_equal(a[n], c[n]), n
    g.close()
    o.close()


def test_sal_stats_refused_without_salinity():
    g = pyafid.Solver(dict(nxm=8, nym=4, nzm=4, istr3=8))
    with pytest.raises(pyafid.AfidError, match="salinity"):
        g.sal_stats()
    g.close()


PF = dict(nxm=16, nym=12, nzm=12, istr3=8, multires=1, phasefield=1, nxmr=32, nymr=24, nzmr=24, istr3r=8,
          rayt=1e6, prat=1.0, pf_a=1.0, pf_s=1.0, pf_tm=0.0)


@pytest.mark.parametrize("form", ["fp64", "mixed"])
@pytest.mark.parametrize("cfg", [PF, dict(PF, nxm=12, nym=8, nzm=6, nxmr=30, nymr=20, nzmr=18, istr3=4, str3=1.4, istr3r=6)])
def test_phi_stats_match_oracle(cfg, form):
    """CalcPhiStats on a random refined phase field (the oracle sums the values
    the GPU holds, FP32 in the mixed form)."""
    g = pyafid.Solver(dict(cfg, precision=pyafid.MIXED if form == "mixed" else pyafid.FP64))
    o = pyoracle.Oracle(cfg)
    g.upload("phi", uniform(60, o.field("phi").shape))
    o.field("phi")[...] = g.download("phi")
    a, b = g.phi_stats(), o.phi_stats()
    for n in ("phibar", "phirms"):
        e = relerr(a[n], b[n])
        assert e <= 1e-12, f"{n}: relative error {e:.2e}"
    g.close()
    o.close()


def test_phi_stats_refused_without_phase_field():
    g = pyafid.Solver(dict(nxm=8, nym=4, nzm=4, istr3=8))
    with pytest.raises(pyafid.AfidError, match="phase field"):
        g.phi_stats()
    g.close()
