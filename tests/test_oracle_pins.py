"""Pins of oracle functions that the step parity tests alone would not catch
(the GPU and the oracle could share a mistake): closed forms and invariants
that share no code with the oracle, each with planted mistakes (made in the
oracle, on the CPU) that must fail it.

  * GlobalQuantities (GlobalQuantities.F90:39-91): kinetic energy, rms
    velocity, mean temperature and the volume Nusselt number of fields whose
    weighted sums have closed forms.
  * The rest state under buoyancy: T = 1 - x, u = 0.  The discrete
    hydrostatic balance udx3c(k) (p_k - p_{k-1}) = byct (T_k + T_{k-1}) / 2
    (the pressure-gradient term of ImplicitAndUpdateVX.F90:37-68 against the
    buoyancy of ExplicitTermsVX.F90:117-127) is reached in the first substep
    and kept, with u = 0 and T unchanged, only if the RHS pressure gradient,
    the buoyancy and CorrectPressure (pressure.F90:270-301) are all right.
  * The x-diffusion of vz (ImplicitAndUpdateVZ.F90:38-70 with
    SolveImpEqnUpdate_YZ): vz = sin(m pi x) at the cell centres is an exact
    eigenvector of the no-slip discrete operator, so one RK3 step multiplies
    it by prod_s (1 + a_s b lam) / (1 - a_s b lam).  (vz has no wall velocity
    in the reference, so a Couette profile is only possible in vy, which
    tests/test_oracle.py pins.)
"""
import numpy as np
import pytest

import pyoracle as po

GAM = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
ROM = (0.0, -17.0 / 60.0, -5.0 / 12.0)


def fresh(cfg):
    o = po.Oracle(cfg)
    for n in ("vx", "vy", "vz", "temp", "pr", "rux", "ruy", "ruz", "rutemp"):
        o.field(n)[...] = 0.0
    o.field("tempbp")[...] = 1.0
    o.field("temptp")[...] = 0.0
    return o


# ------------------------------------------------------------ GlobalQuantities
def quantities_error(istr3, nxm=16, ny=8, nz=6):
    """Largest relative error of (KE, vrms, tmean, Nu) against their closed
    forms on a grid symmetric about x = 1/2 (there sum_k g3rc(k) = dx alx3,
    from g3rc(1) = xc(2) dx and g3rc(k) = (xc(k+1) - xc(k-1)) dx / 2)."""
    ra, pr = 4e6, 0.7
    cfg = dict(nxm=nxm, nym=ny, nzm=nz, istr3=istr3, str3=1.5, rayt=ra, prat=pr, ylen=1.5, zlen=0.75)
    o = fresh(cfg)
    A, B = 0.3, -0.2
    j = np.arange(ny)
    i = np.arange(nz)
    # sum_j sin^2(2 pi j / ny) = ny / 2 and sum_i cos^2(2 pi i / nz) = nz / 2 (ny, nz > 2)
    o.field("vy")[...] = A * np.sin(2 * np.pi * j / ny)[None, :, None]
    o.field("vz")[...] = B * np.cos(2 * np.pi * i / nz)[:, None, None]
    q = o.global_quantities()
    ke = (A * A + B * B) / 4.0
    errs = [abs(q["ke"] - ke) / ke,
            abs(q["vrms"] - np.sqrt(2 * ke) * np.sqrt(ra / pr)) / (np.sqrt(2 * ke) * np.sqrt(ra / pr))]
    # constant T: tmean = T
    o.field("temp")[..., :nxm] = 0.625
    errs.append(abs(o.global_quantities()["tmean"] - 0.625) / 0.625)
    o.close()
    return max(errs)


def nusselt_error(nxm=12):
    """Uniform grid (g3rc = 1), T = 1 - xm, vx = C on the interior faces
    (vx = 0 on both walls): sum_k g3rc T_k (vx_k + vx_{k+1}) / 2 = C (n - 1) / 2,
    so Nu = 1 + sqrt(Pr Ra) C (n - 1) / (2 n); and tmean = 1/2."""
    ra, pr, Cv = 2e5, 2.0, 0.01
    o = fresh(dict(nxm=nxm, nym=4, nzm=4, istr3=0, rayt=ra, prat=pr))
    xm = (np.arange(nxm) + 0.5) / nxm
    o.field("temp")[..., :nxm] = 1.0 - xm
    o.field("vx")[..., 1:nxm] = Cv
    q = o.global_quantities()
    nu = 1.0 + np.sqrt(pr * ra) * Cv * (nxm - 1) / (2.0 * nxm)
    ke = 0.5 * Cv * Cv * (nxm - 1) / nxm          # interior faces k = 2..n carry g3rc = 1 each
    o.close()
    return max(abs(q["nu"] - nu) / nu, abs(q["tmean"] - 0.5) / 0.5, abs(q["ke"] - ke) / ke)


@pytest.mark.parametrize("istr3", [0, 4, 8])
def test_global_quantities_closed_forms(istr3):
    assert quantities_error(istr3) < 1e-13


def test_nusselt_closed_form():
    assert nusselt_error() < 1e-13


@pytest.mark.parametrize("plant,check", [(4, lambda: quantities_error(8)), (5, nusselt_error)],
                         ids=["ke_without_vy", "nu_lower_face"])
def test_global_quantities_pins_catch_planted_mistakes(plant, check):
    with po.planted(plant):
        assert check() > 1e-7                     # the pins hold to 1e-13


# ------------------------------------------------------------ hydrostatic rest
def hydrostatic_error(istr3=4, nxm=16, steps=2, dt=1e-2):
    """Start from rest with T = 1 - xm (the conductive state, itself steady)
    and p = 0; after `steps` RK3 steps return the largest of max|u|, the
    change of T and the relative error of the discrete hydrostatic balance."""
    cfg = dict(nxm=nxm, nym=6, nzm=4, istr3=istr3, str3=1.5, rayt=400.0, prat=1.0, gaxis=1, active_t=1)
    o = fresh(cfg)
    xm = o.grid("xm")[:nxm]
    T0 = np.zeros_like(o.field("temp"))
    T0[..., :nxm] = 1.0 - xm
    o.field("temp")[...] = T0
    for _ in range(steps):
        o.time_marcher(dt)
    u = max(np.abs(o.field(n)).max() for n in ("vx", "vy", "vz"))
    dT = np.abs(o.field("temp")[..., :nxm] - T0[..., :nxm]).max()
    p = o.field("pr")[..., :nxm]
    T = o.field("temp")[..., :nxm]
    udx3c = o.grid("udx3c")                 # 1-based kc -> index kc - 1
    byct = o.scalars()["byct"]
    # faces kc = 2..nxm (0-based k = 1..nxm-1)
    lhs = udx3c[1:nxm][None, None, :] * (p[..., 1:] - p[..., :-1])
    rhs = byct * 0.5 * (T[..., 1:] + T[..., :-1])
    bal = np.abs(lhs - rhs).max() / np.abs(rhs).max()
    o.close()
    return max(u, dT, bal)


@pytest.mark.parametrize("istr3", [0, 4, 8])
def test_rest_state_is_hydrostatic(istr3):
    assert hydrostatic_error(istr3) < 1e-12


@pytest.mark.parametrize("plant", [1, 2, 3], ids=["pressure_gradient_sign", "no_alpha_beta_in_correct_pressure",
                                                   "buoyancy_without_half"])
def test_rest_state_pin_catches_planted_mistakes(plant):
    with po.planted(plant):
        assert hydrostatic_error() > 1e-7         # the pin holds to 1e-12


# ------------------------------------------------------------ vz x-diffusion
def vz_decay_error(m=3, nxm=24, dt=2e-3):
    """Uniform grid, no-slip walls: vz_k = sin(m pi xm_k) has ghost values
    -vz at both walls, so the discrete x-Laplacian maps it to lam vz with
    lam = -4 n^2 sin^2(m pi / (2 n)).  Nothing else acts on a field that
    depends on x only (no advection, no divergence, no buoyancy), and one RK3
    step multiplies it by prod_s (1 + a_s b lam / Re) / (1 - a_s b lam / Re),
    a_s = gamma_s + rho_s, b = dt / 2."""
    ra, pr = 900.0, 1.0
    re = np.sqrt(ra / pr)
    o = fresh(dict(nxm=nxm, nym=4, nzm=4, istr3=0, rayt=ra, prat=pr, gaxis=1, active_t=0))
    xm = (np.arange(nxm) + 0.5) / nxm
    v0 = np.sin(m * np.pi * xm)
    o.field("vz")[..., :nxm] = v0
    o.time_marcher(dt)
    lam = -4.0 * nxm * nxm * np.sin(m * np.pi / (2 * nxm)) ** 2 / re
    g = 1.0
    for ga, ro in zip(GAM, ROM):
        a = ga + ro
        g *= (1 + a * 0.5 * dt * lam) / (1 - a * 0.5 * dt * lam)
    err = np.abs(o.field("vz")[..., :nxm] - g * v0).max()
    other = max(np.abs(o.field(n)).max() for n in ("vx", "vy"))
    o.close()
    return max(err, other) / (1.0 - g)        # relative to the decay itself


def test_vz_diffusion_eigenmode():
    assert vz_decay_error() < 1e-10


def test_vz_pin_catches_planted_mistake():
    with po.planted(6):
        assert vz_decay_error() > 1e-7            # the pin holds to 1e-10


# ------------------------------------------- refined -> coarse, general path
# Base grid clustered at the walls (cosine), refined grid uniform: the first
# refined cell centre lies above the first coarse one, so InterpSalMultigrid,
# InterpPhiMultigrid and AddLatentHeat take interpolate_xyz_to_coarse
# (HermiteInterpolations.F90:218-250) instead of the fast loop.
GEN = dict(nxm=24, nym=8, nzm=6, istr3=8, multires=1, salinity=1, nxmr=40, nymr=16, nzmr=12, istr3r=0,
           rayt=1e6, prat=1.0, rays=-1e6, pras=4.0)


def test_general_path_is_taken():
    o = po.Oracle(GEN)
    assert o.grid("xmr")[0] >= o.grid("xm")[0]
    o.close()


@pytest.mark.parametrize("gaxis", [1, 2, 3])
def test_general_path_reproduces_constants_and_lines(gaxis):
    """Partition of unity and linear exactness at every coarse centre, wall
    cells included: S = a + b x with its wall values as the Dirichlet ghosts
    (salinity.F90:466-482) lies on one line, which the linear wall intervals
    and the Hermite interior intervals both reproduce.  The general path puts
    salc at the coarse cell centres for every gAxis (its weights are the
    cxphic / cyphic / czphic of the cell centres)."""
    o = po.Oracle(dict(GEN, gaxis=gaxis))
    xmr = o.grid("xmr")[:GEN["nxmr"]]
    xm = o.grid("xm")[:GEN["nxm"]]
    for a, b in ((0.3, 0.0), (0.5, -1.0), (-0.2, 2.5)):
        S = np.zeros_like(o.field("sal"))
        S[..., :GEN["nxmr"]] = a + b * xmr
        o.field("sal")[...] = S
        o.field("salbp")[...] = a
        o.field("saltp")[...] = a + b
        o.call("interp_sal_multigrid")
        err = np.abs(o.field("salc")[..., :GEN["nxm"]] - (a + b * xm)).max()
        assert err < 1e-13, (a, b, err)
    o.close()


def test_general_path_phi_constant():
    """phi with zero-gradient ghosts (phasefield.F90:479-487): a constant
    transfers exactly, through the general path."""
    cfg = dict(GEN, salinity=0, phasefield=1, pf_a=1.0, pf_s=1.0, pf_tm=0.0)
    o = po.Oracle(cfg)
    o.field("phi")[...] = 0.375
    o.call("interp_phi_multigrid")
    assert np.abs(o.field("phic")[..., :GEN["nxm"]] - 0.375).max() < 1e-14
    o.close()
