"""ctypes view of the FP64 CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the measured or
shipped path.
"""
import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

FIELDS = ["vx", "vy", "vz", "temp", "pr", "rux", "ruy", "ruz", "rutemp",
          "qcap", "dph", "dq", "hro", "rhs",
          "sal", "rusal", "hsal", "rhsr", "vxr", "vyr", "vzr", "salc",
          "tempbp", "temptp", "salbp", "saltp",
          "phi", "ruphi", "hphi", "phic", "tempr",
          "humid", "ruhumid", "hhumid", "qsat", "humbp", "humtp"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}

GRID = ["xc", "xm", "g3rc", "g3rm", "d3xc", "d3xm", "udx3c", "udx3m",
        "ap3ck", "ac3ck", "am3ck", "ap3sk", "ac3sk", "am3sk", "ap3ssk", "ac3ssk", "am3ssk",
        "ak1", "ak2", "amphk", "acphk", "apphk",
        "xcr", "xmr", "udx3mr", "g3rmr", "ap3sskr", "ac3sskr", "am3sskr",
        "cxrs", "cyrs", "czrs", "cxsalc", "cysalc", "czsalc",
        "ap3spkr", "ac3spkr", "am3spkr", "cxphic", "cyphic", "czphic"]
GRID_ID = {n: i for i, n in enumerate(GRID)}


class OrcConfig(C.Structure):
    """Mirror of orc_config (the bou.in value lines, ReadInputFile.F90:20-76)."""
    _fields_ = [
        ("nxm", C.c_int32), ("nym", C.c_int32), ("nzm", C.c_int32), ("nsst", C.c_int32),
        ("multires", C.c_int32), ("nxmr", C.c_int32), ("nymr", C.c_int32), ("nzmr", C.c_int32),
        ("salinity", C.c_int32), ("phasefield", C.c_int32),
        ("alx3", C.c_double), ("ylen", C.c_double), ("zlen", C.c_double),
        ("istr3", C.c_int32), ("str3", C.c_double), ("istr3r", C.c_int32),
        ("rayt", C.c_double), ("prat", C.c_double), ("rays", C.c_double), ("pras", C.c_double),
        ("ffscales", C.c_int32),
        ("inslws", C.c_int32), ("inslwn", C.c_int32), ("tfixs", C.c_int32), ("tfixn", C.c_int32),
        ("sfixs", C.c_int32), ("sfixn", C.c_int32),
        ("active_t", C.c_int32), ("active_s", C.c_int32), ("gaxis", C.c_int32),
        ("xplusu", C.c_double), ("xminusu", C.c_double), ("dpdy", C.c_double), ("dpdz", C.c_double),
        ("pf_a", C.c_double), ("pf_s", C.c_double), ("pf_tm", C.c_double), ("pf_eps", C.c_double),
        ("moist", C.c_int32), ("q_alpha", C.c_double), ("q_beta", C.c_double), ("q_gamma", C.c_double),
        ("q_tau", C.c_double), ("q_ifdiff", C.c_int32), ("q_fixn", C.c_int32), ("q_fixs", C.c_int32),
        ("q_sm", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(OrcConfig)]
        L.orc_destroy.argtypes = [vp]
        L.orc_field.restype = C.POINTER(C.c_double)
        L.orc_field.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
        L.orc_grid.restype = C.POINTER(C.c_double)
        L.orc_grid.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
        L.orc_int_grid.restype = C.c_int
        L.orc_int_grid.argtypes = [vp, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.POINTER(C.c_int32))]
        L.orc_scalars.argtypes = [vp, C.POINTER(C.c_double)]
        L.orc_set_substep.argtypes = [vp, C.c_double, C.c_int]
        L.orc_time_marcher.argtypes = [vp, C.c_double]
        for name in ["explicit_terms_vx", "explicit_terms_vy", "explicit_terms_vz", "explicit_terms_temp",
                     "implicit_and_update_vx", "implicit_and_update_vy", "implicit_and_update_vz",
                     "implicit_and_update_temp", "calc_local_divergence", "solve_pressure_correction",
                     "correct_velocity", "correct_pressure", "explicit_salinity", "add_sal_buoyancy",
                     "implicit_salinity", "interp_vel_mgrd", "interp_sal_multigrid",
                     "explicit_phase", "add_volume_penalty", "implicit_phase", "add_latent_heat",
                     "interp_temp_multigrid", "interp_phi_multigrid",
                     "set_humidity_bcs", "explicit_humidity", "add_condensation", "implicit_humidity",
                     "update_saturation"]:
            getattr(L, "orc_" + name).argtypes = [vp]
        L.orc_calc_max_cfl.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_check_divergence.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_global_quantities.argtypes = [vp, C.POINTER(C.c_double)]
        L.orc_set_plant.argtypes = [C.c_int]
        L.orc_set_ibm.argtypes = [vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_int64]
        L.orc_interp_input_vel.argtypes = [vp, C.POINTER(OldGrid)] + [C.POINTER(C.c_double)] * 4
        L.orc_interp_input_sal.argtypes = [vp, C.POINTER(OldGrid), C.POINTER(C.c_double)]
        L.orc_interp_input_phi.argtypes = [vp, C.POINTER(OldGrid), C.POINTER(C.c_double)]
        L.orc_interp_input_humid.argtypes = [vp, C.POINTER(OldGrid), C.POINTER(C.c_double)]
        L.orc_mean_profiles.argtypes = [vp, C.POINTER(C.c_double)]
        L.orc_sal_stats.argtypes = [vp, C.POINTER(C.c_double)]
        L.orc_phi_stats.argtypes = [vp, C.POINTER(C.c_double)]
        L.orc_moist_stats.argtypes = [vp, C.POINTER(C.c_double)]
        dp = C.POINTER(C.c_double)
        L.orc_fft.argtypes = [dp, dp, C.c_int, C.c_int]
        L.orc_dft_naive.argtypes = [dp, dp, dp, dp, C.c_int, C.c_int]
        L.orc_tridiag_solve.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_pf_scalars.argtypes = [vp, C.POINTER(C.c_double)]
        L.orc_moist_scalars.argtypes = [vp, C.POINTER(C.c_double)]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One oracle instance; fields are numpy views (nz, ny, nx) of its state."""

    def __init__(self, cfg: dict):
        self.cfg = make_config(cfg)
        self.L = lib()
        self.h = self.L.orc_create(C.byref(self.cfg))
        if not self.h:
            raise RuntimeError("orc_create failed")

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field(self, name):
        dims = (C.c_int64 * 3)()
        p = self.L.orc_field(self.h, FIELD_ID[name], dims)
        if not p:
            raise KeyError(name)
        shape = tuple(int(d) for d in dims)
        return np.ctypeslib.as_array(p, shape=shape)

    def grid(self, name):
        n = C.c_int64()
        p = self.L.orc_grid(self.h, GRID_ID[name], C.byref(n))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(int(n.value),)).copy()

    def int_grid(self, which):
        n = C.c_int64()
        out = C.POINTER(C.c_int32)()
        rc = self.L.orc_int_grid(self.h, which, C.byref(n), C.byref(out))
        if rc:
            raise KeyError(which)
        return np.ctypeslib.as_array(out, shape=(int(n.value),)).copy()

    def scalars(self):
        out = (C.c_double * 8)()
        self.L.orc_scalars(self.h, out)
        return dict(zip(["ren", "pect", "pecs", "byct", "bycs", "dx", "dy", "dz"], list(out)))

    def pf_scalars(self):
        out = (C.c_double * 4)()
        self.L.orc_pf_scalars(self.h, out)
        return dict(zip(["pf_d", "pf_eps", "pf_eta", "pf_b"], list(out)))

    def moist_scalars(self):
        out = (C.c_double * 5)()
        self.L.orc_moist_scalars(self.h, out)
        return dict(zip(["alpha_q", "beta_q", "gamma_q", "tau_q", "pecq"], list(out)))

    def set_substep(self, dt, ns):
        self.L.orc_set_substep(self.h, dt, ns)

    def call(self, name):
        getattr(self.L, "orc_" + name)(self.h)

    def set_ibm(self, field, mask, dist=None):
        """The immersed-boundary mask (nz, ny, nx) of a solve and its boundary
        distances in loop order (orc_set_ibm)."""
        m = np.ascontiguousarray(mask, dtype=np.int32)
        d = np.ascontiguousarray(dist if dist is not None else np.zeros(1), dtype=np.float64)
        if self.L.orc_set_ibm(self.h, FIELD_ID[field], m.ctypes.data_as(C.POINTER(C.c_int32)), _dp(d),
                              0 if dist is None else d.size):
            raise RuntimeError("orc_set_ibm failed")

    def interp_input_vel(self, og, vx, vy, vz, temp):
        """InterpInputVel: old fields (nzo-1, nyo-1, nxo) onto this grid."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (vx, vy, vz, temp)]
        rc = self.L.orc_interp_input_vel(self.h, C.byref(og), *[_dp(x) for x in a])
        if rc:
            raise RuntimeError("orc_interp_input_vel failed")

    def interp_input_ref(self, og, name, f):
        """InterpInputSal (name 'sal') / InterpInputPhi ('phi') onto this refined grid."""
        a = np.ascontiguousarray(f, dtype=np.float64)
        fn = self.L.orc_interp_input_sal if name == "sal" else self.L.orc_interp_input_phi
        if fn(self.h, C.byref(og), _dp(a)):
            raise RuntimeError("orc_interp_input_" + name + " failed")

    def interp_input_humid(self, og, q):
        """InterpInputHum: the old humidity (nzo-1, nyo-1, nxo) onto this base grid."""
        a = np.ascontiguousarray(q, dtype=np.float64)
        if self.L.orc_interp_input_humid(self.h, C.byref(og), _dp(a)):
            raise RuntimeError("orc_interp_input_humid failed")

    def time_marcher(self, dt):
        self.L.orc_time_marcher(self.h, dt)

    def calc_max_cfl(self):
        a, b = C.c_double(), C.c_double()
        self.L.orc_calc_max_cfl(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def check_divergence(self):
        a, b = C.c_double(), C.c_double()
        self.L.orc_check_divergence(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def global_quantities(self):
        out = (C.c_double * 4)()
        self.L.orc_global_quantities(self.h, out)
        return dict(zip(["nu", "ke", "vrms", "tmean"], list(out)))

    MEAN_PROFILES = ["Tbar", "vybar", "vzbar", "vxT", "vyT", "vzT", "Trms", "vxrms", "vyrms", "vzrms",
                     "vxvy", "vxvz", "vyvz", "chiT", "epsilon"]

    def mean_profiles(self):
        """CalcMeanProfiles.F90:11-229: dict of the 15 profiles (nxm values each)."""
        nxm = int(self.cfg.nxm)
        out = np.zeros(15 * nxm)
        self.L.orc_mean_profiles(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return {n: out[q * nxm:(q + 1) * nxm].copy() for q, n in enumerate(self.MEAN_PROFILES)}

    SAL_PROFILES = ["Sbar", "Srms", "vxS", "vyS", "vzS", "chiS"]

    def phi_stats(self):
        """CalcPhiStats (phasefield.F90:750-798): dict of phibar, phirms."""
        n = int(self.cfg.nxmr)
        out = np.zeros(2 * n)
        self.L.orc_phi_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return {"phibar": out[:n].copy(), "phirms": out[n:].copy()}

    def moist_stats(self):
        """CalcMoistStats (moisture.F90:422-505): dict of the 7 profiles (nxm values each)."""
        n = int(self.cfg.nxm)
        out = np.zeros(7 * n)
        self.L.orc_moist_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        names = ["qbar", "qsbar", "qrms", "qrel", "vxq", "vyq", "vzq"]
        return {m: out[q * n:(q + 1) * n].copy() for q, m in enumerate(names)}

    def sal_stats(self):
        """CalcSalStats (salinity.F90:510-690): dict of the 6 refined profiles (nxmr values each)."""
        n = int(self.cfg.nxmr)
        out = np.zeros(6 * n)
        self.L.orc_sal_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return {m: out[q * n:(q + 1) * n].copy() for q, m in enumerate(self.SAL_PROFILES)}


DEFAULTS = dict(nxm=16, nym=16, nzm=16, nsst=3, multires=0, nxmr=0, nymr=0, nzmr=0,
                salinity=0, phasefield=0, alx3=1.0, ylen=1.0, zlen=1.0,
                istr3=8, str3=0.0, istr3r=8, rayt=1e5, prat=1.0, rays=-1e5, pras=1.0, ffscales=0,
                inslws=1, inslwn=1, tfixs=1, tfixn=1, sfixs=1, sfixn=1,
                active_t=1, active_s=0, gaxis=1, xplusu=0.0, xminusu=0.0, dpdy=0.0, dpdz=0.0,
                pf_a=1.0, pf_s=1.0, pf_tm=0.0, pf_eps=0.0,
                moist=0, q_alpha=3.0, q_beta=1.2, q_gamma=0.19, q_tau=5e-5, q_ifdiff=1, q_fixn=1, q_fixs=1, q_sm=1.0)


def make_config(d):
    v = dict(DEFAULTS)
    v.update(d)
    c = OrcConfig()
    for k, _ in OrcConfig._fields_:
        setattr(c, k, v[k])
    return c


def fft(x, sign=-1):
    re = np.ascontiguousarray(np.real(x), dtype=np.float64).copy()
    im = np.ascontiguousarray(np.imag(x), dtype=np.float64).copy()
    lib().orc_fft(_dp(re), _dp(im), len(re), sign)
    return re + 1j * im


def dft_naive(x, sign=-1):
    re = np.ascontiguousarray(np.real(x), dtype=np.float64)
    im = np.ascontiguousarray(np.imag(x), dtype=np.float64)
    ore, oim = np.empty_like(re), np.empty_like(re)
    lib().orc_dft_naive(_dp(re), _dp(im), _dp(ore), _dp(oim), len(re), sign)
    return ore + 1j * oim


def tridiag_solve(a, b, c, d):
    a, b, c = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c))
    d = np.ascontiguousarray(d, dtype=np.float64).copy()
    lib().orc_tridiag_solve(len(d), _dp(a), _dp(b), _dp(c), _dp(d))
    return d


def set_threads(n):
    lib().orc_set_threads(int(n))


class OldGrid(C.Structure):
    """The old run's grid of a checkpoint (continua_master.h5 scalars)."""
    _fields_ = [("nxo", C.c_int32), ("nyo", C.c_int32), ("nzo", C.c_int32), ("istr3o", C.c_int32),
                ("str3o", C.c_double), ("nxro", C.c_int32), ("nyro", C.c_int32), ("nzro", C.c_int32),
                ("istr3ro", C.c_int32)]


def old_grid(cfg):
    """OldGrid of a run with config dict cfg (the grid its checkpoint records)."""
    mr = cfg.get("multires") or cfg.get("salinity") or cfg.get("phasefield")
    return OldGrid(cfg["nxm"] + 1, cfg["nym"] + 1, cfg["nzm"] + 1, cfg.get("istr3", 0), float(cfg.get("str3", 0.0)),
                   (cfg["nxmr"] + 1) if mr else 0, (cfg["nymr"] + 1) if mr else 0, (cfg["nzmr"] + 1) if mr else 0,
                   cfg.get("istr3r", 0) if mr else 0)


class planted:
    """Context manager: the oracle runs with planted mistake n (see
    orc_set_plant in afid_oracle.c) inside the block."""

    def __init__(self, n):
        self.n = int(n)

    def __enter__(self):
        lib().orc_set_plant(self.n)
        return self

    def __exit__(self, *exc):
        lib().orc_set_plant(0)
        return False
