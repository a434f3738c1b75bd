"""Python binding of libafid_gpu.so (the C ABI declared in include/afid_gpu.h).

The reference's host loop (src/main.F90:281-422) drives TimeMarcher and the
CFL / divergence reductions; `Solver` exposes the same calls with the same
names so a host written in Python reads like the Fortran one.  There is no
CPU fallback: constructing a Solver without the built library, or without a
GPU, raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libafid_gpu.so")

FIELDS = ["vx", "vy", "vz", "temp", "pr", "rux", "ruy", "ruz", "rutemp",
          "qcap", "dph", "dq", "hro", "rhs",
          "sal", "rusal", "hsal", "rhsr", "vxr", "vyr", "vzr", "salc",
          "tempbp", "temptp", "salbp", "saltp",
          "phi", "ruphi", "hphi", "phic", "tempr",
          "humid", "ruhumid", "hhumid", "qsat", "humbp", "humtp"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
REFINED = {"sal", "rusal", "hsal", "rhsr", "vxr", "vyr", "vzr", "salbp", "saltp", "phi", "ruphi", "hphi", "tempr"}
PLANES = {"tempbp", "temptp", "salbp", "saltp", "humbp", "humtp"}

# every entry point declared in include/afid_gpu.h
EXPORTS = [
    "afid_gpu_init", "afid_gpu_finalize", "afid_last_error", "afid_device_bytes",
    "afid_upload", "afid_download", "afid_set_substep",
    "afid_explicit_terms_vx", "afid_explicit_terms_vy", "afid_explicit_terms_vz", "afid_explicit_terms_temp",
    "afid_explicit_salinity", "afid_add_sal_buoyancy",
    "afid_implicit_and_update_vx", "afid_implicit_and_update_vy", "afid_implicit_and_update_vz",
    "afid_implicit_and_update_temp", "afid_implicit_salinity",
    "afid_calc_local_divergence", "afid_solve_pressure_correction", "afid_correct_velocity", "afid_correct_pressure",
    "afid_explicit_phase", "afid_add_volume_penalty", "afid_implicit_phase", "afid_add_latent_heat",
    "afid_interp_temp_multigrid", "afid_interp_phi_multigrid",
    "afid_interp_vel_mgrd", "afid_interp_sal_multigrid", "afid_time_marcher", "afid_time_marcher_graph", "afid_refinement_residuals",
    "afid_calc_max_cfl", "afid_check_divergence", "afid_global_quantities", "afid_calc_mean_profiles", "afid_calc_mean_profiles_noalloc", "afid_calc_sal_stats", "afid_calc_phi_stats", "afid_calc_moist_stats",
    "afid_synchronize", "afid_set_ibm",
    "afid_set_humidity_bcs", "afid_explicit_humidity", "afid_add_condensation", "afid_implicit_humidity",
    "afid_update_saturation",
    "afid_interp_input_vel", "afid_interp_input_sal", "afid_interp_input_phi", "afid_interp_input_humid",
    "afid_create_initial_conditions",
    "afid_set_timing", "afid_get_timing", "afid_grid_array",
]

# the profiles of afid_calc_mean_profiles, in the order of means.h5 (CalcMeanProfiles.F90:230-283)
MEAN_PROFILES = ["Tbar", "vybar", "vzbar", "vxT", "vyT", "vzT", "Trms", "vxrms", "vyrms", "vzrms",
                 "vxvy", "vxvz", "vyvz", "chiT", "epsilon"]
# the refined-grid profiles of afid_calc_sal_stats (salinity.F90:583-597)
SAL_PROFILES = ["Sbar", "Srms", "vxS", "vyS", "vzS", "chiS"]

# every entry point declared in include/afid_ensemble.h
ENS_EXPORTS = [
    "afid_ens_init", "afid_ens_finalize", "afid_ens_last_error", "afid_ens_device_bytes",
    "afid_ens_upload", "afid_ens_download", "afid_ens_create_initial_conditions", "afid_ens_reset",
    "afid_ens_step", "afid_ens_step_timed", "afid_ens_run_state", "afid_ens_global_quantities",
    "afid_ens_autotune", "afid_ens_set_tuning", "afid_ens_get_tuning", "afid_ens_synchronize",
]

ROUTINES = ["explicit_terms_vx", "explicit_terms_vy", "explicit_terms_vz", "explicit_terms_temp",
            "explicit_salinity", "add_sal_buoyancy",
            "implicit_and_update_vx", "implicit_and_update_vy", "implicit_and_update_vz",
            "implicit_and_update_temp", "implicit_salinity",
            "calc_local_divergence", "solve_pressure_correction", "correct_velocity", "correct_pressure",
            "interp_vel_mgrd", "interp_sal_multigrid",
            "explicit_phase", "add_volume_penalty", "implicit_phase", "add_latent_heat",
            "interp_temp_multigrid", "interp_phi_multigrid",
            "set_humidity_bcs", "explicit_humidity", "add_condensation", "implicit_humidity", "update_saturation"]

FP64, MIXED = 0, 1

# libafid_io.so: the solver's HDF5 files (include/afid_io.h)
IO_LIB_PATH = os.path.join(HERE, "libafid_io.so")
IO_EXPORTS = [
    "afid_h5_last_error", "afid_h5_create_file", "afid_h5_create_group",
    "afid_h5_write_field3d", "afid_h5_read_field3d", "afid_h5_dataset_dims",
    "afid_h5_write_int_scalar", "afid_h5_write_real_scalar", "afid_h5_read_int_scalar", "afid_h5_read_real_scalar",
    "afid_h5_write_real1d", "afid_h5_read_real1d",
]
# checkpoint file of each field (HdfReadContinua.F90:44-61, WriteFlowField.F90:26-50)
CONTINUA = {"vx": "vx", "vy": "vy", "vz": "vz", "temp": "temp", "sal": "sal", "phi": "phi", "humid": "humid", "pr": "pr"}
# the phase-field profiles of afid_calc_phi_stats (phasefield.F90:788-793)
PHI_PROFILES = ["phibar", "phirms"]
# the moisture profiles of afid_calc_moist_stats (moisture.F90:485-500)
MOIST_PROFILES = ["qbar", "qsbar", "qrms", "qrel", "vxq", "vyq", "vzq"]


class OldGrid(C.Structure):
    """Mirror of afid_old_grid: the old run's grid as continua_master.h5
    records it (sizes count the faces)."""
    _fields_ = [("nxo", C.c_int32), ("nyo", C.c_int32), ("nzo", C.c_int32), ("istr3o", C.c_int32),
                ("str3o", C.c_double), ("ylen", C.c_double), ("zlen", C.c_double),
                ("nxro", C.c_int32), ("nyro", C.c_int32), ("nzro", C.c_int32), ("istr3ro", C.c_int32)]


class AfidConfig(C.Structure):
    """Mirror of afid_config: the bou.in value lines (ReadInputFile.F90:20-76)."""
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
        ("precision", C.c_int32),
        ("div_tol", C.c_double), ("max_sweeps", C.c_int32),
        ("pf_a", C.c_double), ("pf_s", C.c_double), ("pf_tm", C.c_double), ("pf_eps", C.c_double),
        ("pf_r0", C.c_double),
        ("moist", C.c_int32), ("q_alpha", C.c_double), ("q_beta", C.c_double), ("q_gamma", C.c_double),
        ("q_tau", C.c_double), ("q_ifdiff", C.c_int32), ("q_fixn", C.c_int32), ("q_fixs", C.c_int32),
        ("q_sm", C.c_double),
    ]


DEFAULTS = dict(nxm=16, nym=16, nzm=16, nsst=3, multires=0, nxmr=0, nymr=0, nzmr=0,
                salinity=0, phasefield=0, alx3=1.0, ylen=1.0, zlen=1.0,
                istr3=8, str3=0.0, istr3r=8, rayt=1e5, prat=1.0, rays=-1e5, pras=1.0, ffscales=0,
                inslws=1, inslwn=1, tfixs=1, tfixn=1, sfixs=1, sfixn=1,
                active_t=1, active_s=0, gaxis=1, xplusu=0.0, xminusu=0.0, dpdy=0.0, dpdz=0.0,
                precision=FP64, div_tol=0.0, max_sweeps=0,
                pf_a=1.0, pf_s=1.0, pf_tm=0.0, pf_eps=0.0, pf_r0=0.0,
                moist=0, q_alpha=3.0, q_beta=1.2, q_gamma=0.19, q_tau=5e-5, q_ifdiff=1, q_fixn=1, q_fixs=1, q_sm=1.0)


def make_config(d):
    v = dict(DEFAULTS)
    v.update(d)
    c = AfidConfig()
    for k, _ in AfidConfig._fields_:
        setattr(c, k, v[k])
    return c


class RunControl(C.Structure):
    """Mirror of afid_run_control: IDTV, DT, RESID, CFLMAX, DTMIN, DTMAX of a
    run's bou.in and its NTST / TMAX (main.F90:287-302, 355-385)."""
    _fields_ = [("variable_dt", C.c_int32), ("dt", C.c_double), ("resid", C.c_double), ("limit_cfl", C.c_double),
                ("dtmin", C.c_double), ("dtmax", C.c_double), ("ntst", C.c_int32), ("tmax", C.c_double)]


class EnsTuning(C.Structure):
    _fields_ = [("threads", C.c_int32), ("workgroups", C.c_int32), ("lines", C.c_int32), ("phased", C.c_int32),
                ("ms_per_step", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class RunState(C.Structure):
    _fields_ = [("time", C.c_double), ("dt", C.c_double), ("cfl", C.c_double), ("max_div", C.c_double),
                ("ntime", C.c_int32), ("status", C.c_int32), ("steps", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


CONTROL_DEFAULTS = dict(variable_dt=0, dt=1e-3, resid=1e-3, limit_cfl=1.0, dtmin=1e-8, dtmax=0.0, ntst=1 << 30, tmax=0.0)


def make_control(d):
    v = dict(CONTROL_DEFAULTS)
    v.update(d)
    c = RunControl()
    for k, _ in RunControl._fields_:
        setattr(c, k, v[k])
    return c


_lib = None


def lib():
    """Load libafid_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libafid_gpu.so is not built: run __graft_entry__.build() "
                               "or python afid-murphfi_amd/build_lib.py")
        L = C.CDLL(LIB_PATH)
        vp, i64p, dp = C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)
        L.afid_gpu_init.argtypes = [C.POINTER(AfidConfig), C.POINTER(vp)]
        L.afid_gpu_finalize.argtypes = [vp]
        L.afid_last_error.argtypes = [vp]
        L.afid_last_error.restype = C.c_char_p
        L.afid_device_bytes.argtypes = [vp]
        L.afid_device_bytes.restype = C.c_int64
        L.afid_upload.argtypes = [vp, C.c_int, dp, i64p, i64p]
        L.afid_download.argtypes = [vp, C.c_int, dp, i64p, i64p]
        L.afid_set_substep.argtypes = [vp, C.c_double, C.c_int]
        L.afid_time_marcher.argtypes = [vp, C.c_double]
        L.afid_time_marcher_graph.argtypes = [vp, C.c_double, C.c_int]
        for r in ROUTINES:
            getattr(L, "afid_" + r).argtypes = [vp]
        L.afid_calc_max_cfl.argtypes = [vp, dp, dp]
        L.afid_check_divergence.argtypes = [vp, dp, dp]
        L.afid_global_quantities.argtypes = [vp, dp]
        L.afid_calc_mean_profiles.argtypes = [vp, dp]
        # out is a device address (an integer), not a host array
        L.afid_calc_mean_profiles_noalloc.argtypes = [vp, C.c_void_p]
        L.afid_calc_sal_stats.argtypes = [vp, dp]
        L.afid_calc_phi_stats.argtypes = [vp, dp]
        L.afid_calc_moist_stats.argtypes = [vp, dp]
        L.afid_synchronize.argtypes = [vp]
        L.afid_interp_input_vel.argtypes = [vp, C.POINTER(OldGrid), dp, dp, dp, dp]
        L.afid_interp_input_sal.argtypes = [vp, C.POINTER(OldGrid), dp]
        L.afid_interp_input_phi.argtypes = [vp, C.POINTER(OldGrid), dp]
        L.afid_interp_input_humid.argtypes = [vp, C.POINTER(OldGrid), dp]
        L.afid_create_initial_conditions.argtypes = [vp, C.c_uint64, C.c_double]
        L.afid_set_timing.argtypes = [vp, C.c_int]
        L.afid_get_timing.argtypes = [vp, C.POINTER(C.c_char_p), dp, C.c_int, C.POINTER(C.c_int)]
        L.afid_grid_array.argtypes = [vp, C.c_char_p, dp, C.c_int64, i64p]
        L.afid_refinement_residuals.argtypes = [vp, dp, C.c_int, C.POINTER(C.c_int)]
        L.afid_ens_init.argtypes = [C.POINTER(AfidConfig), C.POINTER(RunControl), C.c_int32, C.POINTER(vp)]
        L.afid_ens_finalize.argtypes = [vp]
        L.afid_ens_last_error.argtypes = [vp]
        L.afid_ens_last_error.restype = C.c_char_p
        L.afid_ens_device_bytes.argtypes = [vp]
        L.afid_ens_device_bytes.restype = C.c_int64
        L.afid_ens_upload.argtypes = [vp, C.c_int32, C.c_int32, dp, i64p, i64p]
        L.afid_ens_download.argtypes = [vp, C.c_int32, C.c_int32, dp, i64p, i64p]
        L.afid_ens_create_initial_conditions.argtypes = [vp, C.c_uint64, C.c_double]
        L.afid_ens_reset.argtypes = [vp]
        L.afid_ens_step.argtypes = [vp, C.c_int32, C.c_int32]
        L.afid_ens_step_timed.argtypes = [vp, C.c_int32, C.c_int32, dp]
        L.afid_ens_run_state.argtypes = [vp, C.c_int32, C.POINTER(RunState)]
        L.afid_ens_global_quantities.argtypes = [vp, C.c_int32, dp]
        L.afid_ens_autotune.argtypes = [vp, C.c_int32, C.POINTER(EnsTuning), C.POINTER(EnsTuning), C.c_int32,
                                        C.POINTER(C.c_int32)]
        L.afid_ens_set_tuning.argtypes = [vp, C.POINTER(EnsTuning)]
        L.afid_ens_get_tuning.argtypes = [vp, C.POINTER(EnsTuning)]
        L.afid_ens_synchronize.argtypes = [vp]
        _lib = L
    return _lib


_io = None


def io_lib():
    """Load libafid_io.so (host-only HDF5 library); raises if it has not been built."""
    global _io
    if _io is None:
        if not os.path.exists(IO_LIB_PATH):
            raise RuntimeError("libafid_io.so is not built: run __graft_entry__.build() "
                               "or python afid-murphfi_amd/build_lib.py")
        L = C.CDLL(IO_LIB_PATH)
        cp, dp, i64 = C.c_char_p, C.POINTER(C.c_double), C.c_int64
        L.afid_h5_last_error.argtypes = []
        L.afid_h5_last_error.restype = cp
        L.afid_h5_create_file.argtypes = [cp]
        L.afid_h5_create_group.argtypes = [cp, cp]
        L.afid_h5_write_field3d.argtypes = [cp, dp, i64, i64, i64]
        L.afid_h5_read_field3d.argtypes = [cp, dp, i64, i64, i64]
        L.afid_h5_dataset_dims.argtypes = [cp, cp, C.POINTER(i64), C.c_int]
        L.afid_h5_write_int_scalar.argtypes = [cp, cp, C.c_int32]
        L.afid_h5_write_real_scalar.argtypes = [cp, cp, C.c_double]
        L.afid_h5_read_int_scalar.argtypes = [cp, cp, C.POINTER(C.c_int32)]
        L.afid_h5_read_real_scalar.argtypes = [cp, cp, dp]
        L.afid_h5_write_real1d.argtypes = [cp, cp, dp, i64]
        L.afid_h5_read_real1d.argtypes = [cp, cp, dp, i64]
        for n in IO_EXPORTS:
            if n != "afid_h5_last_error":
                getattr(L, n).restype = C.c_int
        _io = L
    return _io


class H5Error(RuntimeError):
    pass


def _h5(rc, what):
    if rc != 0:
        raise H5Error(f"{what} failed ({rc}): {io_lib().afid_h5_last_error().decode()}")


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def h5_create_file(path):
    """HdfCreateBlankFile (HdfRoutines.F90:11-22)."""
    _h5(io_lib().afid_h5_create_file(os.fsencode(path)), "afid_h5_create_file " + str(path))


def h5_create_group(path, name):
    _h5(io_lib().afid_h5_create_group(os.fsencode(path), name.encode()), f"afid_h5_create_group {name}")


def h5_write_field(path, arr):
    """HdfWriteRealHalo3D (HdfRoutines.F90:428-520): `arr` is [nz][ny][nx]
    (x fastest), stored as dataset "var" of Fortran dims (nx, ny, nz)."""
    a = np.ascontiguousarray(arr, dtype=np.float64)
    nz, ny, nx = a.shape
    _h5(io_lib().afid_h5_write_field3d(os.fsencode(path), _dptr(a), nx, ny, nz), "afid_h5_write_field3d " + str(path))


def h5_read_field(path, shape):
    """HdfReadContinua (HdfReadContinua.F90:11-107): the [nz][ny][nx] array of
    `path`; raises if the file's dims differ from `shape`."""
    a = np.empty(shape, dtype=np.float64)
    nz, ny, nx = shape
    _h5(io_lib().afid_h5_read_field3d(os.fsencode(path), _dptr(a), nx, ny, nz), "afid_h5_read_field3d " + str(path))
    return a


def h5_dims(path, name="var"):
    """Fortran-order dims (fastest first) of dataset `name`."""
    d = (C.c_int64 * 8)()
    r = io_lib().afid_h5_dataset_dims(os.fsencode(path), name.encode(), d, 8)
    if r < 0:
        raise H5Error(f"afid_h5_dataset_dims {name} failed ({r}): {io_lib().afid_h5_last_error().decode()}")
    return tuple(d[i] for i in range(r))


def h5_write_scalar(path, name, value):
    """HdfSerialWriteIntScalar / HdfSerialWriteRealScalar (HdfRoutines.F90:25-96):
    Python ints are written as 32-bit integers, floats as doubles."""
    L = io_lib()
    if isinstance(value, (int, np.integer)):
        rc = L.afid_h5_write_int_scalar(os.fsencode(path), name.encode(), int(value))
    else:
        rc = L.afid_h5_write_real_scalar(os.fsencode(path), name.encode(), float(value))
    _h5(rc, f"writing scalar {name}")


def h5_read_int(path, name):
    v = C.c_int32()
    _h5(io_lib().afid_h5_read_int_scalar(os.fsencode(path), name.encode(), C.byref(v)), f"reading {name}")
    return int(v.value)


def h5_read_real(path, name):
    v = C.c_double()
    _h5(io_lib().afid_h5_read_real_scalar(os.fsencode(path), name.encode(), C.byref(v)), f"reading {name}")
    return float(v.value)


def h5_write_1d(path, name, arr):
    """HdfSerialWriteReal1D (HdfRoutines.F90:139-177)."""
    a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
    _h5(io_lib().afid_h5_write_real1d(os.fsencode(path), name.encode(), _dptr(a), a.size), f"writing {name}")


def h5_read_1d(path, name, n):
    a = np.empty(n, dtype=np.float64)
    _h5(io_lib().afid_h5_read_real1d(os.fsencode(path), name.encode(), _dptr(a), n), f"reading {name}")
    return a


def fortran_nint(x):
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(np.sign(x) * np.floor(abs(x) + 0.5))


class AfidError(RuntimeError):
    pass


class Solver:
    """One device context: the module state of the reference on the GPU."""

    def __init__(self, cfg: dict):
        self.cfgd = dict(DEFAULTS)
        self.cfgd.update(cfg)
        self.cfg = make_config(self.cfgd)
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.afid_gpu_init(C.byref(self.cfg), C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.L.afid_last_error(h).decode() if h else "no context"
            if h:
                self.L.afid_gpu_finalize(h)
            self.h = None
            raise AfidError(f"afid_gpu_init failed ({rc}): {msg}")

    # -- lifetime
    def close(self):
        if getattr(self, "h", None):
            self.L.afid_gpu_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise AfidError(f"{what} failed ({rc}): {self.L.afid_last_error(self.h).decode()}")

    # -- shapes
    def shape(self, name):
        c = self.cfgd
        ref = name in REFINED
        ny, nz = (c["nymr"], c["nzmr"]) if ref else (c["nym"], c["nzm"])
        if name in PLANES:
            return (nz, ny, 1)
        nx = (c["nxmr"] + 1) if ref else (c["nxm"] + 1)
        return (nz, ny, nx)

    # -- data movement
    def upload(self, name, arr, lbound=(1, 1, 1)):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        ext = (C.c_int64 * 3)(a.shape[2], a.shape[1], a.shape[0])
        lb = (C.c_int64 * 3)(*lbound)
        self._check(self.L.afid_upload(self.h, FIELD_ID[name], a.ctypes.data_as(C.POINTER(C.c_double)), lb, ext),
                    "afid_upload " + name)

    def download(self, name, shape=None, lbound=(1, 1, 1)):
        shape = shape or self.shape(name)
        a = np.zeros(shape, dtype=np.float64)
        ext = (C.c_int64 * 3)(shape[2], shape[1], shape[0])
        lb = (C.c_int64 * 3)(*lbound)
        self._check(self.L.afid_download(self.h, FIELD_ID[name], a.ctypes.data_as(C.POINTER(C.c_double)), lb, ext),
                    "afid_download " + name)
        return a

    # -- the TimeMarcher call tree
    def set_substep(self, dt, ns):
        self._check(self.L.afid_set_substep(self.h, dt, ns), "afid_set_substep")

    def call(self, routine):
        self._check(getattr(self.L, "afid_" + routine)(self.h), "afid_" + routine)

    def time_marcher(self, dt):
        self._check(self.L.afid_time_marcher(self.h, dt), "afid_time_marcher")

    def time_marcher_graph(self, dt, nsteps):
        """nsteps RK3 steps replayed from a captured HIP graph."""
        self._check(self.L.afid_time_marcher_graph(self.h, dt, nsteps), "afid_time_marcher_graph")

    def create_initial_conditions(self, seed, eps=1e-3):
        self._check(self.L.afid_create_initial_conditions(self.h, seed, eps), "afid_create_initial_conditions")

    def synchronize(self):
        self._check(self.L.afid_synchronize(self.h), "afid_synchronize")

    # -- reductions
    def calc_max_cfl(self):
        a, b = C.c_double(), C.c_double()
        self._check(self.L.afid_calc_max_cfl(self.h, C.byref(a), C.byref(b)), "afid_calc_max_cfl")
        return a.value, b.value

    def check_divergence(self):
        a, b = C.c_double(), C.c_double()
        self._check(self.L.afid_check_divergence(self.h, C.byref(a), C.byref(b)), "afid_check_divergence")
        return a.value, b.value

    def global_quantities(self):
        out = (C.c_double * 4)()
        self._check(self.L.afid_global_quantities(self.h, out), "afid_global_quantities")
        return dict(zip(["nu", "ke", "vrms", "tmean"], list(out)))

    def mean_profiles(self):
        """CalcMeanProfiles (CalcMeanProfiles.F90:11-229): dict of the 15 profiles
        (each nxm values) in the reference's means.h5 order."""
        nxm = int(self.cfg.nxm)
        out = np.zeros(len(MEAN_PROFILES) * nxm)
        self._check(self.L.afid_calc_mean_profiles(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), "afid_calc_mean_profiles")
        return {n: out[q * nxm:(q + 1) * nxm].copy() for q, n in enumerate(MEAN_PROFILES)}

    def moist_stats(self):
        """CalcMoistStats (moisture.F90:422-505): dict of the 7 profiles (nxm
        values each) in the reference's means.h5 order."""
        n = int(self.cfg.nxm)
        out = np.zeros(len(MOIST_PROFILES) * n)
        self._check(self.L.afid_calc_moist_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), "afid_calc_moist_stats")
        return {m: out[q * n:(q + 1) * n].copy() for q, m in enumerate(MOIST_PROFILES)}

    def sal_stats(self):
        """CalcSalStats (salinity.F90:510-690): dict of the 6 refined profiles
        (each nxmr values) in the reference's means.h5 order."""
        n = int(self.cfg.nxmr)
        out = np.zeros(len(SAL_PROFILES) * n)
        self._check(self.L.afid_calc_sal_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), "afid_calc_sal_stats")
        return {m: out[q * n:(q + 1) * n].copy() for q, m in enumerate(SAL_PROFILES)}

    def phi_stats(self):
        """CalcPhiStats (phasefield.F90:750-798): dict of phibar, phirms (nxmr values each)."""
        n = int(self.cfg.nxmr)
        out = np.zeros(2 * n)
        self._check(self.L.afid_calc_phi_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_double))), "afid_calc_phi_stats")
        return {"phibar": out[:n].copy(), "phirms": out[n:].copy()}

    # -- files (libafid_io.so)
    def _grid_scalars(self):
        c = self.cfgd
        g = dict(nx=c["nxm"] + 1, ny=c["nym"] + 1, nz=c["nzm"] + 1, ylen=float(c["ylen"]), zlen=float(c["zlen"]),
                 istr3=int(c["istr3"]), str3=float(c["str3"]))
        if c["multires"]:
            g.update(nxr=c["nxmr"] + 1, nyr=c["nymr"] + 1, nzr=c["nzmr"] + 1, istr3r=int(c["istr3r"]))
        return g

    def _continua_fields(self, final):
        c = self.cfgd
        names = ["temp", "vx", "vy", "vz"]
        if c["salinity"]:
            names.append("sal")
        if c["phasefield"]:
            names.append("phi")
        if c["moist"]:
            names.append("humid")            # WriteFlowField.F90:34-38
        if final:
            names.append("pr")
        return names

    def write_flow_field(self, outdir, time, final=True, save_3D=None):
        """WriteFlowField (WriteFlowField.F90:11-86): outdir/continua_<field>.h5
        for temp, vx, vy, vz (+ sal, phi on the refined grid; + pr when final)
        and, when final, outdir/continua_master.h5 with the grid and the time.
        Not final: outdir/fields/<nint(time/save_3D) as 5 digits>_<field>.h5,
        without pr and master.  Returns the list of files written."""
        os.makedirs(outdir, exist_ok=True)      # main.F90:77 creates outputdir at start-up
        if final:
            base = os.path.join(outdir, "continua")
        else:
            if not save_3D:
                raise ValueError("write_flow_field(final=False) needs save_3D")
            os.makedirs(os.path.join(outdir, "fields"), exist_ok=True)
            base = os.path.join(outdir, "fields", "%05d" % fortran_nint(time / save_3D))
        out = []
        for n in self._continua_fields(final):
            path = f"{base}_{CONTINUA[n]}.h5"
            h5_write_field(path, self.download(n))
            out.append(path)
        if final:
            path = os.path.join(outdir, "continua_master.h5")
            h5_create_file(path)
            g = self._grid_scalars()
            for k in ["nx", "ny", "nz", "ylen", "zlen"]:
                h5_write_scalar(path, k, g[k])
            h5_write_scalar(path, "time", float(time))
            h5_write_scalar(path, "istr3", g["istr3"])
            h5_write_scalar(path, "str3", g["str3"])
            if self.cfgd["multires"]:
                for k in ["nxr", "nyr", "nzr", "istr3r"]:
                    h5_write_scalar(path, k, g[k])
            out.append(path)
        return out

    def set_ibm(self, field, mask, dist=None):
        """ibm_param.F90: the immersed-boundary mask (nz, ny, nx) of the solve
        of `field` (vx, vy, vz, temp, sal) and its boundary distances in the
        reference's loop order; mask None switches the solve back."""
        if mask is None:
            self.L.afid_set_ibm.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_int64]
            self._check(self.L.afid_set_ibm(self.h, FIELD_ID[field], None, None, 0), "afid_set_ibm")
            return
        m = np.ascontiguousarray(mask, dtype=np.int32)
        d = np.ascontiguousarray(dist if dist is not None else np.zeros(1), dtype=np.float64)
        self.L.afid_set_ibm.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.c_int64]
        self._check(self.L.afid_set_ibm(self.h, FIELD_ID[field], m.ctypes.data_as(C.POINTER(C.c_int32)),
                                        d.ctypes.data_as(C.POINTER(C.c_double)), 0 if dist is None else d.size),
                    "afid_set_ibm")

    def interp_input(self, old, fields):
        """The interpolation of ReadFlowInterp.F90:90-141 onto this context's
        grids: `old` a dict of the old grid (nx, ny, nz, istr3, str3, ylen,
        zlen, and nxr, nyr, nzr, istr3r with multires; sizes count the faces)
        and `fields` the old arrays (nz-1, ny-1, nx) by name: vx, vy, vz, temp
        (InterpInputVel), sal / phi (InterpInputSal / InterpInputPhi) and humid
        (InterpInputHum, after set_humidity_bcs)."""
        og = OldGrid(int(old["nx"]), int(old["ny"]), int(old["nz"]), int(old["istr3"]), float(old["str3"]),
                     float(old.get("ylen", self.cfgd["ylen"])), float(old.get("zlen", self.cfgd["zlen"])),
                     int(old.get("nxr", 0)), int(old.get("nyr", 0)), int(old.get("nzr", 0)), int(old.get("istr3r", 0)))
        a = [np.ascontiguousarray(fields[n], dtype=np.float64) for n in ("vx", "vy", "vz", "temp")]
        dpp = [x.ctypes.data_as(C.POINTER(C.c_double)) for x in a]
        self._check(self.L.afid_interp_input_vel(self.h, C.byref(og), *dpp), "afid_interp_input_vel")
        for n, fn in (("sal", self.L.afid_interp_input_sal), ("phi", self.L.afid_interp_input_phi),
                      ("humid", self.L.afid_interp_input_humid)):
            if n in fields:
                b = np.ascontiguousarray(fields[n], dtype=np.float64)
                self._check(fn(self.h, C.byref(og), b.ctypes.data_as(C.POINTER(C.c_double))), "afid_interp_input_" + n)

    def read_flow_field(self, outdir, interpolate=True):
        """ReadFlowInterp (ReadFlowInterp.F90:1-179): reads continua_master.h5
        and returns the checkpoint's time.  On the configured grid (:84-89) it
        uploads vx, vy, vz, temp (+ sal, phi; + pr if continua_pr.h5 exists,
        :143-171).  On another grid it reads the old fields with the old
        grid's dims and interpolates them (InterpInputVel / InterpInputSal /
        InterpInputPhi, :90-141); the domain lengths must match (:76-82).
        With `interpolate`, the start-up transfers that main.F90:205-213 runs
        after either initial state follow: InterpVelMgrd and InterpSalMultigrid
        (salinity), InterpTempMultigrid and InterpPhiMultigrid (phase field)."""
        master = os.path.join(outdir, "continua_master.h5")
        g = self._grid_scalars()
        old = {k: h5_read_int(master, k) for k in ["nx", "ny", "nz", "istr3"]}
        old["str3"] = h5_read_real(master, "str3")
        if self.cfgd["multires"]:
            old.update({k: h5_read_int(master, k) for k in ["nxr", "nyr", "nzr", "istr3r"]})
        t = h5_read_real(master, "time")
        # :76-81: the domain lengths must match, on the same grid as on a new one
        yleno, zleno = h5_read_real(master, "ylen"), h5_read_real(master, "zlen")
        if abs(g["ylen"] - yleno) > 1e-8 or abs(g["zlen"] - zleno) > 1e-8:
            raise AfidError(f"continua domain size does not match the configuration: old Ly, Lz = {yleno}, {zleno}; "
                            f"current {g['ylen']}, {g['zlen']}")
        same = all(old[k] == g[k] for k in old if k != "str3") and abs(g["str3"] - old["str3"]) <= 1e-8
        if not same:
            old["ylen"], old["zlen"] = yleno, zleno
            fields = {}
            for n in self._continua_fields(final=False):
                ref = n in REFINED
                dims = (old["nzr"] - 1, old["nyr"] - 1, old["nxr"]) if ref else (old["nz"] - 1, old["ny"] - 1, old["nx"])
                fields[n] = h5_read_field(os.path.join(outdir, f"continua_{CONTINUA[n]}.h5"), dims)
            self.interp_input(old, fields)
        else:
            for n in self._continua_fields(final=False):
                self.upload(n, h5_read_field(os.path.join(outdir, f"continua_{CONTINUA[n]}.h5"), self.shape(n)))
            pr = os.path.join(outdir, "continua_pr.h5")
            if os.path.exists(pr):
                self.upload("pr", h5_read_field(pr, self.shape("pr")))
        if interpolate:
            if self.cfgd["salinity"]:
                self.call("interp_vel_mgrd")
                self.call("interp_sal_multigrid")
            if self.cfgd["phasefield"]:
                self.call("interp_temp_multigrid")
                self.call("interp_phi_multigrid")
            if self.cfgd["moist"]:
                self.call("update_saturation")          # main.F90:202
        return t

    def write_means(self, outdir, time, tout):
        """The file part of CalcMeanProfiles (CalcMeanProfiles.F90:209-283),
        CalcSalStats (salinity.F90:588-603) and CalcPhiStats
        (phasefield.F90:786-793): outdir/means.h5 gets <profile>/<nint(time/tout)
        as 5 digits> for every profile; the file and its groups are created on
        the first call.  Returns the profiles written."""
        os.makedirs(outdir, exist_ok=True)
        path = os.path.join(outdir, "means.h5")
        c = self.cfgd
        groups = list(MEAN_PROFILES)
        if c["salinity"]:
            groups += SAL_PROFILES
        if c["phasefield"]:
            groups += PHI_PROFILES
        if c["moist"]:
            groups += MOIST_PROFILES
        if not os.path.exists(path):
            h5_create_file(path)
            for gname in groups:
                h5_create_group(path, gname)
        prof = self.mean_profiles()
        if c["salinity"]:
            prof.update(self.sal_stats())
        if c["phasefield"]:
            prof.update(self.phi_stats())
        if c["moist"]:
            prof.update(self.moist_stats())
        nstat = "%05d" % fortran_nint(time / tout)
        for gname in groups:
            h5_write_1d(path, f"{gname}/{nstat}", prof[gname])
        return prof

    def refinement_residuals(self):
        """Mixed form: max|div u| before each refinement sweep of the last substep."""
        out = (C.c_double * 4)()
        n = C.c_int()
        self._check(self.L.afid_refinement_residuals(self.h, out, 4, C.byref(n)), "afid_refinement_residuals")
        return list(out)[:n.value]

    # -- diagnostics
    def device_bytes(self):
        return int(self.L.afid_device_bytes(self.h))

    def set_timing(self, on=True):
        self._check(self.L.afid_set_timing(self.h, 1 if on else 0), "afid_set_timing")

    def get_timing(self):
        names = (C.c_char_p * 64)()
        ms = (C.c_double * 64)()
        n = C.c_int()
        self._check(self.L.afid_get_timing(self.h, names, ms, 64, C.byref(n)), "afid_get_timing")
        return {names[q].decode(): ms[q] for q in range(min(n.value, 64))}

    def grid_array(self, name):
        n = C.c_int64()
        self._check(self.L.afid_grid_array(self.h, name.encode(), None, 0, C.byref(n)), "afid_grid_array")
        out = np.zeros(n.value)
        self._check(self.L.afid_grid_array(self.h, name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)),
                                           n.value, C.byref(n)), "afid_grid_array")
        return out


class Ensemble:
    """Many independent 2-D runs (nym = 1) advanced together: each run is a
    bou.in (config dict) plus its time-stepping control (RunControl fields).
    Mirrors the reference's driver loop (main.F90:281-422) per run."""

    def __init__(self, cfgs, controls):
        assert len(cfgs) == len(controls) and cfgs
        self.cfgds = []
        for c in cfgs:
            d = dict(DEFAULTS)
            d.update(c)
            self.cfgds.append(d)
        self.n = len(cfgs)
        arr = (AfidConfig * self.n)(*[make_config(d) for d in self.cfgds])
        ctl = (RunControl * self.n)(*[make_control(k) for k in controls])
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.afid_ens_init(arr, ctl, self.n, C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.L.afid_ens_last_error(h).decode() if h else "no ensemble"
            if h:
                self.L.afid_ens_finalize(h)
            self.h = None
            raise AfidError(f"afid_ens_init failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.L.afid_ens_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise AfidError(f"{what} failed ({rc}): {self.L.afid_ens_last_error(self.h).decode()}")

    def shape(self, run, name):
        c = self.cfgds[run]
        if name in PLANES:
            return (c["nzm"], 1, 1)
        return (c["nzm"], 1, c["nxm"] + 1)

    def upload(self, run, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        ext = (C.c_int64 * 3)(a.shape[2], a.shape[1], a.shape[0])
        lb = (C.c_int64 * 3)(1, 1, 1)
        self._check(self.L.afid_ens_upload(self.h, run, FIELD_ID[name], a.ctypes.data_as(C.POINTER(C.c_double)), lb, ext),
                    f"afid_ens_upload {name} run {run}")

    def download(self, run, name):
        shape = self.shape(run, name)
        a = np.zeros(shape, dtype=np.float64)
        ext = (C.c_int64 * 3)(shape[2], shape[1], shape[0])
        lb = (C.c_int64 * 3)(1, 1, 1)
        self._check(self.L.afid_ens_download(self.h, run, FIELD_ID[name], a.ctypes.data_as(C.POINTER(C.c_double)), lb, ext),
                    f"afid_ens_download {name} run {run}")
        return a

    def create_initial_conditions(self, seed0, eps=1e-3):
        self._check(self.L.afid_ens_create_initial_conditions(self.h, seed0, eps), "afid_ens_create_initial_conditions")

    def reset(self):
        self._check(self.L.afid_ens_reset(self.h), "afid_ens_reset")

    def step(self, nsteps=1, graph=True):
        self._check(self.L.afid_ens_step(self.h, nsteps, 1 if graph else 0), "afid_ens_step")

    def step_timed(self, nsteps=1, graph=True):
        """Steps bracketed by HIP events on the ensemble's stream; returns ms."""
        ms = C.c_double()
        self._check(self.L.afid_ens_step_timed(self.h, nsteps, 1 if graph else 0, C.byref(ms)), "afid_ens_step_timed")
        return ms.value

    def synchronize(self):
        self._check(self.L.afid_ens_synchronize(self.h), "afid_ens_synchronize")

    def run_state(self, run):
        s = RunState()
        self._check(self.L.afid_ens_run_state(self.h, run, C.byref(s)), "afid_ens_run_state")
        return s.as_dict()

    def global_quantities(self, run):
        out = (C.c_double * 4)()
        self._check(self.L.afid_ens_global_quantities(self.h, run, out), "afid_ens_global_quantities")
        return dict(zip(["nu", "ke", "vrms", "tmean"], list(out)))

    def autotune(self, steps=2):
        best = EnsTuning()
        log = (EnsTuning * 64)()
        n = C.c_int32()
        self._check(self.L.afid_ens_autotune(self.h, steps, C.byref(best), log, 64, C.byref(n)), "afid_ens_autotune")
        return best.as_dict(), [log[q].as_dict() for q in range(min(n.value, 64))]

    def set_tuning(self, threads=1024, workgroups=0, lines=8, phased=0):
        t = EnsTuning(threads, workgroups, lines, phased, 0.0)
        self._check(self.L.afid_ens_set_tuning(self.h, C.byref(t)), "afid_ens_set_tuning")

    def get_tuning(self):
        t = EnsTuning()
        self._check(self.L.afid_ens_get_tuning(self.h, C.byref(t)), "afid_ens_get_tuning")
        return t.as_dict()

    def device_bytes(self):
        return int(self.L.afid_ens_device_bytes(self.h))
