"""The HDF5 files of libafid_io.so (include/afid_io.h) and the host logic of
WriteFlowField / ReadFlowInterp / the means.h5 writes in pyafid, on the CPU.

The on-disk layout is pinned without this library's own reader: the files are
contiguous, unfiltered HDF5, so the raw bytes of a [nz][ny][nx] array (x
fastest, the reference's Fortran order (nx, ny, nz)) must appear verbatim in
the file, and the bytes of the transposed array must not.  The checkpoint
names, the master scalars and the means.h5 dataset names follow
WriteFlowField.F90:11-86, HdfReadContinua.F90:44-61 and
CalcMeanProfiles.F90:209-283.  No GPU is used: the solver-level tests drive
the host methods through a stand-in whose upload/download keep host arrays.
"""
import os
import re

import numpy as np
import pytest

import pyafid
from helpers import uniform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDF5_SIGNATURE = b"\x89HDF\r\n\x1a\n"
pytestmark = pytest.mark.skipif(not os.path.exists(pyafid.IO_LIB_PATH), reason="libafid_io.so not built")


def header_functions():
    txt = open(os.path.join(ROOT, "include", "afid_io.h")).read()
    return sorted(set(re.findall(r"\b(afid_h5_\w+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    L = pyafid.io_lib()
    names = header_functions()
    assert names == sorted(pyafid.IO_EXPORTS)
    for n in names:
        assert hasattr(L, n), n


def test_field_roundtrip_and_layout(tmp_path):
    nx, ny, nz = 7, 6, 5
    a = uniform(3, (nz, ny, nx)) - 0.5
    p = tmp_path / "continua_vx.h5"
    pyafid.h5_write_field(p, a)
    raw = p.read_bytes()
    assert raw.startswith(HDF5_SIGNATURE)
    # the array's bytes in C order [nz][ny][nx] (x fastest) are stored verbatim
    assert raw.find(a.tobytes()) > 0
    assert raw.find(np.ascontiguousarray(a.transpose(2, 1, 0)).tobytes()) < 0
    assert pyafid.h5_dims(p) == (nx, ny, nz)          # Fortran order, as the reference declares it
    b = pyafid.h5_read_field(p, (nz, ny, nx))
    assert np.array_equal(a, b)
    # rewriting truncates: a smaller field replaces the old one
    c = uniform(4, (2, 3, 4))
    pyafid.h5_write_field(p, c)
    assert pyafid.h5_dims(p) == (4, 3, 2)
    assert np.array_equal(pyafid.h5_read_field(p, (2, 3, 4)), c)


def test_field_special_values_roundtrip(tmp_path):
    a = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, 1.7976931348623157e308, -1e-310],
                 dtype=np.float64).reshape(2, 2, 2)
    p = tmp_path / "f.h5"
    pyafid.h5_write_field(p, a)
    b = pyafid.h5_read_field(p, (2, 2, 2))
    assert a.tobytes() == b.tobytes()


def test_field_errors(tmp_path):
    p = tmp_path / "f.h5"
    pyafid.h5_write_field(p, np.zeros((3, 4, 5)))
    with pytest.raises(pyafid.H5Error, match="dim"):
        pyafid.h5_read_field(p, (3, 4, 6))
    with pytest.raises(pyafid.H5Error, match="cannot open"):
        pyafid.h5_read_field(tmp_path / "missing.h5", (3, 4, 5))
    with pytest.raises(pyafid.H5Error, match="no dataset"):
        pyafid.h5_dims(p, "nothere")
    # 1-D datasets have rank 1, a 3-D read of them fails on the rank
    m = tmp_path / "m.h5"
    pyafid.h5_create_file(m)
    pyafid.h5_write_1d(m, "var", np.arange(4.0))
    with pytest.raises(pyafid.H5Error, match="rank"):
        pyafid.h5_read_field(m, (1, 1, 4))
    with pytest.raises(ValueError):
        pyafid.h5_write_field(p, np.zeros((2, 2)))


def test_master_scalars(tmp_path):
    p = tmp_path / "continua_master.h5"
    pyafid.h5_create_file(p)
    pyafid.h5_write_scalar(p, "nx", 385)
    pyafid.h5_write_scalar(p, "time", 12.5)
    pyafid.h5_write_scalar(p, "str3", 0.0)
    assert pyafid.h5_read_int(p, "nx") == 385
    assert pyafid.h5_read_real(p, "time") == 12.5
    assert pyafid.h5_dims(p, "nx") == (1,)            # a one-element 1-D dataset (HdfRoutines.F90:75-88)
    raw = p.read_bytes()
    assert raw.find(np.int32(385).tobytes()) > 0
    assert raw.find(np.float64(12.5).tobytes()) > 0
    # a second write of the same name replaces the dataset
    pyafid.h5_write_scalar(p, "time", 13.25)
    assert pyafid.h5_read_real(p, "time") == 13.25
    with pytest.raises(pyafid.H5Error):
        pyafid.h5_read_int(p, "ny")
    with pytest.raises(pyafid.H5Error, match="cannot open"):
        pyafid.h5_write_scalar(tmp_path / "none.h5", "nx", 1)


def test_means_groups_and_datasets(tmp_path):
    p = tmp_path / "means.h5"
    pyafid.h5_create_file(p)
    for g in pyafid.MEAN_PROFILES:
        pyafid.h5_create_group(p, g)
    pyafid.h5_create_group(p, "Tbar")                 # an existing group is kept
    a, b = uniform(1, (9,)), uniform(2, (9,))
    pyafid.h5_write_1d(p, "Tbar/00001", a)
    pyafid.h5_write_1d(p, "Tbar/00002", b)
    pyafid.h5_write_1d(p, "Tbar/00001", b)            # replaced, as HdfSerialWriteReal1D deletes first
    assert np.array_equal(pyafid.h5_read_1d(p, "Tbar/00001", 9), b)
    assert np.array_equal(pyafid.h5_read_1d(p, "Tbar/00002", 9), b)
    with pytest.raises(pyafid.H5Error):
        pyafid.h5_write_1d(p, "nogroup/00001", a)      # the group must exist (created up front in the reference)
    with pytest.raises(pyafid.H5Error, match="dim"):
        pyafid.h5_read_1d(p, "Tbar/00001", 8)


def test_fortran_nint():
    assert [pyafid.fortran_nint(x) for x in (2.5, -2.5, 0.4999, 2.9999999, 3.0000001, -0.5)] == [3, -3, 0, 3, 3, -1]


class HostSolver(pyafid.Solver):
    """Stand-in for a device context: the module arrays live in a dict."""

    def __init__(self, cfg):
        self.cfgd = dict(pyafid.DEFAULTS)
        self.cfgd.update(cfg)
        self.cfg = pyafid.make_config(self.cfgd)
        self.arrays = {}
        self.calls = []
        self.h = None

    def call(self, routine):
        self.calls.append(routine)

    def interp_input(self, old, fields):
        self.calls.append("interp_input")
        self.interp_args = (dict(old), {n: np.array(a) for n, a in fields.items()})

    def close(self):
        pass

    def upload(self, name, arr, lbound=(1, 1, 1)):
        assert tuple(arr.shape) == self.shape(name), name
        self.arrays[name] = np.array(arr, dtype=np.float64)

    def download(self, name, shape=None, lbound=(1, 1, 1)):
        return self.arrays[name].copy()

    def mean_profiles(self):
        n = self.cfgd["nxm"]
        return {m: uniform(50 + q, (n,)) for q, m in enumerate(pyafid.MEAN_PROFILES)}

    def sal_stats(self):
        n = self.cfgd["nxmr"]
        return {m: uniform(80 + q, (n,)) for q, m in enumerate(pyafid.SAL_PROFILES)}

    def phi_stats(self):
        n = self.cfgd["nxmr"]
        return {m: uniform(90 + q, (n,)) for q, m in enumerate(pyafid.PHI_PROFILES)}


MR = dict(nxm=8, nym=6, nzm=4, istr3=6, str3=1.2, multires=1, salinity=1, nxmr=16, nymr=12, nzmr=8, istr3r=4,
          ylen=2.0, zlen=1.5)


def filled(cfg, seed):
    s = HostSolver(cfg)
    for q, n in enumerate(["vx", "vy", "vz", "temp", "pr"] + (["sal"] if cfg.get("salinity") else [])
                          + (["phi"] if cfg.get("phasefield") else [])):
        s.upload(n, uniform(seed * 100 + q, s.shape(n)))
    return s


def test_write_read_flow_field_host(tmp_path):
    a = filled(MR, 1)
    files = a.write_flow_field(tmp_path, time=0.75, final=True)
    names = sorted(os.path.basename(f) for f in files)
    assert names == sorted(["continua_temp.h5", "continua_vx.h5", "continua_vy.h5", "continua_vz.h5",
                            "continua_sal.h5", "continua_pr.h5", "continua_master.h5"])
    m = tmp_path / "continua_master.h5"
    assert [pyafid.h5_read_int(m, k) for k in ("nx", "ny", "nz", "istr3")] == [9, 7, 5, 6]
    assert [pyafid.h5_read_int(m, k) for k in ("nxr", "nyr", "nzr", "istr3r")] == [17, 13, 9, 4]
    assert [pyafid.h5_read_real(m, k) for k in ("ylen", "zlen", "time", "str3")] == [2.0, 1.5, 0.75, 1.2]
    # the checkpoint's field files have the reference's Fortran dims
    assert pyafid.h5_dims(tmp_path / "continua_vx.h5") == (9, 6, 4)
    assert pyafid.h5_dims(tmp_path / "continua_sal.h5") == (17, 12, 8)
    b = filled(MR, 2)
    t = b.read_flow_field(tmp_path)
    assert t == 0.75
    for n in ("vx", "vy", "vz", "temp", "pr", "sal"):
        assert np.array_equal(a.arrays[n], b.arrays[n]), n
    # then the start-up transfers of main.F90:205-209
    assert b.calls == ["interp_vel_mgrd", "interp_sal_multigrid"]
    c = filled(dict(MR, salinity=0, phasefield=1), 3)
    os.makedirs(tmp_path / "pf")
    c.write_flow_field(tmp_path / "pf", time=0.5)
    d = filled(dict(MR, salinity=0, phasefield=1), 4)
    d.read_flow_field(tmp_path / "pf")
    assert d.calls == ["interp_temp_multigrid", "interp_phi_multigrid"]
    assert np.array_equal(c.arrays["phi"], d.arrays["phi"])


def test_read_flow_field_without_pressure_keeps_it(tmp_path):
    a = filled(MR, 1)
    a.write_flow_field(tmp_path, time=1.0, final=True)
    os.remove(tmp_path / "continua_pr.h5")
    b = filled(MR, 2)
    pr = b.arrays["pr"].copy()
    b.read_flow_field(tmp_path)
    assert np.array_equal(b.arrays["pr"], pr)       # ReadFlowInterp.F90:163-168 reads pr only if present
    assert np.array_equal(b.arrays["vx"], a.arrays["vx"])


@pytest.mark.parametrize("change", [dict(nym=8), dict(istr3=4), dict(str3=1.3), dict(nxmr=20), dict(istr3r=6)])
def test_read_flow_field_another_grid_interpolates(tmp_path, change):
    """A checkpoint of another grid (ReadFlowInterp.F90:84-89) is read with
    the old grid's dims and handed to the interpolation (:90-141) with the
    old grid of the master file; pr is not read; the start-up transfers
    follow."""
    a = filled(MR, 1)
    a.write_flow_field(tmp_path, time=1.0, final=True)
    b = filled(dict(MR, **change), 2)
    t = b.read_flow_field(tmp_path)
    assert t == 1.0
    assert b.calls == ["interp_input", "interp_vel_mgrd", "interp_sal_multigrid"]
    old, fields = b.interp_args
    assert old == dict(nx=9, ny=7, nz=5, istr3=6, str3=1.2, nxr=17, nyr=13, nzr=9, istr3r=4, ylen=2.0, zlen=1.5)
    assert sorted(fields) == ["sal", "temp", "vx", "vy", "vz"]
    for n in fields:
        assert np.array_equal(fields[n], a.arrays[n]), n


@pytest.mark.parametrize("change", [dict(ylen=2.5), dict(zlen=1.0)])
def test_read_flow_field_rejects_another_domain(tmp_path, change):
    """ReadFlowInterp.F90:76-81 compares the domain lengths before anything
    else: another Ly or Lz stops the restart on the same grid as on a new
    one, and nothing is uploaded or interpolated."""
    a = filled(MR, 1)
    a.write_flow_field(tmp_path, time=1.0, final=True)
    for grid in ({}, dict(nym=8)):
        b = filled(dict(MR, **change, **grid), 2)
        before = {n: v.copy() for n, v in b.arrays.items()}
        with pytest.raises(pyafid.AfidError, match="domain size"):
            b.read_flow_field(tmp_path)
        assert b.calls == []
        for n, v in before.items():
            assert np.array_equal(b.arrays[n], v), n


def test_read_flow_field_str3_tolerance(tmp_path):
    a = filled(MR, 1)
    a.write_flow_field(tmp_path, time=1.0, final=True)
    b = filled(dict(MR, str3=1.2 + 5e-9), 2)       # |str3 - str3o| <= 1e-8 is the same grid (ReadFlowInterp.F90:85)
    b.read_flow_field(tmp_path)
    assert np.array_equal(b.arrays["temp"], a.arrays["temp"])


def test_write_flow_field_creates_the_output_folder(tmp_path):
    a = filled(MR, 1)
    a.write_flow_field(tmp_path / "outputdir", time=0.5)
    a.write_means(tmp_path / "other", time=0.5, tout=0.5)
    assert (tmp_path / "outputdir" / "continua_master.h5").exists()
    assert (tmp_path / "other" / "means.h5").exists()


def test_write_flow_field_snapshot_names(tmp_path):
    a = filled(dict(MR, salinity=0, phasefield=1), 1)
    files = a.write_flow_field(tmp_path, time=0.3, final=False, save_3D=0.1)
    names = sorted(os.path.relpath(f, tmp_path) for f in files)
    assert names == sorted(os.path.join("fields", f"00003_{n}.h5") for n in ("temp", "vx", "vy", "vz", "phi"))
    with pytest.raises(ValueError):
        a.write_flow_field(tmp_path, time=0.3, final=False)


def test_write_means_host(tmp_path):
    s = filled(dict(MR, phasefield=1), 1)
    p1 = s.write_means(tmp_path, time=0.2, tout=0.1)
    p2 = s.write_means(tmp_path, time=0.3, tout=0.1)
    f = tmp_path / "means.h5"
    for name in pyafid.MEAN_PROFILES:
        assert np.array_equal(pyafid.h5_read_1d(f, f"{name}/00002", MR["nxm"]), p1[name])
        assert np.array_equal(pyafid.h5_read_1d(f, f"{name}/00003", MR["nxm"]), p2[name])
    for name in pyafid.SAL_PROFILES + pyafid.PHI_PROFILES:
        assert pyafid.h5_dims(f, f"{name}/00003") == (MR["nxmr"],)


def test_fortran_binding_covers_the_header():
    src = open(os.path.join(ROOT, "afid-murphfi_amd", "fortran", "afid_io_mod.F90")).read()
    bound = sorted(set(re.findall(r"bind\(C, name='(afid_h5_\w+)'\)", src)))
    assert bound == header_functions()
