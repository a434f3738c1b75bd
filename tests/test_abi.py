"""CPU checks of the C ABI: the built library loads and exports every entry
point declared in include/afid_gpu.h and include/afid_ensemble.h (no
compute calls: no GPU here)."""
import os
import re
import subprocess

import pytest

import pyafid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "afid_gpu.h")
ENS_HEADER = os.path.join(ROOT, "include", "afid_ensemble.h")


def declared(header=None):
    files = [header] if header else [HEADER, ENS_HEADER]
    txt = "".join(open(f).read() for f in files)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*\*?\s*(afid_\w+)\s*\(", txt, re.M)))


def test_header_and_binding_agree():
    assert declared(HEADER) == sorted(pyafid.EXPORTS)
    assert declared(ENS_HEADER) == sorted(pyafid.ENS_EXPORTS)


@pytest.mark.skipif(not os.path.exists(pyafid.LIB_PATH), reason="library not built")
def test_library_exports_every_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", pyafid.LIB_PATH], capture_output=True, text=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = [s for s in declared() if s not in syms]
    assert not missing, missing
    L = pyafid.lib()
    for s in declared():
        assert hasattr(L, s)


def test_config_struct_layout():
    # the C struct and the ctypes mirror have the same field order
    txt = open(HEADER).read()
    body = txt[txt.index("typedef struct afid_config"):txt.index("} afid_config;")]
    names = re.findall(r"\b(?:int32_t|double)\s+([^;]+);", body)
    flat = [n.strip() for group in names for n in group.split(",")]
    assert flat == [f for f, _ in pyafid.AfidConfig._fields_]


def _struct_fields(header, name):
    txt = open(header).read()
    body = txt[txt.index("typedef struct " + name):txt.index("} " + name + ";")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"\b(?:int32_t|double)\s+([^;]+);", body)
    return [n.strip() for group in names for n in group.split(",")]


def test_ensemble_struct_layouts():
    assert _struct_fields(ENS_HEADER, "afid_run_control") == [f for f, _ in pyafid.RunControl._fields_]
    assert _struct_fields(ENS_HEADER, "afid_ens_tuning") == [f for f, _ in pyafid.EnsTuning._fields_]
    assert _struct_fields(ENS_HEADER, "afid_run_state") == [f for f, _ in pyafid.RunState._fields_]


def test_fortran_config_type_mirrors_the_struct():
    """type(afid_config) of the Fortran binding lists the C struct's members in
    the same order (afid_gpu_init reads sizeof(afid_config) bytes from it)."""
    src = open(os.path.join(ROOT, "afid-murphfi_amd", "fortran", "afid_gpu_mod.F90")).read()
    start = src.index("type, bind(C) :: afid_config")
    body = src[src.index("\n", start):src.index("end type afid_config")]
    body = re.sub(r"!.*", "", body)
    names = re.findall(r"::\s*([^\n]+)", body)
    flat = [n.strip() for group in names for n in group.split(",")]
    assert flat == [f for f, _ in pyafid.AfidConfig._fields_]
