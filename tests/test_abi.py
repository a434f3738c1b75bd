"""CPU checks of the C ABI: the built library loads and exports every entry
point declared in include/afid_gpu.h (no compute calls: no GPU here)."""
import os
import re
import subprocess

import pytest

import pyafid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "afid_gpu.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*\*?\s*(afid_\w+)\s*\(", txt, re.M)))


def test_header_and_binding_agree():
    assert declared() == sorted(pyafid.EXPORTS)


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
