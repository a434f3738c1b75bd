"""Build step for the HIP library libafid_gpu.so (gfx950 only), the host-only
HDF5 library libafid_io.so and the Fortran host.

Used by __graft_entry__.build(); can also be run directly:
    python afid-murphfi_amd/build_lib.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libafid_gpu.so")
SRC = os.path.join(HERE, "csrc", "afid_gpu.hip")


def hipcc_cmd(out=LIB, extra=()):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    return [hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function",
            "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
            *extra, "-o", out, SRC]


IO_LIB = os.path.join(HERE, "libafid_io.so")
IO_SRC = os.path.join(HERE, "csrc", "afid_io.c")
# serial HDF5 C library of the image (the reference needs HDF5 for its files)
HDF5_PREFIX = os.environ.get("AFID_HDF5_PREFIX", "/opt/conda")


def build_io(verbose=True):
    """Host-only HDF5 library libafid_io.so (checkpoints, means.h5)."""
    cc = os.environ.get("CC", "gcc")
    inc, libdir = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    cmd = [cc, "-O2", "-std=c11", "-Wall", "-Wextra", "-fPIC", "-shared",
           "-I", os.path.join(ROOT, "include"), "-I", inc, "-o", IO_LIB + ".tmp", IO_SRC,
           "-L" + libdir, "-lhdf5", "-Wl,-rpath," + libdir]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("building libafid_io.so failed (HDF5 under %s)" % HDF5_PREFIX)
    os.replace(IO_LIB + ".tmp", IO_LIB)
    return IO_LIB


FORTRAN_DIR = os.path.join(HERE, "fortran")
HOST_EXE = os.path.join(HERE, "afid_host")


def build_fortran_host(verbose=True):
    """Fortran host (driver loop + TimeMarcher call tree) over the C ABI."""
    import tempfile
    fc = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    with tempfile.TemporaryDirectory() as tmp:
        cmd = [fc, "-O2", "-module-dir", tmp, "-o", HOST_EXE + ".tmp",
               os.path.join(FORTRAN_DIR, "afid_gpu_mod.F90"), os.path.join(FORTRAN_DIR, "afid_io_mod.F90"),
               os.path.join(FORTRAN_DIR, "afid_host.F90"),
               "-L" + HERE, "-lafid_gpu", "-lafid_io", "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("amdflang failed building the Fortran host")
    os.replace(HOST_EXE + ".tmp", HOST_EXE)
    return HOST_EXE


def build(verbose=True):
    cmd = hipcc_cmd(out=LIB + ".tmp")
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libafid_gpu.so")
    if verbose and r.stderr.strip():
        # only the harmless packed-fp32 host-pass note is expected here
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
    os.replace(LIB + ".tmp", LIB)
    build_io(verbose)
    if os.path.exists("/opt/rocm/bin/amdflang") or os.environ.get("FC"):
        build_fortran_host(verbose)
    return LIB


if __name__ == "__main__":
    build()
