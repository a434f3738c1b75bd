"""afid_calc_mean_profiles_noalloc against afid_calc_mean_profiles.

The no-allocation entry point keeps its partial sums in the library's static
device array and writes the 15 profiles to a device buffer the caller owns.
It runs the same kernels in the same order as afid_calc_mean_profiles, so the
two agree bit for bit.  Two calls in a row on two contexts, the second on a
smaller grid (the array then holds the first call's partials beyond the
second's range), and a grid whose partials exceed the array is refused
before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

import pyafid
from helpers import random_state, load_state

pytestmark = pytest.mark.gpu

BIG = dict(nxm=96, nym=48, nzm=24, istr3=8, rayt=1e6, prat=1.0, inslws=0, inslwn=1, xplusu=0.3)
SMALL = dict(nxm=33, nym=12, nzm=8, istr3=6, str3=1.5, rayt=1e5, prat=7.0)


H2D, D2H = 1, 2


def hip():
    """The HIP runtime libafid_gpu.so itself links (already loaded, so the
    same instance): the test's own device buffers for the output."""
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return L


class DevBuf:
    def __init__(self, host):
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.L = hip()
        self.p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(self.p), self.host.nbytes) == 0
        self.put(self.host)

    def put(self, a):
        assert self.L.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.L.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.L.hipFree(self.p)


def noalloc_profiles(g):
    nxm = int(g.cfg.nxm)
    buf = DevBuf(np.full(len(pyafid.MEAN_PROFILES) * nxm, np.nan))
    rc = g.L.afid_calc_mean_profiles_noalloc(g.h, buf.p.value)
    assert rc == 0, g.L.afid_last_error(g.h)
    g.synchronize()
    a = buf.get()
    buf.free()
    return {n: a[q * nxm:(q + 1) * nxm] for q, n in enumerate(pyafid.MEAN_PROFILES)}


@pytest.mark.parametrize("mixed", [False, True])
def test_noalloc_matches_mean_profiles(mixed):
    prec = pyafid.MIXED if mixed else pyafid.FP64
    for cfg, seed in ((BIG, 3), (SMALL, 4)):            # the second call on a smaller grid
        g = pyafid.Solver(dict(cfg, precision=prec))
        load_state(g, random_state(cfg, seed), upload=True)
        ref = g.mean_profiles()
        got = noalloc_profiles(g)
        for n in pyafid.MEAN_PROFILES:
            assert np.all(np.isfinite(got[n])), n
            assert np.array_equal(got[n], ref[n]), (cfg["nxm"], n)
        g.close()


def test_noalloc_refuses_grids_beyond_the_scratch():
    """15 * 1000 * ceil(576 / 4) = 2.16M doubles > the 2^21 of the array:
    the call returns AFID_E_ARG (1) from its host-side size check and
    launches nothing; the output buffer keeps its contents."""
    cfg = dict(nxm=1000, nym=576, nzm=2, istr3=0, rayt=1e5, prat=1.0)
    assert 15 * 1000 * 144 > 2 * 1024 * 1024
    g = pyafid.Solver(cfg)
    buf = DevBuf(np.zeros(15 * 1000))
    rc = g.L.afid_calc_mean_profiles_noalloc(g.h, buf.p.value)
    assert rc == 1
    assert b"static scratch" in g.L.afid_last_error(g.h)
    g.synchronize()
    assert not np.any(buf.get())
    buf.free()
    # the allocating entry point still serves this grid
    assert np.all(np.isfinite(g.mean_profiles()["Tbar"]))
    g.close()
