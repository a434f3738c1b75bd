"""Shared helpers for the parity tests: seeded inputs and error norms."""
import numpy as np


def splitmix64(x):
    """Counter-based generator (splitmix64), vectorised over uint64 arrays."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return z ^ (z >> np.uint64(31))


def uniform(seed, shape):
    """U[0,1) from splitmix64(seed ^ linear index): identical on any host."""
    n = int(np.prod(shape))
    with np.errstate(over="ignore"):
        idx = np.arange(n, dtype=np.uint64) ^ np.uint64(seed)
        r = splitmix64(idx)
    return ((r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)).reshape(shape)


def relerr(a, b):
    """Max-norm difference relative to max|b|; complex inputs are compared as
    complex numbers (both parts)."""
    a, b = np.asarray(a), np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a, b = a.astype(dt), b.astype(dt)
    den = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / den)


BASE_FIELDS = ["vx", "vy", "vz", "temp", "pr", "rux", "ruy", "ruz", "rutemp"]
REF_FIELDS = ["sal", "rusal"]


def random_state(cfg, seed=1):
    """Random but admissible state: vx vanishes on the walls, T and S carry
    wall planes; returns dict of arrays shaped (nz, ny, nx)."""
    nx, ny, nz = cfg["nxm"] + 1, cfg["nym"], cfg["nzm"]
    st = {}
    for q, name in enumerate(BASE_FIELDS):
        a = uniform(seed * 1000 + q, (nz, ny, nx)) - 0.5
        a[..., nx - 1] = 0.0          # index nx of cell-centred arrays is unused
        st[name] = a
    st["vx"][..., 0] = 0.0
    st["vx"][..., nx - 1] = 0.0
    st["tempbp"] = np.full((nz, ny, 1), 1.0)
    st["temptp"] = np.full((nz, ny, 1), 0.0)
    if cfg.get("salinity"):
        nxr, nyr, nzr = cfg["nxmr"] + 1, cfg["nymr"], cfg["nzmr"]
        for q, name in enumerate(REF_FIELDS):
            a = uniform(seed * 1000 + 100 + q, (nzr, nyr, nxr)) - 0.5
            a[..., nxr - 1] = 0.0
            st[name] = a
        st["salbp"] = np.full((nzr, nyr, 1), 0.5)
        st["saltp"] = np.full((nzr, nyr, 1), -0.5)
    return st


def load_state(obj, st, upload):
    """Copy a state into an oracle (upload=False: write its arrays) or a solver."""
    for name, a in st.items():
        if upload:
            obj.upload(name, a)
        else:
            obj.field(name)[...] = a
