#!/usr/bin/env python3
"""Benchmark of the RK3 step (TimeMarcher.F90:13-165) on one MI355X.

Workload (the largest one-GPU configuration): multi-resolution Rayleigh-Benard
convection with a passive scalar, base grid 384 (x, cosine-clustered) x 768 x
768, refined grid 768 x 1536 x 1536, Ly = Lz = 2, Ra = 1e9, Pr = 1, Sc = 16,
dt = 5e-5, u = 0 and T = 1-x, S = 1-x plus seeded 1e-3 noise.  Synthetic data
generated on the device.

Prints ONE JSON line: Mcell-steps/s = (base + refined cells) x steps / s,
with the HBM roofline of the dominant kernel group (HIP-event timing on the
library's stream inside the timed region) and a CPU baseline (the FP64 CPU
oracle, OpenMP, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "afid-murphfi_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = "Mcell-steps/s per RK3 step, mixed vs FP64 on GPU at stated div/Nu error; % HBM peak"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "rbc_multires_384x768x768_768x1536x1536": dict(
        nxm=384, nym=768, nzm=768, nsst=3, multires=1, nxmr=768, nymr=1536, nzmr=1536,
        salinity=1, alx3=1.0, ylen=2.0, zlen=2.0, istr3=8, istr3r=8, rayt=1e9, prat=1.0,
        rays=-1e9, pras=16.0, active_t=1, active_s=0, gaxis=1, dt=5e-5),
    "rbc_64x64x64": dict(nxm=64, nym=64, nzm=64, nsst=3, alx3=1.0, ylen=1.0, zlen=1.0, istr3=8,
                         rayt=1e5, prat=1.0, active_t=1, gaxis=1, dt=1e-3),
}
DEFAULT = "rbc_multires_384x768x768_768x1536x1536"


def cells(cfg):
    n = cfg["nxm"] * cfg["nym"] * cfg["nzm"]
    if cfg.get("salinity"):
        n += cfg["nxmr"] * cfg["nymr"] * cfg["nzmr"]
    return n


def kernel_bytes(cfg):
    """Algorithmic HBM bytes per launch of each kernel group (FP64, per
    substep): every array the operation must read or write once."""
    nb = cfg["nxm"] * cfg["nym"] * cfg["nzm"]
    nr = cfg["nxmr"] * cfg["nymr"] * cfg["nzmr"] if cfg.get("salinity") else 0
    ns = cfg["nxm"] * (cfg["nym"] // 2 + 1) * cfg["nzm"]       # complex spectral values
    d = 8
    return {
        # explicit terms + RHS + x-solve of vx, vy, vz, T in one sweep:
        # read u, v, w, T, p and the 4 old explicit terms, write u, v, w, T and the 4 new terms
        "column_base": 17 * d * nb,
        # salinity: read S, vxr, vyr, vzr, old H_S; write S, H_S
        "column_sal": 7 * d * nr,
        # read u,v,w; write complex half spectrum
        "poisson_fwd_y": 3 * d * nb + 16 * ns,
        "poisson_fft_z": 2 * 2 * 16 * ns,
        "poisson_tridiag": 2 * 16 * ns,
        "poisson_inv_y": 16 * ns + d * nb,
        # read phi, u,v,w,p; write u,v,w,p
        "project": 9 * d * nb,
        "mgrd_gradient": 3 * 2 * d * nb,
        "mgrd_vxr": d * nr + d * nb,
        "mgrd_vyr": d * nr + d * nb,
        "mgrd_vzr": d * nr + d * nb,
        # read S, write salc
        "mgrd_interp_sal": d * nr + d * nb,
    }


def cpu_baseline(args):
    """Time the FP64 CPU oracle on a bounded sample of the workload: the same
    wall-normal grids and grid spacings, periodic extent cut by 4 in y and z."""
    import pyoracle
    from helpers import uniform
    wl = dict(WORKLOADS[args.workload])
    dt = wl.pop("dt")
    f = args.cpu_shrink
    if wl.get("salinity"):
        wl.update(nym=wl["nym"] // f, nzm=wl["nzm"] // f, nymr=wl["nymr"] // f, nzmr=wl["nzmr"] // f,
                  ylen=wl["ylen"] / f, zlen=wl["zlen"] / f)
    else:
        wl.update(nym=max(wl["nym"] // f, 1), nzm=max(wl["nzm"] // f, 1), ylen=wl["ylen"] / f, zlen=wl["zlen"] / f)
    threads = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")), 16))
    pyoracle.set_threads(threads)
    o = pyoracle.Oracle(wl)
    nxm = wl["nxm"]
    xm = o.grid("xm")[:nxm]
    T = o.field("temp")
    o.field("tempbp")[...] = 1.0
    o.field("temptp")[...] = 0.0
    eps = 1e-3
    amp = np.where(np.abs(xm - 0.5) + eps > 0.5, 0.5 - np.abs(xm - 0.5), eps)
    r = uniform(args.seed, T.shape)[..., :nxm]
    T[..., :nxm] = 1.0 - xm + amp * (2 * r - 1)
    if wl.get("salinity"):
        nxmr = wl["nxmr"]
        xmr = o.grid("xmr")[:nxmr]
        S = o.field("sal")
        o.field("salbp")[...] = 1.0
        o.field("saltp")[...] = 0.0
        ampr = np.where(np.abs(xmr - 0.5) + eps > 0.5, 0.5 - np.abs(xmr - 0.5), eps)
        rr = uniform(args.seed + 1, S.shape)[..., :nxmr]
        S[..., :nxmr] = 1.0 - xmr + ampr * (2 * rr - 1)
        o.call("interp_sal_multigrid")
    times = []
    t_start = time.perf_counter()
    for _ in range(args.cpu_steps):
        t0 = time.perf_counter()
        o.time_marcher(dt)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() - t_start > args.cpu_budget:
            break
    o.close()
    n = cells(wl)
    best = min(times)
    return {"value": round(n / best / 1e6, 3), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": (f"FP64 CPU oracle (C, OpenMP {threads} threads), {len(times)} RK3 step(s) of the same "
                       f"workload with the periodic extent cut by {f} in y and z "
                       f"({wl['nxm']}x{wl['nym']}x{wl['nzm']}" +
                       (f" + {wl['nxmr']}x{wl['nymr']}x{wl['nzmr']}" if wl.get('salinity') else "") +
                       f" cells, same spacing); minimum step time {best:.3f} s")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-shrink", type=int, default=4)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import pyafid
    wl = dict(WORKLOADS[args.workload])
    dt = wl.pop("dt")
    g = pyafid.Solver(wl)
    nz, ny = wl["nzm"], wl["nym"]
    g.upload("tempbp", np.ones((nz, ny, 1)))
    g.upload("temptp", np.zeros((nz, ny, 1)))
    if wl.get("salinity"):
        g.upload("salbp", np.ones((wl["nzmr"], wl["nymr"], 1)))
        g.upload("saltp", np.zeros((wl["nzmr"], wl["nymr"], 1)))
    g.create_initial_conditions(args.seed, 1e-3)
    g.synchronize()
    print(f"bench: {args.workload}, device memory {g.device_bytes() / 2**30:.1f} GiB", file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        g.time_marcher(dt)
    g.synchronize()

    g.set_timing(True)
    step_ms, groups = [], {}
    t_total0 = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        g.time_marcher(dt)
        tm = g.get_timing()          # synchronises the stream
        step_ms.append((time.perf_counter() - t0) * 1e3)
        for k, v in tm.items():
            groups[k] = groups.get(k, 0.0) + v
    g.synchronize()
    total_s = time.perf_counter() - t_total0
    g.set_timing(False)

    div, divr = g.check_divergence()
    gq = g.global_quantities()
    ncell = cells(wl)
    ms_per_step = total_s * 1e3 / args.steps
    value = ncell * args.steps / total_s / 1e6

    # roofline of the dominant kernel group
    nsub = wl["nsst"] * args.steps
    launches = {k: nsub * (3 if k == "mgrd_gradient" else (2 if k == "poisson_fft_z" else 1)) for k in groups}
    kb = kernel_bytes(dict(wl))
    dom = max((k for k in groups if k in kb), key=groups.get)
    per_launch_bytes = kb[dom] / (3 if dom == "mgrd_gradient" else (2 if dom == "poisson_fft_z" else 1))
    avg_ms = groups[dom] / launches[dom]
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": int(per_launch_bytes)}
    total_alg = sum(kb.values()) * wl["nsst"]
    step_frac = total_alg / (min(step_ms) * 1e-3) / 1e9 / HBM_PEAK_GBS

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(args)
        except Exception as e:   # reported, never replaces the GPU number
            cpu = {"error": str(e)}

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "min_ms_per_step": round(min(step_ms), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "form": "fp64",
                   "base": [wl["nxm"], wl["nym"], wl["nzm"]],
                   "refined": [wl["nxmr"], wl["nymr"], wl["nzmr"]] if wl.get("salinity") else None,
                   "Ra": wl["rayt"], "Pr": wl["prat"], "Sc": wl.get("pras") if wl.get("salinity") else None,
                   "dt": dt, "cells": ncell},
        "roofline": roofline,
        "step_hbm_frac": round(step_frac, 4),
        "kernel_ms_per_step": {k: round(v / args.steps, 3) for k, v in sorted(groups.items(), key=lambda t: -t[1])},
        "max_div": div, "max_div_refined": divr, "nu_vol": gq["nu"], "ke": gq["ke"],
        "cpu_baseline": cpu,
    }
    g.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
