#!/usr/bin/env python3
"""Benchmark of the RK3 step (TimeMarcher.F90:13-165) on one MI355X, in the
mixed FP32/FP64 form against the plain FP64 form.

Workload (the largest one-GPU configuration): multi-resolution Rayleigh-Benard
convection with a passive scalar, base grid 384 (x, cosine-clustered) x 768 x
768, refined grid 768 x 1536 x 1536, Ly = Lz = 2, Ra = 1e9, Pr = 1, Sc = 16,
dt = 5e-5, u = 0 and T = 1-x, S = 1-x plus seeded 1e-3 noise.  Synthetic data
generated on the device.

Both forms start from the same seeded state and run the same steps; the
mixed form's error is then reported against the FP64 form: max|div u|,
Nu_vol and KE relative errors and relative L2 errors of the fields.  The
error leg then runs both forms and the FP64 CPU oracle for 50 steps on the
same grids cut by 4 in y and z (same spacing) and reports each form against
the oracle: relative L2 of the fields, max|div u|, Nu and KE relative error.

Prints ONE JSON line: value = Mcell-steps/s of the mixed form =
(base + refined cells) x steps / s, with the HBM roofline of its dominant
kernel group (HIP-event timing on the library's stream inside the timed
region), the FP64 form's numbers beside it, and a CPU baseline: the FP64 CPU
oracle (OpenMP) on the cut grids, its minimum step time over the error leg.
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
    # melting of a solid sphere (radius 0.15, phase-field width 1.5 refined
    # spacings) in RBC at Ra = 1e8: base 256 x 512 x 512, refined phi/T grid
    # 512 x 1024 x 1024, Stefan number 1, A = 1, T_m = 0
    "pf_sphere_256x512x512_512x1024x1024": dict(
        nxm=256, nym=512, nzm=512, nsst=3, multires=1, nxmr=512, nymr=1024, nzmr=1024, phasefield=1,
        alx3=1.0, ylen=2.0, zlen=2.0, istr3=8, istr3r=8, rayt=1e8, prat=1.0, active_t=1, gaxis=1,
        pf_a=1.0, pf_s=1.0, pf_tm=0.0, pf_eps=1.5 / 512, pf_r0=0.15, dt=1e-4),
}
DEFAULT = "rbc_multires_384x768x768_768x1536x1536"
ENSEMBLE = "ensemble_2d_256"


def ensemble_runs(n=256, seed0=3000):
    """The 2-D RBC ensemble: run r draws its grid and physics from seed
    seed0 + r: nx in {64, 96, 128} wall-normal cells (cosine-clustered), nz in
    {128, 256, 512} periodic cells, aspect ratio 2 (Lz = 2), Ra log-uniform in
    [1e6, 1e9], Pr in {1, 7}; IDTV = 1, CFLMAX = 0.5, DT = 1e-4.

    DTMAX is 1e-2 or less where the explicit z diffusion needs it: the RK3
    step treats only the wall-normal diffusion implicitly, and its explicit
    part is stable for nu dt / dz^2 up to about 0.63 (real-axis bound 2.51
    of the scheme over the eigenvalue 4 nu / dz^2).  With nu = 1 / min(Re,
    Pe) (Re = sqrt(Ra / Pr), Pe = Re Pr), DTMAX = min(1e-2, 0.5 dz^2 /
    nu).  A fixed 1e-2 lets the Pr = 7 runs on nz = 512 grids blow up within
    20 steps (KE from 1e-12 to 3e-2, the oracle included), since the CFL
    limit alone does not bound dt while the velocity is small."""
    cfgs, ctls = [], []
    for r in range(n):
        g = np.random.default_rng(seed0 + r)
        nxm = int(g.choice([64, 96, 128]))
        nzm = int(g.choice([128, 256, 512]))
        ra = float(10.0 ** g.uniform(6.0, 9.0))
        pr = float(g.choice([1.0, 7.0]))
        cfgs.append(dict(nxm=nxm, nym=1, nzm=nzm, nsst=3, alx3=1.0, ylen=1.0, zlen=2.0, istr3=8,
                         rayt=ra, prat=pr, active_t=1, gaxis=1))
        re = np.sqrt(ra / pr)
        dz = 2.0 / nzm
        dtmax = min(1e-2, 0.5 * dz * dz * min(re, re * pr))
        ctls.append(dict(variable_dt=1, dt=1e-4, resid=1e-3, limit_cfl=0.5, dtmin=1e-8, dtmax=dtmax, ntst=1 << 30))
    return cfgs, ctls


def ensemble_bytes_per_cell():
    """Algorithmic FP64 bytes per cell and substep of the ensemble kernel:
    columns read u, w, T, p and the 3 explicit terms and write u, w, T and
    the terms (13); divergence reads u, w and writes the half spectrum (3);
    the mode solves read and write it plus 3 factors per mode row (3.5);
    the inverse reads it and writes dph (2); the projection reads dph, u, w,
    p and writes u, w, p (7); the step's CFL / divergence check reads u, w
    once per step (2/3)."""
    return 8.0 * (13 + 3 + 3.5 + 2 + 7 + 2.0 / 3.0)


def _ensemble_cpu_worker(job):
    """Advance a subset of the ensemble's runs on one core (the oracle with
    one OpenMP thread) for a fixed number of steps; returns (cell-steps, s)."""
    import pyoracle
    from helpers import uniform
    runs, cfgs, ctls, steps = job
    pyoracle.set_threads(1)
    orcs, dts = [], []
    for r in runs:
        c = cfgs[r]
        o = pyoracle.Oracle(c)
        nxm = c["nxm"]
        xm = o.grid("xm")[:nxm]
        o.field("tempbp")[...] = 1.0
        o.field("temptp")[...] = 0.0
        T = o.field("temp")
        amp = np.where(np.abs(xm - 0.5) + 1e-3 > 0.5, 0.5 - np.abs(xm - 0.5), 1e-3)
        T[..., :nxm] = 1.0 - xm + amp * (2 * uniform(3000 + r, T.shape)[..., :nxm] - 1)
        orcs.append(o)
        dts.append(ctls[r]["dt"])
    t0 = time.perf_counter()
    for s in range(steps):
        for q, o in enumerate(orcs):
            k = ctls[runs[q]]
            if s > 1:          # main.F90:287-297
                cfl = max(o.calc_max_cfl()[0], 1e-8)
                dts[q] = min(k["limit_cfl"] / cfl, k["dtmax"])
            o.time_marcher(dts[q])
    el = time.perf_counter() - t0
    n = sum(cfgs[r]["nxm"] * cfgs[r]["nzm"] for r in runs) * steps
    for o in orcs:
        o.close()
    return n, el


def ensemble_cpu_baseline(cfgs, ctls, args):
    """The FP64 CPU oracle on every run of the ensemble, the runs dealt
    largest-first over one worker process per core (the ensemble's own
    parallelism on a CPU), each with the same CFL control; throughput over
    the slowest worker.  Runs before the GPU is touched."""
    import multiprocessing as mp
    workers = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")), 16))
    order = sorted(range(len(cfgs)), key=lambda r: -cfgs[r]["nxm"] * cfgs[r]["nzm"])
    load = [0] * workers
    parts = [[] for _ in range(workers)]
    for r in order:
        w = load.index(min(load))
        parts[w].append(r)
        load[w] += cfgs[r]["nxm"] * cfgs[r]["nzm"]
    steps = args.cpu_steps
    with mp.get_context("fork").Pool(workers) as pool:
        res = pool.map(_ensemble_cpu_worker, [(p, cfgs, ctls, steps) for p in parts if p])
    n = sum(a for a, _ in res)
    el = max(b for _, b in res)
    ncell = sum(c["nxm"] * c["nzm"] for c in cfgs)
    return {"value": round(n / el / 1e6, 3), "unit": "Mcell-steps/s", "cores": workers, "kind": "port",
            "sample": (f"FP64 CPU oracle (C), all {len(cfgs)} runs ({ncell} cells) dealt over {workers} single-thread "
                       f"worker processes, {steps} RK3 steps each with the same CFL control; slowest worker {el:.1f} s")}


def ensemble_form(cfgs, ctls, form, steps, args):
    """Initialise, tune and time one form of the ensemble."""
    import pyafid
    prec = pyafid.MIXED if form == "mixed" else pyafid.FP64
    ens = pyafid.Ensemble([dict(c, precision=prec) for c in cfgs], ctls)
    for r, c in enumerate(cfgs):
        ens.upload(r, "tempbp", np.ones((c["nzm"], 1, 1)))
        ens.upload(r, "temptp", np.zeros((c["nzm"], 1, 1)))
    ens.create_initial_conditions(3000, 1e-3)
    log(f"ensemble {form}: {len(cfgs)} runs, device memory {ens.device_bytes() / 2**30:.2f} GiB")
    best, cands = ens.autotune(steps=5)
    log(f"ensemble {form}: autotuner picked {best} from {len(cands)} launch shapes")
    ens.create_initial_conditions(3000, 1e-3)     # the timed run starts from step 0
    ens.step(args.warmup, graph=True)
    ens.synchronize()
    t0 = time.perf_counter()
    ms = ens.step_timed(steps, graph=True)
    wall = time.perf_counter() - t0
    states = [ens.run_state(r) for r in range(len(cfgs))]
    nus = [ens.global_quantities(r)["nu"] for r in range(len(cfgs))]
    ens.close()
    return dict(ms=ms, wall=wall, best=best, cands=cands, states=states, nus=nus)


def run_ensemble(args):
    cfgs, ctls = ensemble_runs()
    ncell = sum(c["nxm"] * c["nzm"] for c in cfgs)
    steps = args.steps if args.steps is not None else 1000
    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (before the GPU is initialised)")
        try:
            cpu = ensemble_cpu_baseline(cfgs, ctls, args)
        except Exception as e:    # reported, never replaces the GPU number
            cpu = {"error": str(e)}
    forms = ["fp64", "mixed"] if args.form == "both" else [args.form]
    runs = {f: ensemble_form(cfgs, ctls, f, steps, args) for f in forms}
    hform = "mixed" if "mixed" in runs else "fp64"
    res = runs[hform]
    ms, wall, best, cands, states, nus = res["ms"], res["wall"], res["best"], res["cands"], res["states"], res["nus"]
    running = sum(1 for s in states if s["status"] == 0)
    value = ncell * steps / (ms * 1e-3) / 1e6
    summary = {f: {"value": round(ncell * steps / (r["ms"] * 1e-3) / 1e6, 2), "ms_per_step": round(r["ms"] / steps, 4),
                   "tuning": r["best"], "max_div": max(s["max_div"] for s in r["states"]),
                   "exit_codes": sorted({s["status"] for s in r["states"]}),
                   "time_reached": [round(min(s["time"] for s in r["states"]), 5), round(max(s["time"] for s in r["states"]), 5)]}
               for f, r in runs.items()}
    if len(runs) == 2:
        summary["speedup_mixed_over_fp64"] = round(runs["fp64"]["ms"] / runs["mixed"]["ms"], 3)
    nsub = 3 * steps
    avg_launch_ms = ms / nsub
    per_launch = ensemble_bytes_per_cell() * ncell
    achieved = per_launch / (avg_launch_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(ms / steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64 (mixed form)" if hform == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": ENSEMBLE, "form": hform, "runs": len(cfgs), "cells": ncell,
                   "nx": [64, 96, 128], "nz": [128, 256, 512], "aspect": 2, "Ra": [1e6, 1e9], "Pr": [1, 7],
                   "cflmax": 0.5, "dt0": 1e-4, "dtmax": 1e-2, "graph": True, "tuning": best},
        "roofline": {"bound": "hbm",
                     "kernel": ("substep: k_ensp_factors, k_ensp_columns, k_ensp_transform x2, k_ensp_modes, "
                                "k_ensp_project, k_ensp_finish" if best.get("phased") else "k_ens_substep"),
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "avg_launch_ms": round(avg_launch_ms, 4), "bytes_per_launch": int(per_launch)},
        "host_wall_s": round(wall, 4),
        "runs_running": running,
        "exit_codes": sorted({s["status"] for s in states}),
        "time_reached": [round(min(s["time"] for s in states), 5), round(max(s["time"] for s in states), 5)],
        "nu_vol": [round(min(nus), 4), round(max(nus), 4)],
        "forms": summary,
        "autotune_candidates": cands,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cells(cfg):
    n = cfg["nxm"] * cfg["nym"] * cfg["nzm"]
    if cfg.get("salinity") or cfg.get("phasefield"):
        n += cfg["nxmr"] * cfg["nymr"] * cfg["nzmr"]
    return n


def kernel_bytes(cfg, form, sweeps=1, checks=0):
    """Algorithmic HBM bytes per substep of each kernel group: every array the
    operation must read or write once, at the form's element sizes.  In the
    mixed form `sweeps` Poisson sweeps ran, after `checks` residual checks."""
    nb = cfg["nxm"] * cfg["nym"] * cfg["nzm"]
    nr = cfg["nxmr"] * cfg["nymr"] * cfg["nzmr"] if (cfg.get("salinity") or cfg.get("phasefield")) else 0
    pf = bool(cfg.get("phasefield"))
    ns = cfg["nxm"] * (cfg["nym"] // 2 + 1) * cfg["nzm"]       # complex spectral values
    nf = cfg["nxm"] * (cfg["nym"] // 2 + 1) * (cfg["nzm"] // 2 + 1)   # factor rows (modes i, nz-i shared)
    d = 8                                                     # FP64 state
    h = 8 if form == "fp64" else 4                            # explicit terms, refined fields
    c = 16 if form == "fp64" else 8                           # complex spectrum
    r = 8 if form == "fp64" else 4                            # real factor
    q = 8 if form == "fp64" else 4                            # pressure correction dph of the Poisson sweeps
    b = {
        # explicit terms + RHS + x-solve of vx, vy, vz, T in one sweep:
        # read u, v, w, T, p and the 4 old explicit terms, write u, v, w, T and the 4 new terms
        # (+ phic and the latent heat with the phase field)
        "column_base": 9 * d * nb + 8 * h * nb + (2 * d * nb if pf else 0),
        # salinity: read S, vxr, vyr, vzr, old H_S; write S, H_S
        "column_sal": 7 * h * nr,
        # read dph, u,v,w,p; write u,v,w,p
        "project": 8 * d * nb + q * nb,
        "mgrd_vxr": h * nr + d * nb,
        "mgrd_vyr": h * nr + d * nb,
        "mgrd_vzr": h * nr + d * nb,
        # read S, write salc
        "mgrd_interp_sal": h * nr + h * nb,
    }
    if pf:
        b.pop("column_sal"); b.pop("mgrd_vxr"); b.pop("mgrd_vyr"); b.pop("mgrd_vzr"); b.pop("mgrd_interp_sal")
        # phi column: read phi, tempr, old H; write phi, H, d phi
        b["pf_column_phi"] = 6 * h * nr
        b["pf_latent_heat"] = h * nr + d * nb       # read d phi, write the coarse source
        b["pf_interp_temp"] = d * nb + h * nr       # read T, write tempr
        b["pf_interp_phi"] = h * nr + d * nb        # read phi, write phic
    # Poisson: forward y (velocity -> spectrum; a residual check also reads
    # dph), two z transforms, the per-mode tridiagonals (+ their factors),
    # inverse y (-> dph; refinement sweeps read and write dph)
    b["poisson_fwd_y"] = (3 * d * nb + c * ns) + checks * (4 * d * nb + c * ns)
    b["poisson_fft_z"] = sweeps * 2 * 2 * c * ns
    b["poisson_tridiag"] = sweeps * (2 * c * ns + 3 * r * nf)
    b["poisson_inv_y"] = (c * ns + q * nb) + (sweeps - 1) * (c * ns + 2 * q * nb)
    return b


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), for the baseline's label."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cut_workload(name, f):
    """The workload with the periodic extent and sizes divided by f (same spacing)."""
    wl = dict(WORKLOADS[name])
    dt = wl.pop("dt")
    wl.update(nym=max(wl["nym"] // f, 1), nzm=max(wl["nzm"] // f, 1), ylen=wl["ylen"] / f, zlen=wl["zlen"] / f)
    if wl.get("salinity") or wl.get("phasefield"):
        wl.update(nymr=wl["nymr"] // f, nzmr=wl["nzmr"] // f)
    return wl, dt


def error_leg(args):
    """Both forms and the FP64 CPU oracle for args.error_steps steps on the
    cut grids, from the same seeded state (made on the device, downloaded
    into the oracle): each form against the oracle.  The oracle's per-step
    times give the CPU baseline."""
    import pyafid
    import pyoracle
    wl, dt = cut_workload(args.workload, args.cpu_shrink)
    threads = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")), 16))
    pyoracle.set_threads(threads)
    names = FIELDS_CMP + (["sal"] if wl.get("salinity") else []) + (["phi"] if wl.get("phasefield") else [])
    gpu = {}
    o = None
    for form in ("fp64", "mixed"):
        g = start_solver(wl, form, args.seed)
        if o is None:
            o = pyoracle.Oracle(wl)
            state = FIELDS_CMP[:4] + ["tempbp", "temptp"] + (["sal", "salbp", "saltp"] if wl.get("salinity") else []) + \
                (["phi"] if wl.get("phasefield") else [])
            for n in state:
                o.field(n)[...] = g.download(n)
            if wl.get("salinity"):
                o.call("interp_sal_multigrid")
            if wl.get("phasefield"):
                o.call("interp_temp_multigrid")
                o.call("interp_phi_multigrid")
        for _ in range(args.error_steps):
            g.time_marcher(dt)
        g.synchronize()
        gpu[form] = {"div": g.check_divergence(), "gq": g.global_quantities(),
                     "fields": {n: interior(wl, n, g.download(n)) for n in names}}
        g.close()
    log(f"error leg: oracle, {args.error_steps} steps on {wl['nxm']}x{wl['nym']}x{wl['nzm']}"
        + (f" + {wl['nxmr']}x{wl['nymr']}x{wl['nzmr']}" if "nxmr" in wl else ""))
    times = []
    for _ in range(args.error_steps):
        t0 = time.perf_counter()
        o.time_marcher(dt)
        times.append(time.perf_counter() - t0)
    od = o.check_divergence()
    oq = o.global_quantities()
    ref = {n: interior(wl, n, o.field(n).copy()) for n in names}
    o.close()
    report = {"grid": {"base": [wl["nxm"], wl["nym"], wl["nzm"]],
                       "refined": [wl["nxmr"], wl["nymr"], wl["nzmr"]] if "nxmr" in wl else None,
                       "cut": args.cpu_shrink},
              "steps": args.error_steps, "oracle": {"max_div": od[0], "max_div_refined": od[1], "nu_vol": oq["nu"], "ke": oq["ke"]}}
    for form, r in gpu.items():
        report[form + "_vs_oracle"] = {
            "max_div": r["div"][0], "max_div_refined": r["div"][1],
            "nu_rel": abs(r["gq"]["nu"] - oq["nu"]) / abs(oq["nu"]),
            "ke_rel": abs(r["gq"]["ke"] - oq["ke"]) / max(abs(oq["ke"]), 1e-300),
            "l2_rel": {n: l2_rel(r["fields"][n], ref[n]) for n in names}}
    n = cells(wl)
    best = min(times)
    cpu = {"value": round(n / best / 1e6, 3), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
           "cpu": cpu_model(),
           "sample": (f"FP64 CPU oracle (C, OpenMP {threads} threads on {cpu_model()}), the error leg's "
                      f"{len(times)} RK3 steps of the same workload with the periodic extent cut by {args.cpu_shrink} "
                      f"in y and z ({wl['nxm']}x{wl['nym']}x{wl['nzm']}" +
                      (f" + {wl['nxmr']}x{wl['nymr']}x{wl['nzmr']}" if 'nxmr' in wl else "") +
                      f" cells, same spacing); minimum step time {best:.3f} s, median {float(np.median(times)):.3f} s")}
    return report, cpu


def l2_rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - np.asarray(b, np.float64)).ravel()) /
                 max(np.linalg.norm(np.asarray(b, np.float64).ravel()), 1e-300))


def cpu_baseline(args):
    """Time the FP64 CPU oracle on a bounded sample of the workload: the same
    wall-normal grids and grid spacings, periodic extent cut by 4 in y and z."""
    import pyoracle
    from helpers import uniform
    wl = dict(WORKLOADS[args.workload])
    dt = wl.pop("dt")
    f = args.cpu_shrink
    if wl.get("salinity") or wl.get("phasefield"):
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
    if wl.get("phasefield"):
        # solid sphere of radius r0 at the centre of the sampled box
        nxmr = wl["nxmr"]
        xmr = o.grid("xmr")[:nxmr]
        ymr = (np.arange(wl["nymr"]) + 0.5) * wl["ylen"] / wl["nymr"]
        zmr = (np.arange(wl["nzmr"]) + 0.5) * wl["zlen"] / wl["nzmr"]
        rr = np.sqrt((xmr[None, None, :] - 0.5) ** 2 + (ymr[None, :, None] - 0.5 * wl["ylen"]) ** 2 +
                     (zmr[:, None, None] - 0.5 * wl["zlen"]) ** 2)
        o.field("phi")[..., :nxmr] = 0.5 * (1 - np.tanh(0.5 * (rr - wl["pf_r0"]) / wl["pf_eps"]))
        o.call("interp_temp_multigrid")
        o.call("interp_phi_multigrid")
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
            "cpu": cpu_model(),
            "sample": (f"FP64 CPU oracle (C, OpenMP {threads} threads on {cpu_model()}), {len(times)} RK3 step(s) of the same "
                       f"workload with the periodic extent cut by {f} in y and z "
                       f"({wl['nxm']}x{wl['nym']}x{wl['nzm']}" +
                       (f" + {wl['nxmr']}x{wl['nymr']}x{wl['nzmr']}" if 'nxmr' in wl else "") +
                       f" cells, same spacing); minimum step time {best:.3f} s")}


FIELDS_CMP = ["vx", "vy", "vz", "temp", "pr"]


def interior(wl, name, a):
    if name in ("sal", "phi", "tempr"):
        return a[..., :wl["nxmr"]]
    if name == "vx":
        return a[..., 1:wl["nxm"]]
    a = a[..., :wl["nxm"]]
    return a - a.mean() if name == "pr" else a


def start_solver(wl, form, seed):
    """A context of the given form holding the seeded initial state: T = 1 -
    x (and S = 1 - x) plus 1e-3 wall-clamped noise, u = 0; with the phase
    field the solid sphere."""
    import pyafid
    g = pyafid.Solver(dict(wl, precision=pyafid.MIXED if form == "mixed" else pyafid.FP64))
    nz, ny = wl["nzm"], wl["nym"]
    g.upload("tempbp", np.ones((nz, ny, 1)))
    g.upload("temptp", np.zeros((nz, ny, 1)))
    if wl.get("salinity"):
        g.upload("salbp", np.ones((wl["nzmr"], wl["nymr"], 1)))
        g.upload("saltp", np.zeros((wl["nzmr"], wl["nymr"], 1)))
    g.create_initial_conditions(seed, 1e-3)
    g.synchronize()
    return g


def run_form(wl, dt, form, args):
    """Initialise, warm up and time one form; returns its measurements and
    its final fields (for the error report)."""
    g = start_solver(wl, form, args.seed)
    log(f"{form}: {args.workload}, device memory {g.device_bytes() / 2**30:.1f} GiB")
    for _ in range(args.warmup):
        g.time_marcher(dt)
    g.synchronize()

    if args.graph:
        # launch-bound workloads: the timed steps are replays of a captured
        # two-step HIP graph (kernel nodes only), one synchronisation at the
        # end; the capture itself happens in an untimed warm-up call
        g.time_marcher_graph(dt, 2)
        g.synchronize()
        t_total0 = time.perf_counter()
        g.time_marcher_graph(dt, args.steps)
        g.synchronize()
        total_s = time.perf_counter() - t_total0
        step_ms = [total_s * 1e3 / args.steps]
        # per-kernel-group times (HIP events) from eager steps after the timed region
        nprof = min(args.steps, 10)
        g.set_timing(True)
        groups = {}
        for _ in range(nprof):
            g.time_marcher(dt)
            for k, v in g.get_timing().items():
                groups[k] = groups.get(k, 0.0) + v
        g.set_timing(False)
        groups = {k: v / nprof for k, v in groups.items()}
    else:
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
        groups = {k: v / args.steps for k, v in groups.items()}
    log(f"{form}: {args.steps} steps in {total_s:.3f} s" + (" (graph replay)" if args.graph else ""))

    res = {"form": form, "total_s": total_s, "step_ms": step_ms, "groups": groups,
           "steps_done": args.warmup + args.steps + ((2 + min(args.steps, 10)) if args.graph else 0)}
    res["div"], res["divr"] = g.check_divergence()
    res["gq"] = g.global_quantities()
    # CalcMeanProfiles (a diagnostic the reference calls every TOUT, outside
    # the step): one untimed call, then the best of three; each call
    # synchronises and copies its 15 x nxm profiles back
    g.mean_profiles()
    tm = []
    for _ in range(3):
        t0 = time.perf_counter()
        g.mean_profiles()
        tm.append((time.perf_counter() - t0) * 1e3)
    res["means_ms"] = min(tm)
    res["sal_stats_ms"] = None
    if wl.get("salinity"):
        g.sal_stats()
        tm = []
        for _ in range(3):
            t0 = time.perf_counter()
            g.sal_stats()
            tm.append((time.perf_counter() - t0) * 1e3)
        res["sal_stats_ms"] = min(tm)
    res["phi_stats_ms"] = None
    if wl.get("phasefield"):
        g.phi_stats()
        tm = []
        for _ in range(3):
            t0 = time.perf_counter()
            g.phi_stats()
            tm.append((time.perf_counter() - t0) * 1e3)
        res["phi_stats_ms"] = min(tm)
    res["residuals"] = g.refinement_residuals()
    if not args.no_errors:
        names = FIELDS_CMP + (["sal"] if wl.get("salinity") else []) + (["phi"] if wl.get("phasefield") else [])
        res["fields"] = {n: interior(wl, n, g.download(n)) for n in names}
    g.close()
    return res


def source_digest():
    """sha256 (16 hex digits) of the kernel sources (afid-murphfi_amd/csrc/*.hip,
    *.hpp), the build a traffic measurement belongs to."""
    import glob
    import hashlib
    h = hashlib.sha256()
    for f in sorted(glob.glob(os.path.join(ROOT, "afid-murphfi_amd", "csrc", "*.hip")) +
                    glob.glob(os.path.join(ROOT, "afid-murphfi_amd", "csrc", "*.hpp"))):
        h.update(os.path.basename(f).encode())
        h.update(open(f, "rb").read())
    return h.hexdigest()[:16]


def measured_traffic(workload, form, group):
    """HBM bytes per launch of a kernel group, measured with rocprofv3
    FETCH_SIZE and WRITE_SIZE (one pass each) and scaled by the same counters
    on a plain copy (--copy-calibration).  Read from the lines
    `traffic-per-launch: <workload> <form> <group> <bytes> sources=<digest>` of
    the summaries in profiles/; a figure measured on other kernel sources than
    these (another digest) is not used, so the result is None until the
    counters are taken again."""
    import glob
    digest = source_digest()
    for f in glob.glob(os.path.join(ROOT, "profiles", "*.md")):
        for line in open(f):
            if not line.startswith("traffic-per-launch:"):
                continue
            w = line.split()
            if len(w) == 6 and w[1:4] == [workload, form, group] and w[5] == "sources=" + digest:
                return float(w[4])
    return None


def roofline(wl, res, form, workload=None):
    """Roofline of the dominant kernel group: algorithmic bytes per launch
    over the average launch time (HIP events on the library's stream)."""
    nsst = wl["nsst"]
    r = res["residuals"]
    sweeps = 1 + sum(1 for q in r if q > 0.0 and q > 1e-10) if form == "mixed" else 1
    checks = sum(1 for q in r if q > 0.0) if form == "mixed" else 0
    kb = kernel_bytes(wl, form, sweeps, checks)
    launches = {"poisson_fft_z": 2 * sweeps, "poisson_tridiag": sweeps, "poisson_inv_y": sweeps,
                "poisson_fwd_y": 1 + checks}
    grp = res["groups"]
    dom = max((k for k in grp if k in kb), key=grp.get)
    nl = launches.get(dom, 1)
    avg_ms = grp[dom] / nsst / nl
    per_launch = kb[dom] / nl
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    step_bytes = sum(kb.values()) * nsst
    traffic = measured_traffic(workload, form, dom)
    return ({"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
             "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
             "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": int(per_launch)},
            round(step_bytes / (min(res["step_ms"]) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed RK3 steps (10; 1000 for the ensemble)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS) + [ENSEMBLE])
    ap.add_argument("--form", default="both", choices=["both", "mixed", "fp64"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy a 2 GiB device array 3 times (the reference kernel for FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--no-errors", action="store_true", help="skip the field downloads of the error report")
    ap.add_argument("--graph", action="store_true", default=None,
                    help="time replays of a captured HIP graph (default: on for workloads under 2e7 cells)")
    ap.add_argument("--no-graph", dest="graph", action="store_false")
    ap.add_argument("--cpu-shrink", type=int, default=None,
                    help="divide y and z of the CPU sample by this (default 4; 1 for workloads under 2e7 cells)")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=25.0)
    ap.add_argument("--error-steps", type=int, default=None,
                    help="steps of the error leg on the cut grids (50; 200, the whole run, for the 64^3 workload)")
    ap.add_argument("--no-error-leg", action="store_true",
                    help="skip the error leg (the CPU baseline then times a few oracle steps on its own)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        # a plain device-to-device copy of known size: the counter readings of
        # its dispatches calibrate FETCH_SIZE / WRITE_SIZE against real bytes
        import torch
        n = 1 << 28
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "dispatches": 3}))
        return

    if args.workload == ENSEMBLE:
        return run_ensemble(args)
    wl = dict(WORKLOADS[args.workload])
    dt = wl.pop("dt")
    ncell = cells(wl)
    small = ncell < 20_000_000          # launch-bound: graph replay, CPU sample at full size
    if args.steps is None:
        args.steps = 200 if small else 10
    if args.graph is None:
        args.graph = small
    if args.cpu_shrink is None:
        args.cpu_shrink = 1 if small else 4
    if args.error_steps is None:
        args.error_steps = 200 if small else 50
    forms = ["fp64", "mixed"] if args.form == "both" else [args.form]
    runs = {f: run_form(wl, dt, f, args) for f in forms}

    head = runs["mixed"] if "mixed" in runs else runs["fp64"]
    hform = head["form"]
    value = ncell * args.steps / head["total_s"] / 1e6
    rl, step_frac = roofline(wl, head, hform, args.workload)

    summary = {}
    for f, r in runs.items():
        rlf, sf = roofline(wl, r, f, args.workload)
        summary[f] = {"value": round(ncell * args.steps / r["total_s"] / 1e6, 2),
                      "ms_per_step": round(r["total_s"] * 1e3 / args.steps, 3),
                      "min_ms_per_step": round(min(r["step_ms"]), 3),
                      "median_ms_per_step": round(float(np.median(r["step_ms"])), 3),
                      "max_div": r["div"], "max_div_refined": r["divr"],
                      "nu_vol": r["gq"]["nu"], "ke": r["gq"]["ke"],
                      "roofline": rlf, "step_hbm_frac": sf,
                      "kernel_ms_per_step": {k: round(v, 3) for k, v in sorted(r["groups"].items(), key=lambda t: -t[1])}}
        if f == "mixed":
            summary[f]["refinement_residuals_last_substep"] = r["residuals"]
        # mean profiles: u, v, w, T of the base grid read once = 32 B per cell
        mb = 32.0 * wl["nxm"] * wl["nym"] * wl["nzm"]
        summary[f]["mean_profiles"] = {"ms": round(r["means_ms"], 3), "bytes": int(mb),
                                       "achieved_GBs": round(mb / (r["means_ms"] * 1e-3) / 1e9, 1)}
        if r["sal_stats_ms"] is not None:
            # S, vxr, vyr, vzr of the refined grid read once, in the form's storage type
            sb = (8.0 if f == "fp64" else 4.0) * 4 * wl["nxmr"] * wl["nymr"] * wl["nzmr"]
            summary[f]["sal_stats"] = {"ms": round(r["sal_stats_ms"], 3), "bytes": int(sb),
                                       "achieved_GBs": round(sb / (r["sal_stats_ms"] * 1e-3) / 1e9, 1)}
        if r["phi_stats_ms"] is not None:
            pb = (8.0 if f == "fp64" else 4.0) * wl["nxmr"] * wl["nymr"] * wl["nzmr"]     # phi read once
            summary[f]["phi_stats"] = {"ms": round(r["phi_stats_ms"], 3), "bytes": int(pb),
                                       "achieved_GBs": round(pb / (r["phi_stats_ms"] * 1e-3) / 1e9, 1)}
    errors = None
    if len(runs) == 2:
        a, b = runs["fp64"], runs["mixed"]
        errors = {"steps": a["steps_done"],
                  "max_div_mixed": b["div"], "max_div_fp64": a["div"],
                  "nu_rel": abs(b["gq"]["nu"] - a["gq"]["nu"]) / abs(a["gq"]["nu"]),
                  "ke_rel": abs(b["gq"]["ke"] - a["gq"]["ke"]) / max(abs(a["gq"]["ke"]), 1e-300)}
        if not args.no_errors:
            errors["l2_rel"] = {n: float(np.linalg.norm((b["fields"][n] - a["fields"][n]).ravel()) /
                                         max(np.linalg.norm(a["fields"][n].ravel()), 1e-300))
                                for n in a["fields"]}
        summary["speedup_mixed_over_fp64"] = round(a["total_s"] / b["total_s"], 3)

    cpu = None
    leg = None
    if not args.no_error_leg:
        log("error leg: both forms and the oracle on the cut grids")
        try:
            leg, cpu = error_leg(args)
        except Exception as e:   # reported, never replaces the GPU number
            leg = {"error": str(e)}
    elif not args.no_cpu_baseline:
        log("CPU baseline")
        try:
            cpu = cpu_baseline(args)
        except Exception as e:   # reported, never replaces the GPU number
            cpu = {"error": str(e)}
    if args.no_cpu_baseline:
        cpu = None

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["total_s"] * 1e3 / args.steps, 3),
        "min_ms_per_step": round(min(head["step_ms"]), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64 (mixed form)" if hform == "mixed" else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "form": hform,
                   "base": [wl["nxm"], wl["nym"], wl["nzm"]],
                   "refined": [wl["nxmr"], wl["nymr"], wl["nzmr"]] if "nxmr" in wl else None,
                   "Ra": wl["rayt"], "Pr": wl["prat"], "Sc": wl.get("pras") if wl.get("salinity") else None,
                   "dt": dt, "cells": ncell, "graph": bool(args.graph)},
        "roofline": rl,
        "step_hbm_frac": step_frac,
        "forms": summary,
        "mixed_vs_fp64": errors,
        "mixed_vs_oracle": leg.get("mixed_vs_oracle") if leg else None,
        "fp64_vs_oracle": leg.get("fp64_vs_oracle") if leg else None,
        "error_leg": {k: v for k, v in leg.items() if not k.endswith("_vs_oracle")} if leg else None,
        "cpu_baseline": cpu,
        "kernel_sources": source_digest(),
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
