#!/usr/bin/env python
"""Benchmark of the batched all-level density build on one MI355X.

One step = one density build (TSC weights on every AMR level, deterministic
fixed-point reduction, all grids in one persistent deposit launch) over a
synthetic clustered snapshot whose particles are already resident in HBM:
binning (count, scan, scatter), deposit, normalisation and FP64 grid sums are
all inside the timed region.

Prints ONE JSON line on stdout; progress goes to stderr.
  value      particles/s of the batched build
  roofline   the persistent deposit kernel against MI355X's 8 TB/s HBM peak:
             algorithmic bytes = 16 B per (particle, grid) pair + 4 B per output
             cell, over its average duration measured with HIP events on the
             stream it runs on
  plain_gpu  the same build as one launch per grid with global FP32 atomics
  cpu_baseline  the oracle's OpenMP restatement of the reference algorithm
             (INTERPOLATE_DENSITY_TSC, FP32) on this host's cores
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particles/s TSC deposit, all AMR levels (batched vs plain GPU), % HBM roofline"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2, help="workload id (asohf_amd.synth.CONFIGS)")
    ap.add_argument("--mode", default="tsc", choices=["tsc", "table"])
    ap.add_argument("--reduction", default="deterministic", choices=["deterministic", "atomic"])
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--tune", default="", help="tx,ty,tz,chunk,waves_per_cu: fixed tuning, no autotune")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="only run a device-to-device copy of this many GiB (counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the density build of one snapshot runs on a single "
              "MI355X (multi-GPU sharding of the build is not implemented); "
              "run it with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    if args.copy_calibration > 0:
        # a plain copy of known size: the reference point that scales the
        # FETCH_SIZE / WRITE_SIZE counters into HBM bytes
        n = int(args.copy_calibration * (1 << 30)) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": 4 * n, "launches": 5}), flush=True)
        return 0

    from asohf_amd import synth
    from asohf_amd.density import DensityBuilder

    spec = synth.CONFIGS[args.config]
    t0 = time.time()
    snap = synth.make_snapshot(spec)
    g = snap.geom
    log(f"[bench] snapshot {spec.name}: {snap.n} particles, {g.npatch_total} patches, "
        f"{g.ncells_total} cells, generated in {time.time() - t0:.1f}s")
    dev = torch.device("cuda:0")
    x, y, z, m = (torch.from_numpy(a).to(dev) for a in (snap.x, snap.y, snap.z, snap.m))
    bld = DensityBuilder(0)
    s1, n11 = bld.out_shapes(g)
    u1 = torch.empty(s1, dtype=torch.float32, device=dev)
    u11 = torch.empty(n11, dtype=torch.float32, device=dev)
    kw = dict(mode=args.mode, reduction=args.reduction, u1=u1, u11=u11)
    tune = bld.get_tune()
    if args.tune:
        tx, ty, tz, ch, wv = (int(v) for v in args.tune.split(","))
        bld.set_tune(tile=(tx, ty, tz), chunk=ch, waves_per_cu=wv)
        tune = bld.get_tune()
    elif not args.no_autotune:
        t0 = time.time()
        tune = bld.autotune(g, x, y, z, m, snap.npart_esp, mode=args.mode, reduction=args.reduction)
        log(f"[bench] autotune picked {tune} in {time.time() - t0:.1f}s")

    for _ in range(max(args.warmup, 1)):
        bld.density(g, x, y, z, m, snap.npart_esp, **kw)
    torch.cuda.synchronize()
    per = {"ms_bin": [], "ms_scatter": [], "ms_deposit": [], "ms_sums": [], "ms_total": []}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = bld.density(g, x, y, z, m, snap.npart_esp, **kw)
        st = bld.stats()
        for k in per:
            per[k].append(st[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1000.0 * wall / args.steps
    value = snap.n / (wall / args.steps)
    avg = {k: float(np.mean(v)) for k, v in per.items()}
    n_pairs, n_cells, n_entries = st["n_pairs"], st["n_cells"], st["n_entries"]
    P = g.npatch_total
    b_fwd = 16.0 * n_pairs + 4.0 * n_cells + 8.0 * (P + 1)
    b_dep = 16.0 * n_pairs + 4.0 * n_cells
    achieved = b_dep / (avg["ms_deposit"] * 1e-3) / 1e9
    log(f"[bench] batched: {ms_step:.3f} ms/step, {value / 1e9:.3f} G particles/s; "
        f"phases {avg}; pairs {n_pairs} entries {n_entries}")
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_deposit_{spec.name}_{args.mode}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out = {
        "metric": METRIC, "value": value, "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": spec.name, "particles": snap.n, "base_grid": g.nx, "levels": g.nl,
                   "patches": P, "cells": int(n_cells), "mode": args.mode, "reduction": args.reduction,
                   "tune": tune},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "deposit_kernel", "algorithmic_bytes": b_dep,
                     "avg_ms": avg["ms_deposit"]},
        "pipeline": {"algorithmic_bytes": b_fwd, "achieved_gbs": b_fwd / (ms_step * 1e-3) / 1e9,
                     "frac": b_fwd / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, "phase_ms": avg,
                     "n_pairs": int(n_pairs), "n_entries": int(n_entries)},
    }
    u1_gpu = r.u1.cpu().numpy()

    if args.mode == "tsc":
        # the hand-written adjoint: d(sum(w*rho))/d(positions, masses)
        gen = torch.Generator(device=dev).manual_seed(spec.seed)
        w1 = torch.randn(s1, generator=gen, device=dev)
        w11 = torch.randn(n11, generator=gen, device=dev)
        bld.adjoint(g, x, y, z, m, w1, w11)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            bld.adjoint(g, x, y, z, m, w1, w11)
        torch.cuda.synchronize()
        aw = 1000.0 * (time.perf_counter() - t0) / args.steps
        b_adj = 16.0 * n_pairs + 4.0 * n_cells + 16.0 * snap.n
        out["adjoint"] = {"ms_per_step": aw, "ratio_to_forward": aw / ms_step,
                          "algorithmic_bytes": b_adj, "achieved_gbs": b_adj / (aw * 1e-3) / 1e9}
        log(f"[bench] adjoint: {aw:.3f} ms/step ({aw / ms_step:.2f}x forward)")

    if not args.no_plain:
        kwp = dict(kw, path="plain", reduction="atomic")
        bld.density(g, x, y, z, m, snap.npart_esp, **kwp)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            bld.density(g, x, y, z, m, snap.npart_esp, **kwp)
        torch.cuda.synchronize()
        pw = (time.perf_counter() - t0) / args.plain_steps
        out["plain_gpu"] = {"value": snap.n / pw, "unit": "particles/s", "ms_per_step": 1000 * pw,
                            "launches": 2 + P, "speedup_batched_vs_plain": (snap.n / (wall / args.steps)) / (snap.n / pw)}
        log(f"[bench] plain: {1000 * pw:.3f} ms/step")

    if not args.no_cpu:
        # cpu_baseline leg: the oracle's restatement of the reference algorithm
        from oracle import oracle as O
        cores = args.cpu_threads or len(os.sched_getaffinity(0))
        os.environ.setdefault("OMP_NUM_THREADS", str(cores))
        mode = "B" if args.mode == "tsc" else "A"
        t0 = time.perf_counter()
        c1, c11 = O.density32(snap, mode, nthreads=cores)
        tc = time.perf_counter() - t0
        sel = c1 > c1.mean()
        diff = float(np.max(np.abs(u1_gpu[sel] - c1[sel]) / c1[sel]))
        out["cpu_baseline"] = {"value": snap.n / tc, "unit": "particles/s", "cores": cores,
                               "kind": "port",
                               "sample": f"the whole {spec.name} build ({snap.n} particles, {P} patches) "
                                         f"by the FP32 OpenMP restatement of "
                                         f"{'INTERPOLATE_DENSITY_TSC' if mode == 'B' else 'DENSITY'}",
                               "seconds": tc, "max_rel_diff_base_vs_gpu": diff}
        log(f"[bench] cpu baseline: {tc:.2f}s on {cores} threads")
    print(json.dumps(out), flush=True)
    bld.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
