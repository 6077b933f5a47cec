#!/usr/bin/env python
"""Benchmark of the batched all-level density build on one MI355X.

One step = one density build (TSC weights on every AMR level, deterministic
fixed-point reduction) over a synthetic clustered snapshot whose particles are
already resident in HBM: binning, deposit, normalisation (/DENBAS, /FDM) and
the FP64 grid sums are all inside the timed region; the geometry plan (built
once per hierarchy on the host and cached) is not.

Prints ONE JSON line on stdout; progress goes to stderr.
  value      particles/s of the batched build (wall clock over K steps)
  roofline   the whole step against MI355X's 8 TB/s HBM peak: B_fwd = 16 B per
             (particle, grid) pair + 4 B per output cell + 8 B per grid sum,
             over the mean device time of a step (HIP events on the stream the
             step runs on); `deposit_phase` gives the deposit kernels alone
             and `b_min_frac` the stricter count with each particle read once
  stream_copy  a device-to-device copy on the same GPU (GB/s, read + write)
  graph      the same step captured once as a hipGraph and replayed
  modes      median ms/step of every mode (TSC, kernel table) x reduction
             (deterministic, atomic) on the same snapshot
  per_level_ms  builds of the hierarchy cut at level L = 0..NL, differenced
  species, peaks, create_mesh, masks   the routines around the build, timed
             on the same snapshot
  adjoint    the hand-written backward (gradients of sum(w * rho))
  plain_gpu  the same build as one launch per grid with global FP32 atomics
  accuracy   per-field max / RMS relative error against the FP64 oracle on a
             seeded 2^17-particle sample built with the full hierarchy, and
             per-grid FP64 masses against the oracle's
  cpu_baseline  the oracle's OpenMP restatement of the reference algorithm
             (INTERPOLATE_DENSITY_TSC, FP32) on this host's cores: whole for
             workloads of up to 2^25 particles, otherwise seeded particle
             samples built with the full hierarchy, fitted as T0 + k * t and
             evaluated at the full particle count (labelled extrapolated)
  secondary  the 256^3-particle, 4-level workload's step, roofline and accuracy
Workloads (--config): asohf_amd.synth.WORKLOADS; 5 (1024^3 particles, 1024^3
base grid, 7 AMR levels, about 150,000 patches) is the default.
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


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=5, help="workload id (asohf_amd.synth.WORKLOADS)")
    ap.add_argument("--secondary", type=int, default=2,
                    help="workload of the secondary leg (0: none; skipped when equal to --config)")
    ap.add_argument("--mode", default="tsc", choices=["tsc", "table"])
    ap.add_argument("--reduction", default="deterministic", choices=["deterministic", "atomic"])
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the FP64 accuracy check")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: OMP_NUM_THREADS, else every core of the affinity set)")
    ap.add_argument("--cpu-budget", type=float, default=10.0,
                    help="seconds of per-particle CPU work in the largest sample when the whole build is too long")
    ap.add_argument("--accuracy-sample", type=int, default=1 << 17,
                    help="particles of the FP64 accuracy sample (0: skip)")
    ap.add_argument("--modes-steps", type=int, default=10, help="steps per mode x reduction (0: skip)")
    ap.add_argument("--graph-steps", type=int, default=10, help="steps of the hipGraph form (0: skip)")
    ap.add_argument("--no-levels", dest="levels", action="store_false", help="skip the per-level timing")
    ap.add_argument("--species-steps", type=int, default=5, help="calls of the species partition (0: skip)")
    ap.add_argument("--peaks-steps", type=int, default=5,
                    help="calls of the peak-candidate scan over every level (0: skip)")
    ap.add_argument("--contrastec", type=float, default=200.0, help="CONTRASTEC of the peak-candidate scan")
    ap.add_argument("--mesh-steps", type=int, default=2, help="calls of the mesh builder (0: skip)")
    ap.add_argument("--masks-steps", type=int, default=3,
                    help="calls of the grid masks and the density dump writer (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=10, help="adjoint steps (0: skip)")
    ap.add_argument("--plain-max-tests", type=float, default=5e12,
                    help="skip the plain form above this many particle-box tests")
    ap.add_argument("--tune", default="", help="tx,ty,tz,chunk,waves_per_cu[,block]: fixed tuning, no autotune")
    ap.add_argument("--copy-calibration", type=float, default=0.0,
                    help="only run a device-to-device copy of this many GiB (counter calibration)")
    return ap.parse_args(argv)


def stream_copy_gbs(torch, gib: float = 4.0, reps: int = 10) -> dict:
    """Achieved bandwidth of a plain device-to-device copy (read + write
    bytes over the mean time of one copy, HIP events on the current stream)."""
    n = int(gib * (1 << 30)) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(reps):
        b.copy_(a)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    del a, b
    torch.cuda.empty_cache()
    return {"gbs": 2.0 * 4 * n / (ms * 1e-3) / 1e9, "bytes_each_way": 4 * n, "ms": ms, "reps": reps,
            "frac_of_peak": 2.0 * 4 * n / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def make_workload(cfg: int, dev):
    import torch

    from asohf_amd import synth
    spec = synth.WORKLOADS[cfg]
    t0 = time.time()
    if cfg != 1:
        # generated on the device (numpy is too slow at these sizes)
        snap = synth.make_snapshot_device(spec, device=dev)
        xyzm = (snap.x, snap.y, snap.z, snap.m)
    else:
        snap = synth.make_snapshot(spec)
        xyzm = tuple(torch.from_numpy(a).to(dev) for a in (snap.x, snap.y, snap.z, snap.m))
    log(f"[bench] snapshot {spec.name}: {int(xyzm[0].numel())} particles, {snap.geom.npatch_total} patches, "
        f"{snap.geom.ncells_total} cells, generated in {time.time() - t0:.1f}s")
    return spec, snap, xyzm


def timed_steps(bld, g, xyzm, esp, kw, steps, warmup):
    """K eager steps after W warm-up steps: wall ms/step and the per-step
    device times (HIP events) of every phase."""
    import numpy as np
    import torch
    x, y, z, m = xyzm
    for _ in range(max(warmup, 1)):
        bld.density(g, x, y, z, m, esp, **kw)
    torch.cuda.synchronize()
    keys = ("ms_bin", "ms_scatter", "ms_deposit", "ms_split", "ms_sums", "ms_total")
    per = {k: [] for k in keys}
    t0 = time.perf_counter()
    r = None
    for _ in range(steps):
        r = bld.density(g, x, y, z, m, esp, **kw)
        st = bld.stats()
        for k in keys:
            per[k].append(st[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return r, st, 1000.0 * wall / steps, {k: float(np.mean(v)) for k, v in per.items()}, \
        {k: float(np.median(v)) for k, v in per.items()}


def roofline_block(spec, args, st, n, P, avg, med, ms_step):
    n_pairs, n_cells, n_entries = st["n_pairs"], st["n_cells"], st["n_entries"]
    b_fwd = 16.0 * n_pairs + 4.0 * n_cells + 8.0 * (P + 1)
    b_min = 16.0 * n + 4.0 * n_cells + 8.0 * (P + 1)
    b_dep = 16.0 * n_pairs + 4.0 * n_cells
    t_step = avg["ms_total"]
    achieved = b_fwd / (t_step * 1e-3) / 1e9
    dep_ms = avg["ms_deposit"] + avg["ms_split"]
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_step_{spec.name}_{args.mode}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_step")
    return {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
        "kernel": "the whole build step (binning, deposit, split tiles, grid sums), device time",
        "algorithmic_bytes": b_fwd, "avg_ms": t_step, "median_ms": med["ms_total"],
        "wall_ms_per_step": ms_step,
        "b_min": b_min, "b_min_frac": b_min / (t_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "n_pairs": int(n_pairs), "n_entries": int(n_entries), "n_cells": int(n_cells),
        "phase_ms": avg, "phase_ms_median": med,
        "deposit_phase": {"kernels": "deposit kernels + split_reduce", "algorithmic_bytes": b_dep,
                          "avg_ms": dep_ms, "achieved": b_dep / (dep_ms * 1e-3) / 1e9,
                          "frac": b_dep / (dep_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }


def run_workload(cfg: int, args, primary: bool) -> dict:
    import numpy as np
    import torch

    from asohf_amd.density import DensityBuilder

    dev = torch.device("cuda:0")
    spec, snap, xyzm = make_workload(cfg, dev)
    x, y, z, m = xyzm
    n = int(x.numel())
    g = snap.geom
    P = g.npatch_total
    bld = DensityBuilder(0)
    s1, n11 = bld.out_shapes(g)
    u1 = torch.empty(s1, dtype=torch.float32, device=dev)
    u11 = torch.empty(n11, dtype=torch.float32, device=dev)
    gmass = torch.empty(P + 1, dtype=torch.float64, device=dev)
    kw = dict(mode=args.mode, reduction=args.reduction, u1=u1, u11=u11, grid_mass=gmass)
    tune = bld.get_tune()
    if args.tune:
        tv = [int(v) for v in args.tune.split(",")]
        tx, ty, tz, ch, wv = tv[:5]
        bld.set_tune(tile=(tx, ty, tz), chunk=ch, waves_per_cu=wv, block=tv[5] if len(tv) > 5 else 256)
        tune = bld.get_tune()
    elif not args.no_autotune:
        t0 = time.time()
        tune = bld.autotune(g, x, y, z, m, snap.npart_esp, mode=args.mode, reduction=args.reduction)
        log(f"[bench] autotune picked {tune} in {time.time() - t0:.1f}s")

    # the first (cold) call sizes every buffer; its time is reported beside the
    # steady state
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bld.density(g, x, y, z, m, snap.npart_esp, **kw)
    torch.cuda.synchronize()
    cold_ms = 1000.0 * (time.perf_counter() - t0)
    r, st, ms_step, avg, med = timed_steps(bld, g, xyzm, snap.npart_esp, kw, args.steps, args.warmup)
    value = n / (ms_step * 1e-3)
    log(f"[bench] {spec.name} batched: {ms_step:.3f} ms/step, {value / 1e9:.3f} G particles/s; "
        f"device {avg['ms_total']:.3f} ms; phases {avg}; pairs {st['n_pairs']} entries {st['n_entries']}")
    out = {
        "metric": METRIC, "value": value, "unit": "particles/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": spec.name, "particles": n, "base_grid": g.nx, "levels": g.nl,
                   "patches": P, "cells": int(st["n_cells"]), "fdm": float(g.fdm), "mode": args.mode,
                   "reduction": args.reduction, "tune": tune},
        "roofline": roofline_block(spec, args, st, n, P, avg, med, ms_step),
        "cold_call_ms": cold_ms,
    }
    total = float(m.double().sum())
    out["checks"] = {"base_grid_mass_rel_err": abs(float(r.grid_mass[0]) - total) / total if total else 0.0}

    if args.graph_steps > 0:
        # the same step as one hipGraph (captured on the first call, replayed)
        kg = dict(kw, use_graph=True)
        bld.density(g, x, y, z, m, snap.npart_esp, **kg)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.graph_steps):
            rg = bld.density(g, x, y, z, m, snap.npart_esp, **kg)
        torch.cuda.synchronize()
        gw = 1000.0 * (time.perf_counter() - t0) / args.graph_steps
        same = bool(torch.equal(rg.u1, r.u1)) if args.reduction == "deterministic" else None
        out["graph"] = {"ms_per_step": gw, "particles_per_s": n / (gw * 1e-3), "steps": args.graph_steps,
                        "bit_identical_to_eager": same}
        log(f"[bench] graph form: {gw:.3f} ms/step")
        del rg

    if primary and args.modes_steps > 0:
        # every mode x reduction on the same snapshot, median of the steps
        table = {}
        for md in ("tsc", "table"):
            for rd in ("deterministic", "atomic"):
                k2 = dict(kw, mode=md, reduction=rd)
                bld.density(g, x, y, z, m, snap.npart_esp, **k2)
                torch.cuda.synchronize()
                walls, dd, tot = [], [], []
                for _ in range(args.modes_steps):
                    t0 = time.perf_counter()
                    bld.density(g, x, y, z, m, snap.npart_esp, **k2)
                    walls.append(1000.0 * (time.perf_counter() - t0))
                    st2 = bld.stats()
                    dd.append(st2["ms_deposit"] + st2["ms_split"])
                    tot.append(st2["ms_total"])
                s2 = bld.stats()
                tw = float(np.median(walls))
                table[f"{md}/{rd}"] = {"ms_per_step_median": tw, "particles_per_s": n / (tw * 1e-3),
                                      "device_ms_median": float(np.median(tot)),
                                      "deposit_ms_median": float(np.median(dd)), "steps": args.modes_steps,
                                      "n_entries": int(s2["n_entries"]), "n_pairs": int(s2["n_pairs"])}
        out["modes"] = table
        log(f"[bench] modes: {table}")
        bld.density(g, x, y, z, m, snap.npart_esp, **kw)
        torch.cuda.synchronize()

    if primary and args.levels and g.nl > 0:
        # time per AMR level: builds of the hierarchy cut at level L, differenced
        cum = []
        reps = max(args.steps // 4, 3)
        for L in range(0, g.nl + 1):
            gl = g.truncated(L)
            bld.density(gl, x, y, z, m, snap.npart_esp, mode=args.mode, reduction=args.reduction)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                bld.density(gl, x, y, z, m, snap.npart_esp, mode=args.mode, reduction=args.reduction)
            torch.cuda.synchronize()
            cum.append(1000 * (time.perf_counter() - t0) / reps)
        out["per_level_ms"] = {"cumulative": cum,
                               "level": [cum[0]] + [cum[i] - cum[i - 1] for i in range(1, len(cum))],
                               "patches": [int(v) for v in g.npatch]}
        log(f"[bench] per level (cumulative ms): {cum}")
        bld.density(g, x, y, z, m, snap.npart_esp, **kw)
        torch.cuda.synchronize()

    if primary:
        side_routines(args, bld, spec, snap, g, xyzm, kw, n, int(st["n_cells"]), out)

    # the forward workspaces are not needed by the adjoint; free them so the
    # adjoint's own buffers fit at any size
    u1_gpu = r.u1
    bld.release_workspaces()
    if primary and args.mode == "tsc" and args.adjoint_steps > 0:
        # the hand-written adjoint: d(sum(w*rho))/d(positions, masses)
        gen = torch.Generator(device=dev).manual_seed(spec.seed)
        w1 = torch.randn(s1, generator=gen, device=dev)
        w11 = torch.randn(n11, generator=gen, device=dev)
        bld.adjoint(g, x, y, z, m, w1, w11)
        torch.cuda.synchronize()
        aws = []
        for _ in range(args.adjoint_steps):
            t0 = time.perf_counter()
            bld.adjoint(g, x, y, z, m, w1, w11)
            aws.append(1000.0 * (time.perf_counter() - t0))
        aw = float(np.median(aws))
        b_adj = 16.0 * st["n_pairs"] + 4.0 * st["n_cells"] + 16.0 * n
        out["adjoint"] = {"ms_per_step_median": aw, "ratio_to_forward": aw / ms_step,
                          "algorithmic_bytes": b_adj, "achieved_gbs": b_adj / (aw * 1e-3) / 1e9,
                          "frac": b_adj / (aw * 1e-3) / 1e9 / HBM_PEAK_GBS, "steps": args.adjoint_steps}
        log(f"[bench] adjoint: {aw:.3f} ms/step ({aw / ms_step:.2f}x forward)")
        del w1, w11
        bld.release_workspaces()

    if primary and not args.no_plain:
        kwp = dict(kw, path="plain", reduction="atomic")
        scan = float(P) * n        # particle-box tests of the one-launch-per-patch form
        if scan > args.plain_max_tests:
            out["plain_gpu"] = {"skipped": f"{P} patch launches over {n} particles "
                                           f"({scan:.1e} particle-box tests) exceed the bench's bound"}
        else:
            steps_p = args.plain_steps if scan < 1e11 else 1
            bld.density(g, x, y, z, m, snap.npart_esp, **kwp)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps_p):
                bld.density(g, x, y, z, m, snap.npart_esp, **kwp)
            torch.cuda.synchronize()
            pw = (time.perf_counter() - t0) / steps_p
            out["plain_gpu"] = {"value": n / pw, "unit": "particles/s", "ms_per_step": 1000 * pw,
                                "steps": steps_p, "launches": 2 + P,
                                "speedup_batched_vs_plain": value / (n / pw)}
            log(f"[bench] plain: {1000 * pw:.3f} ms/step")
            bld.release_workspaces()

    if not args.no_cpu:
        cb, acc = cpu_baseline(args, bld, snap, xyzm, u1_gpu, n, P, primary)
        if primary:
            out["cpu_baseline"] = cb
        out["accuracy"] = acc
    bld.close()
    del r, u1, u11, gmass, u1_gpu, xyzm, x, y, z, m, snap
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return out


def side_routines(args, bld, spec, snap, g, xyzm, kw, n, n_cells, out):
    """The routines around the build, timed on the same snapshot."""
    import numpy as np
    import torch
    x, y, z, m = xyzm
    if args.species_steps > 0:
        # the species partition that feeds NPART_ESP (reader.f:140-336), on a
        # grid of NMAX = NX cells per axis (a build sized for this run)
        from asohf_amd.geometry import background
        rodo, re0 = (float(v) for v in background(spec.h, spec.omega_m, 0.0))
        skw = dict(nmax=(g.nx,) * 3, d=(float(g.dx),) * 3, rodo=rodo, re0=re0)
        bld.species(x, y, z, m, **skw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.species_steps):
            _, order, esp = bld.species(x, y, z, m, **skw)
        torch.cuda.synchronize()
        ts = (time.perf_counter() - t0) / args.species_steps
        out["species"] = {"ms_per_call": 1000 * ts, "particles_per_s": n / ts, "npart_esp": esp.tolist(),
                          "routine": "SORT_DM_PARTICLES_LOCALDENSITY (reader.f:140-336), NMAX = NX"}
        log(f"[bench] species: {1000 * ts:.3f} ms/call")
        del order

    if args.peaks_steps > 0:
        # the grid halo finder's peak candidates on every level of the field
        # (HALOFIND_GRID, haloes_grids.f:316-404 and 700-790, SORT_CELLS_AMR order)
        r = bld.density(g, x, y, z, m, snap.npart_esp, **kw)

        def scan_all():
            return [bld.peak_candidates(g, lev, u1=r.u1 if lev == 0 else None, u11=r.u11 if lev else None,
                                        contrastec=args.contrastec) for lev in range(g.nl + 1)]
        res = scan_all()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.peaks_steps):
            res = scan_all()
        torch.cuda.synchronize()
        tp = (time.perf_counter() - t0) / args.peaks_steps
        counts = [int(c[0].numel()) for c in res]
        out["peaks"] = {"ms_per_call_all_levels": 1000 * tp, "cells_per_s": n_cells / tp,
                        "candidates_per_level": counts, "contrastec": args.contrastec,
                        "routine": "HALOFIND_GRID peak candidates (haloes_grids.f:316-404, 700-790) "
                                   "+ SORT_CELLS_AMR (haloes_grids.f:1383-1425)"}
        log(f"[bench] peaks: {1000 * tp:.3f} ms/call over {g.nl + 1} levels, candidates {counts}")
        del res, r

    if args.mesh_steps > 0 and g.ncells_base > (1 << 26):
        out["create_mesh"] = {"skipped": f"base grid of {g.ncells_base} cells: the host growth walks every "
                                         "refinable base cell; the bench times it up to 2^26 cells"}
    elif args.mesh_steps > 0:
        # the mesh builder CREATE_MESH (grids.f:2-707) on the same particles
        base = g.truncated(0)
        mkw = dict(nl_mesh=max(g.nl, 1), npalev=1 << 20)
        try:
            mg = bld.create_mesh(base, x, y, z, snap.npart_esp, **mkw)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.mesh_steps):
                mg = bld.create_mesh(base, x, y, z, snap.npart_esp, **mkw)
            torch.cuda.synchronize()
            tm = (time.perf_counter() - t0) / args.mesh_steps
            out["create_mesh"] = {"seconds": tm, "levels": mg.nl, "npatch": mg.npatch.tolist(),
                                  "params": "asohf.dat mesh block (REFINE_THR 3, FRAC 0.05, MIN_PATCHSIZE 14, "
                                            "BOR 4, BORAMR 0, BOR_OVLP 0), NAMRX..Z 32, NPALEV 2^20",
                                  "routine": "CREATE_MESH (grids.f:2-707)"}
            log(f"[bench] create_mesh: {tm:.3f} s, npatch {mg.npatch.tolist()}")
            # the density build on the hierarchy CREATE_MESH made for these
            # particles, both modes (deterministic)
            for md in ("tsc", "table"):
                bld.density(mg, x, y, z, m, snap.npart_esp, mode=md)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(args.steps):
                    bld.density(mg, x, y, z, m, snap.npart_esp, mode=md)
                torch.cuda.synchronize()
                tb = (time.perf_counter() - t0) / args.steps
                st_m = bld.stats()
                out["create_mesh"][f"build_{md}"] = {"ms_per_step": 1000 * tb, "particles_per_s": n / tb,
                                                     "n_pairs": int(st_m["n_pairs"]),
                                                     "n_entries": int(st_m["n_entries"])}
                log(f"[bench] build on the CREATE_MESH hierarchy ({md}): {1000 * tb:.3f} ms/step, "
                    f"{n / tb / 1e9:.3f} G particles/s")
        except RuntimeError as err:
            out["create_mesh"] = {"error": str(err)}
            log(f"[bench] create_mesh: {err}")

    if args.masks_steps > 0 and g.pcx is not None:
        # the grid masks and the FW2 density dump (asohf.f:515-545)
        import tempfile
        r = bld.density(g, x, y, z, m, snap.npart_esp, **kw)
        c0, c11 = bld.refinement_masks(g)
        so = bld.overlap_mask(g)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.masks_steps):
            c0, c11 = bld.refinement_masks(g)
        torch.cuda.synchronize()
        t_cr = (time.perf_counter() - t0) / args.masks_steps
        t0 = time.perf_counter()
        for _ in range(args.masks_steps):
            so = bld.overlap_mask(g)
        torch.cuda.synchronize()
        t_vg = (time.perf_counter() - t0) / args.masks_steps
        size, t_w = 12 * n_cells, float("nan")
        if size <= (2 << 30):   # the dump of larger workloads is not written by the bench
            with tempfile.TemporaryDirectory() as td:
                path = os.path.join(td, "density_asohf00001.res")
                t0 = time.perf_counter()
                bld.write_density_dump(path, g, r.u1, c0, r.u11, c11, so)
                t_w = time.perf_counter() - t0
                size = os.path.getsize(path)
        out["masks"] = {"ms_refinement_masks": 1000 * t_cr, "ms_overlap_mask": 1000 * t_vg,
                        "dump_seconds": t_w if t_w == t_w else None, "dump_bytes": size,
                        "dump_gbs": size / t_w / 1e9 if t_w == t_w else None,
                        "zeros": {"cr0amr": int((c0 == 0).sum()), "cr0amr11": int((c11 == 0).sum()),
                                  "solap": int((so == 0).sum())},
                        "routines": "COMPUTE_CR0AMR (grids.f:2076-2226), VEINSGRID (grids.f:2229-2506), "
                                    "FW2 dump (asohf.f:530-545)"}
        log(f"[bench] masks: CR0AMR {1000 * t_cr:.3f} ms, SOLAP {1000 * t_vg:.3f} ms, "
            + (f"dump {size / 1e6:.1f} MB in {t_w:.3f} s" if t_w == t_w else
               f"dump of {size / 1e6:.1f} MB not written at this size"))
        del c0, c11, so, r
    _ = np


def cpu_threads(args) -> int:
    if args.cpu_threads:
        return args.cpu_threads
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def sample_snapshot(snap, xyzm, k, rng):
    """A seeded sample of k particles (kept in reference order, species
    counts recomputed) with the full hierarchy; host arrays plus the device
    copies."""
    import types

    import numpy as np
    import torch
    n = int(xyzm[0].numel())
    idx = np.sort(rng.choice(n, size=k, replace=False))
    ti = torch.from_numpy(idx).to(xyzm[0].device)
    dev = [a[ti].contiguous() for a in xyzm]
    host = [a.cpu().numpy() for a in dev]
    bounds = np.concatenate([[0], np.cumsum(snap.npart_esp.astype(np.int64))])
    esp = np.bincount(np.searchsorted(bounds, idx, side="right") - 1, minlength=len(snap.npart_esp))
    sn = types.SimpleNamespace(x=host[0], y=host[1], z=host[2], m=host[3],
                               npart_esp=esp[: len(snap.npart_esp)].astype(np.int32), geom=snap.geom)
    return sn, dev


def accuracy(args, bld, snap, xyzm) -> dict:
    """GPU build of a seeded particle sample (full hierarchy) against the
    FP64 oracle: max / RMS relative error on cells above each field's mean,
    per level, and the per-grid FP64 masses."""
    import numpy as np

    from oracle import oracle as O
    n = int(xyzm[0].numel())
    k = min(args.accuracy_sample, n)
    rng = np.random.default_rng(1234)
    sn, dv = sample_snapshot(snap, xyzm, k, rng)
    g = snap.geom
    mode = "B" if args.mode == "tsc" else "A"
    t0 = time.perf_counter()
    d1, d11, dgm = O.density64(sn, mode, nthreads=0)
    d1 = d1.reshape(-1)
    # the deterministic deposit's quantum: a fixed-point unit of at most
    # max(m) * 2^-29 per term, in the output units of each level
    mmax = float(sn.m.max())
    per = np.array([mmax * 2.0 ** -29 / (float(g.denbas(int(L))) * float(g.fdm)) for L in range(g.nl + 1)])
    q1 = per[0]
    q11 = np.repeat(per[g.level_of_patch()], g.patch_cells())
    off = g.patch_offsets()
    lev = g.level_of_patch()

    def errs(a, b, floor=None):
        if b.size == 0:
            return {"max": 0.0, "rms": 0.0, "cells": 0}
        sel = b > b.mean()
        if floor is not None:
            sel &= b >= floor
        if not sel.any():
            return {"max": 0.0, "rms": 0.0, "cells": 0}
        e = np.abs(a[sel].astype(np.float64) - b[sel]) / b[sel]
        return {"max": float(e.max()), "rms": float(np.sqrt(np.mean(e * e))), "cells": int(sel.sum())}

    res = {"sample_particles": k, "seed": 1234, "mode": args.mode,
           "target": "max rel err <= 1e-5 on cells above the mean; base mass to 1e-12",
           "note": "the sample thins the density 1:N/k, so many cells above the field mean hold one small "
                   "corner term; the deterministic deposit rounds every term to a fixed-point unit (at most "
                   "max(m) 2^-29), so its figure is also given on cells of at least 2^21 units (2^-8 of a "
                   "particle's mass); the atomic (FP32) deposit has no quantum"}
    for red in ("deterministic", "atomic"):
        r = bld.density(g, *dv, sn.npart_esp, mode=args.mode, reduction=red)
        u1, u11, gm = r.u1.cpu().numpy().reshape(-1), r.u11.cpu().numpy(), r.grid_mass.cpu().numpy()
        per_level = {"0": errs(u1, d1)}
        per_level_q = {"0": errs(u1, d1, 2.0 ** 21 * q1)}
        for L in range(1, g.nl + 1):
            ps = np.nonzero(lev == L)[0]
            if len(ps) == 0:
                continue
            # the patches of a level are consecutive in the packed layout
            a, b = off[ps[0]], off[ps[-1] + 1]
            per_level[str(L)] = errs(u11[a:b], d11[a:b])
            per_level_q[str(L)] = errs(u11[a:b], d11[a:b], 2.0 ** 21 * q11[a:b])
        tot = float(sn.m.astype(np.float64).sum())
        nz = dgm != 0
        gm_rel = np.abs(gm - dgm)[nz] / np.abs(dgm[nz]) if nz.any() else np.zeros(1)
        # grids holding at least 2^21 fixed-point units of mass: below that a
        # grid reached by a few corner terms carries the quantum, not an error
        big = dgm >= 2.0 ** 21 * mmax * 2.0 ** -29
        gm_rel_q = np.abs(gm - dgm)[big] / dgm[big] if big.any() else np.zeros(1)
        res[red] = {"fields": {"base": per_level["0"], "patches_all": errs(u11, d11)},
                    "fields_2p21_units": {"base": per_level_q["0"],
                                          "patches_all": errs(u11, d11, 2.0 ** 21 * q11)},
                    "per_level": per_level, "per_level_2p21_units": per_level_q,
                    "grid_mass_max_rel_diff": float(gm_rel.max()),
                    "grid_mass_max_rel_diff_2p21_units": float(gm_rel_q.max()),
                    "grid_mass_grids_2p21_units": int(big.sum()),
                    "grid_mass_grids_compared": int(nz.sum()),
                    "grid_mass_zero_grids_match": bool(np.all(gm[~nz] == 0)),
                    "base_mass_rel_err": abs(float(gm[0]) - tot) / tot if tot else 0.0}
        log(f"[bench] accuracy {red} ({k} particles): base {per_level['0']}, patches "
            f"{res[red]['fields']['patches_all']}, above 2^21 units {res[red]['fields_2p21_units']}, "
            f"grid mass max rel diff {res[red]['grid_mass_max_rel_diff']:.2e} "
            f"({res[red]['grid_mass_max_rel_diff_2p21_units']:.2e} on {int(big.sum())} grids above 2^21 units)")
    res["seconds"] = time.perf_counter() - t0
    return res


def cpu_baseline(args, bld, snap, xyzm, u1_gpu, n, P, primary):
    """The oracle's OpenMP restatement of the reference build on this host,
    and the FP64 accuracy check of the GPU build.

    Workloads of up to 2^25 particles are built whole.  For the larger ones
    (whose CPU build takes hours: every particle is tested against every patch
    box of a level, and every thread keeps a private copy of the base grid),
    seeded particle samples are built with the full hierarchy, T(k) = T0 + k*t
    is fitted through them by least squares, and the rate is N / (T0 + N*t),
    labelled extrapolated."""
    import numpy as np

    from oracle import oracle as O

    acc = accuracy(args, bld, snap, xyzm) if args.accuracy_sample > 0 else None
    if not primary:
        return None, acc
    cores = cpu_threads(args)
    os.environ["OMP_NUM_THREADS"] = str(cores)
    mode = "B" if args.mode == "tsc" else "A"
    name = "INTERPOLATE_DENSITY_TSC" if mode == "B" else "DENSITY"
    info = {"cpu_model": cpu_model(), "cores_available": len(os.sched_getaffinity(0)), "threads": cores}
    if n <= (1 << 25):
        import types
        host = [a.cpu().numpy() for a in xyzm]
        sn = types.SimpleNamespace(x=host[0], y=host[1], z=host[2], m=host[3],
                                   npart_esp=snap.npart_esp, geom=snap.geom)
        t0 = time.perf_counter()
        c1, _ = O.density32(sn, mode, nthreads=cores)
        tc = time.perf_counter() - t0
        ug = u1_gpu.cpu().numpy()
        sel = c1 > c1.mean()
        diff = float(np.max(np.abs(ug[sel] - c1[sel]) / c1[sel]))
        log(f"[bench] cpu baseline: {tc:.2f}s on {cores} threads")
        return dict(info, value=n / tc, unit="particles/s", cores=cores, kind="port",
                    sample=f"the whole build ({n} particles, {P} patches) by the FP32 OpenMP "
                           f"restatement of {name}",
                    seconds=tc, max_rel_diff_base_fp32_restatement_vs_gpu=diff), acc

    rng = np.random.default_rng(snap.spec.seed if snap.spec else 0)

    def timed(k):
        sn, _ = sample_snapshot(snap, xyzm, k, rng)
        t0 = time.perf_counter()
        O.density32(sn, mode, nthreads=cores)
        return time.perf_counter() - t0

    k0 = 1 << 15
    ks, ts = [k0, 2 * k0], []
    ts += [timed(k0), timed(2 * k0)]
    slope0 = max((ts[1] - ts[0]) / k0, 1e-9)
    # the largest sample spends about cpu_budget seconds on per-particle work
    ka = int(min(max(2 * k0, args.cpu_budget / slope0 / 2), n // 8))
    for k in (ka, 2 * ka):
        ks.append(k)
        ts.append(timed(k))
    A = np.stack([np.ones(len(ks)), np.array(ks, np.float64)], axis=1)
    (fixed, slope), *_ = np.linalg.lstsq(A, np.array(ts), rcond=None)
    if slope <= 0:
        fixed, slope = 0.0, ts[-1] / ks[-1]
    fixed = max(float(fixed), 0.0)
    t_full = fixed + n * slope
    resid = float(np.max(np.abs(A @ np.array([fixed, slope]) - np.array(ts))))
    log(f"[bench] cpu baseline: samples {ks} took {[round(t, 2) for t in ts]} s on {cores} threads; "
        f"extrapolated whole build {t_full:.0f}s")
    return dict(info, value=n / t_full, unit="particles/s", cores=cores, kind="port",
                sample=f"extrapolated: seeded samples of {ks} of the {n} particles, each built with the "
                       f"full {P}-patch hierarchy by the FP32 OpenMP restatement of {name}; time fitted "
                       f"as T0 + k*t (least squares) and evaluated at k = {n}",
                seconds_samples=ts, sample_particles=ks, seconds_extrapolated=t_full, fixed_s=fixed,
                per_particle_s=float(slope), per_particle_patch_s=float(slope) / max(P, 1),
                fit_max_residual_s=resid), acc


def main() -> int:
    args = parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the density build of one snapshot runs on a single "
              "MI355X (multi-GPU sharding of the build is not implemented); "
              "run it with --gpus 1", file=sys.stderr)
        return 2

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

    sc = stream_copy_gbs(torch)
    log(f"[bench] stream copy: {sc['gbs']:.0f} GB/s")
    out = run_workload(args.config, args, primary=True)
    out["stream_copy"] = sc
    out["roofline"]["frac_of_stream_copy"] = out["roofline"]["achieved"] / sc["gbs"]
    if args.secondary and args.secondary != args.config:
        sec = run_workload(args.secondary, args, primary=False)
        keep = ("value", "ms_per_step", "config", "roofline", "cold_call_ms", "checks", "graph", "accuracy")
        out["secondary"] = {k: sec[k] for k in keep if k in sec}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
