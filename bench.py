#!/usr/bin/env python
"""Benchmark of the batched all-level density build on one MI355X.

One step = one density build (TSC weights on every AMR level, deterministic
fixed-point reduction, all grids in one persistent deposit launch) over a
synthetic clustered snapshot whose particles are already resident in HBM:
binning (count, scan, scatter), deposit, normalisation and FP64 grid sums are
all inside the timed region.

Prints ONE JSON line on stdout; progress goes to stderr.
  value      particles/s of the batched build
  roofline   the deposit phase (the persistent deposit kernel and split_reduce,
             which sums the partials of tiles split over several work items)
             against MI355X's 8 TB/s HBM peak: algorithmic bytes = 16 B per
             (particle, grid) pair + 4 B per output cell, over their average
             duration measured with HIP events on the stream they run on
  pipeline   the whole step against the same peak (B_fwd = 16 B per pair + 4 B
             per cell + 8 B per grid sum; B_min counts each particle once)
  modes      ms/step of every mode (TSC, kernel table) x reduction
             (deterministic, atomic) on the same snapshot
  per_level_ms  builds of the hierarchy cut at level L = 0..NL, differenced
  species    the species partition that feeds NPART_ESP (reader.f:140-336)
  peaks      the grid halo finder's peak candidates on every level of the
             field (haloes_grids.f:316-404, 700-790, SORT_CELLS_AMR order)
  create_mesh  the mesh builder CREATE_MESH on the same particles
  masks      the grid masks of the density dump (COMPUTE_CR0AMR, VEINSGRID)
             and the FW2 dump writer (asohf.f:515-545)
  adjoint    the hand-written backward (gradients of sum(w * rho))
  plain_gpu  the same build as one launch per grid with global FP32 atomics
  cpu_baseline  the oracle's OpenMP restatement of the reference algorithm
             (INTERPOLATE_DENSITY_TSC, FP32) on this host's cores; whole for
             the 2M/16.8M-particle workloads, extrapolated from two particle
             samples for the 134M/1.07G ones
Workloads (--config): asohf_amd.synth.WORKLOADS; 2 (256^3 particles, 256^3 base
grid, 4 AMR levels) is the default.
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
    ap.add_argument("--config", type=int, default=2, help="workload id (asohf_amd.synth.WORKLOADS)")
    ap.add_argument("--mode", default="tsc", choices=["tsc", "table"])
    ap.add_argument("--reduction", default="deterministic", choices=["deterministic", "atomic"])
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="seconds of CPU work per sample when the whole CPU build is too long")
    ap.add_argument("--modes-steps", type=int, default=5, help="steps per mode x reduction (0: skip)")
    ap.add_argument("--no-levels", dest="levels", action="store_false", help="skip the per-level timing")
    ap.add_argument("--species-steps", type=int, default=5, help="calls of the species partition (0: skip)")
    ap.add_argument("--peaks-steps", type=int, default=5,
                    help="calls of the peak-candidate scan over every level (0: skip)")
    ap.add_argument("--contrastec", type=float, default=200.0, help="CONTRASTEC of the peak-candidate scan")
    ap.add_argument("--mesh-steps", type=int, default=2, help="calls of the mesh builder (0: skip)")
    ap.add_argument("--masks-steps", type=int, default=3,
                    help="calls of the grid masks and the density dump writer (0: skip)")
    ap.add_argument("--plain-max-tests", type=float, default=5e12,
                    help="skip the plain form above this many particle-box tests")
    ap.add_argument("--tune", default="", help="tx,ty,tz,chunk,waves_per_cu[,block]: fixed tuning, no autotune")
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

    spec = synth.WORKLOADS[args.config]
    dev = torch.device("cuda:0")
    t0 = time.time()
    if spec.n > (1 << 25):
        # 134M / 1.07G particles: generated on the device (numpy is too slow)
        snap = synth.make_snapshot_device(spec, device=dev)
        x, y, z, m = snap.x, snap.y, snap.z, snap.m
    else:
        snap = synth.make_snapshot(spec)
        x, y, z, m = (torch.from_numpy(a).to(dev) for a in (snap.x, snap.y, snap.z, snap.m))
    n = int(x.numel())
    g = snap.geom
    log(f"[bench] snapshot {spec.name}: {n} particles, {g.npatch_total} patches, "
        f"{g.ncells_total} cells, generated in {time.time() - t0:.1f}s")
    bld = DensityBuilder(0)
    s1, n11 = bld.out_shapes(g)
    u1 = torch.empty(s1, dtype=torch.float32, device=dev)
    u11 = torch.empty(n11, dtype=torch.float32, device=dev)
    kw = dict(mode=args.mode, reduction=args.reduction, u1=u1, u11=u11)
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

    for _ in range(max(args.warmup, 1)):
        bld.density(g, x, y, z, m, snap.npart_esp, **kw)
    torch.cuda.synchronize()
    per = {"ms_bin": [], "ms_scatter": [], "ms_deposit": [], "ms_split": [], "ms_sums": [], "ms_total": []}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = bld.density(g, x, y, z, m, snap.npart_esp, **kw)
        st = bld.stats()
        for k in per:
            per[k].append(st[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1000.0 * wall / args.steps
    value = n / (wall / args.steps)
    avg = {k: float(np.mean(v)) for k, v in per.items()}
    med = {k: float(np.median(v)) for k, v in per.items()}
    n_pairs, n_cells, n_entries = st["n_pairs"], st["n_cells"], st["n_entries"]
    P = g.npatch_total
    b_fwd = 16.0 * n_pairs + 4.0 * n_cells + 8.0 * (P + 1)
    b_min = 16.0 * n + 4.0 * n_cells + 8.0 * (P + 1)
    b_dep = 16.0 * n_pairs + 4.0 * n_cells
    dep_ms = avg["ms_deposit"] + avg["ms_split"]
    achieved = b_dep / (dep_ms * 1e-3) / 1e9
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
        "config": {"workload": spec.name, "particles": n, "base_grid": g.nx, "levels": g.nl,
                   "patches": P, "cells": int(n_cells), "mode": args.mode, "reduction": args.reduction,
                   "tune": tune},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "deposit_kernel + split_reduce (the deposit phase)",
                     "algorithmic_bytes": b_dep, "avg_ms": dep_ms,
                     "avg_ms_deposit_kernel": avg["ms_deposit"], "avg_ms_split_reduce": avg["ms_split"]},
        "pipeline": {"algorithmic_bytes": b_fwd, "achieved_gbs": b_fwd / (ms_step * 1e-3) / 1e9,
                     "frac": b_fwd / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "b_min": b_min, "b_min_frac": b_min / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "phase_ms": avg, "phase_ms_median": med,
                     "n_pairs": int(n_pairs), "n_entries": int(n_entries)},
    }
    u1_gpu = r.u1.cpu().numpy()
    # size-independent check of the timed result: every particle's mass is
    # deposited exactly once on the base grid (FP64 grid sum vs FP64 total)
    total = float(m.double().sum())
    out["checks"] = {"base_grid_mass_rel_err": abs(float(r.grid_mass[0]) - total) / total if total else 0.0}

    if args.modes_steps > 0:
        # every mode x reduction on the same snapshot (the headline line above is one of them)
        table = {}
        for md in ("tsc", "table"):
            for rd in ("deterministic", "atomic"):
                k2 = dict(kw, mode=md, reduction=rd)
                bld.density(g, x, y, z, m, snap.npart_esp, **k2)
                torch.cuda.synchronize()
                dd = []
                t0 = time.perf_counter()
                for _ in range(args.modes_steps):
                    bld.density(g, x, y, z, m, snap.npart_esp, **k2)
                    st2 = bld.stats()
                    dd.append(st2["ms_deposit"] + st2["ms_split"])
                torch.cuda.synchronize()
                tw = (time.perf_counter() - t0) / args.modes_steps
                s2 = bld.stats()
                table[f"{md}/{rd}"] = {"ms_per_step": 1000 * tw, "particles_per_s": n / tw,
                                      "deposit_ms": float(np.mean(dd)), "n_entries": int(s2["n_entries"])}
        out["modes"] = table
        log(f"[bench] modes: {table}")
        bld.density(g, x, y, z, m, snap.npart_esp, **kw)
        torch.cuda.synchronize()

    if args.levels and g.nl > 0:
        # time per AMR level: builds of the hierarchy cut at level L, differenced
        cum = []
        for L in range(0, g.nl + 1):
            gl = g.truncated(L)
            bld.density(gl, x, y, z, m, snap.npart_esp, mode=args.mode, reduction=args.reduction)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(max(args.steps // 4, 3)):
                bld.density(gl, x, y, z, m, snap.npart_esp, mode=args.mode, reduction=args.reduction)
            torch.cuda.synchronize()
            cum.append(1000 * (time.perf_counter() - t0) / max(args.steps // 4, 3))
        out["per_level_ms"] = {"cumulative": cum,
                               "level": [cum[0]] + [cum[i] - cum[i - 1] for i in range(1, len(cum))],
                               "patches": [int(v) for v in g.npatch]}
        log(f"[bench] per level (cumulative ms): {cum}")
        bld.density(g, x, y, z, m, snap.npart_esp, **kw)
        torch.cuda.synchronize()

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
        if not args.no_cpu and n <= (1 << 25):
            from oracle import oracle as O
            hx = [a.cpu().numpy() if not isinstance(a, np.ndarray) else a for a in (x, y, z, m)]
            t0 = time.perf_counter()
            _, esp_o, _ = O.species32(*hx, g.nx, float(g.dx), rodo, re0)
            tc = time.perf_counter() - t0
            out["species"]["cpu_baseline"] = {"seconds": tc, "particles_per_s": n / tc, "cores": 1,
                                              "kind": "port", "same_counts": bool(np.array_equal(esp_o, esp))}
        log(f"[bench] species: {1000 * ts:.3f} ms/call")
        del order

    if args.peaks_steps > 0:
        # the grid halo finder's peak candidates on every level of the field
        # just built (HALOFIND_GRID, haloes_grids.f:316-404 and 700-790, in
        # SORT_CELLS_AMR order); one call per level
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
        if not args.no_cpu and n_cells <= (1 << 26):
            from oracle import oracle as O
            hu1 = r.u1.cpu().numpy().reshape(g.nz, g.ny, g.nx)
            hu11 = r.u11.cpu().numpy()
            offs = np.concatenate([[0], np.cumsum(g.patch_cells())])
            grids, q = [[hu1]], 0
            for lev in range(1, g.nl + 1):
                grids.append([hu11[offs[p]:offs[p + 1]].reshape(g.pnz[p], g.pny[p], g.pnx[p])
                              for p in range(q, q + int(g.npatch[lev]))])
                q += int(g.npatch[lev])
            t0 = time.perf_counter()
            cpu = [O.peaks32(fl, args.contrastec) if fl else (np.zeros(0),) for fl in grids]
            tc = time.perf_counter() - t0
            same = all(np.array_equal(c[0], gp[0].cpu().numpy()) and
                       (len(c) == 1 or np.array_equal(c[1], gp[1].cpu().numpy())) for c, gp in zip(cpu, res))
            out["peaks"]["cpu_baseline"] = {"seconds": tc, "cells_per_s": n_cells / tc, "cores": 1,
                                            "kind": "port", "same_lists": bool(same)}
        log(f"[bench] peaks: {1000 * tp:.3f} ms/call over {g.nl + 1} levels, candidates {counts}")
        del res

    if args.mesh_steps > 0 and g.ncells_base > (1 << 26):
        out["create_mesh"] = {"skipped": f"base grid of {g.ncells_base} cells: the host growth walks every "
                                         "refinable base cell; the bench times it up to 2^26 cells"}
    elif args.mesh_steps > 0:
        # the mesh builder CREATE_MESH (grids.f:2-707) on the same particles,
        # asohf.dat's mesh block, as many levels as the workload has
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
        if not args.no_cpu and n_cells <= (1 << 26):
            from oracle import oracle as O
            t0 = time.perf_counter()
            oc0, oc11 = O.cr0amr(g)
            os_ = O.veinsgrid(g)
            tc = time.perf_counter() - t0
            out["masks"]["cpu_baseline"] = {
                "seconds": tc, "cores": 1, "kind": "port",
                "same_masks": bool(np.array_equal(oc0, c0.cpu().numpy()) and np.array_equal(oc11, c11.cpu().numpy())
                                   and np.array_equal(os_, so.cpu().numpy()))}
        log(f"[bench] masks: CR0AMR {1000 * t_cr:.3f} ms, SOLAP {1000 * t_vg:.3f} ms, "
            f"dump {size / 1e6:.1f} MB in {t_w:.3f} s")
        del c0, c11, so

    # the forward workspaces (sized by the largest mode above) are not needed by
    # the adjoint; free them so the adjoint's own buffers fit at any size
    bld.release_workspaces()
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
        b_adj = 16.0 * n_pairs + 4.0 * n_cells + 16.0 * n
        out["adjoint"] = {"ms_per_step": aw, "ratio_to_forward": aw / ms_step,
                          "algorithmic_bytes": b_adj, "achieved_gbs": b_adj / (aw * 1e-3) / 1e9}
        log(f"[bench] adjoint: {aw:.3f} ms/step ({aw / ms_step:.2f}x forward)")
        del w1, w11

    if not args.no_plain:
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

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, snap, (x, y, z, m), u1_gpu, n, P)
    print(json.dumps(out), flush=True)
    bld.close()
    return 0


def cpu_baseline(args, snap, dev_xyzm, u1_gpu, n, P):
    """The oracle's OpenMP restatement of the reference build on this host.

    Small workloads are built whole. For the large ones (whose CPU build is
    hours: every particle is tested against every patch box of a level), two
    seeded particle samples of the same snapshot are built with the full
    hierarchy, T(k) = T0 + k * t is fitted through them, and the rate is
    N / (T0 + N * t), labelled as extrapolated."""
    import numpy as np
    import types

    from oracle import oracle as O

    cores = args.cpu_threads or min(len(os.sched_getaffinity(0)), 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    mode = "B" if args.mode == "tsc" else "A"
    name = "INTERPOLATE_DENSITY_TSC" if mode == "B" else "DENSITY"
    x, y, z, m = dev_xyzm
    if n <= (1 << 25):
        host = [a if isinstance(a, np.ndarray) else a.cpu().numpy() for a in (snap.x, snap.y, snap.z, snap.m)]
        sn = types.SimpleNamespace(x=host[0], y=host[1], z=host[2], m=host[3],
                                   npart_esp=snap.npart_esp, geom=snap.geom)
        t0 = time.perf_counter()
        c1, _ = O.density32(sn, mode, nthreads=cores)
        tc = time.perf_counter() - t0
        sel = c1 > c1.mean()
        diff = float(np.max(np.abs(u1_gpu[sel] - c1[sel]) / c1[sel])) if u1_gpu is not None else None
        log(f"[bench] cpu baseline: {tc:.2f}s on {cores} threads")
        return {"value": n / tc, "unit": "particles/s", "cores": cores, "kind": "port",
                "sample": f"the whole build ({n} particles, {P} patches) by the FP32 OpenMP "
                          f"restatement of {name}",
                "seconds": tc, "max_rel_diff_base_vs_gpu": diff}

    import torch
    rng = np.random.default_rng(snap.spec.seed if snap.spec else 0)
    bounds = np.concatenate([[0], np.cumsum(snap.npart_esp.astype(np.int64))])

    def sample(k):
        idx = np.sort(rng.choice(n, size=k, replace=False))
        ti = torch.from_numpy(idx).to(x.device)
        arrs = [a[ti].cpu().numpy() for a in (x, y, z, m)]
        esp = np.bincount(np.searchsorted(bounds, idx, side="right") - 1, minlength=len(snap.npart_esp))
        return types.SimpleNamespace(x=arrs[0], y=arrs[1], z=arrs[2], m=arrs[3],
                                     npart_esp=esp[: len(snap.npart_esp)].astype(np.int32), geom=snap.geom)

    def timed(sn):
        t0 = time.perf_counter()
        O.density32(sn, mode, nthreads=cores)
        return time.perf_counter() - t0

    k0 = 4096
    t_probe = timed(sample(k0))
    ka = int(min(max(k0, k0 * args.cpu_budget / 2 / max(t_probe, 1e-3)), n // 4))
    ta = timed(sample(ka))
    tb = timed(sample(2 * ka))
    slope = (tb - ta) / ka
    fixed = max(ta - ka * slope, 0.0)
    if slope <= 0:
        slope, fixed = tb / (2 * ka), 0.0
    t_full = fixed + n * slope
    log(f"[bench] cpu baseline: samples {ka}/{2 * ka} particles took {ta:.2f}/{tb:.2f}s on {cores} "
        f"threads; extrapolated whole build {t_full:.0f}s")
    return {"value": n / t_full, "unit": "particles/s", "cores": cores, "kind": "port",
            "sample": f"extrapolated: two seeded samples of {ka} and {2 * ka} of the {n} particles, "
                      f"each built with the full {P}-patch hierarchy by the FP32 OpenMP restatement "
                      f"of {name}; time fitted as T0 + k*t and evaluated at k = {n}",
            "seconds_samples": [ta, tb], "seconds_extrapolated": t_full, "fixed_s": fixed,
            "per_particle_s": slope}


if __name__ == "__main__":
    sys.exit(main())
