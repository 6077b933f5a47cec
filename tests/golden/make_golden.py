"""Writes the golden fixtures of the density build (run in a container that has
the reference tree; the fixtures themselves travel, the reference does not).

Each fixture is an .npz holding the inputs of one density build (particles,
NPART_ESP, geometry) and the fields the reference's own Fortran routines
produce on them, run through oracle/_ref/asohf_ref_density (the reference's
DENSITY and INTERPOLATE_DENSITY_TSC compiled from its sources, no OpenMP):
  u1_A, u11_A  -- DENSITY (grids.f:710-994), then /FDM (grids.f:2668-2722)
  u1_B, u11_B  -- INTERPOLATE_DENSITY_TSC (grids.f:1070-1398), then /FDM
  dx, radx     -- MALLA's constants (grids.f:2509-2573)

Cases:
  seeded_*   -- small clustered snapshots from asohf_amd.synth (sides <= 32 so
                they fit the reference's NAMRX = 32);
  edges      -- hand-placed particles on base-cell faces, on the box faces
                +-LADO0/2, on both patch-box faces of both modes (and one ulp
                either side), and on fine-cell faces inside patches, with every
                species present so every kernel width occurs;
  seeded_14_fdm -- 64,000 particles on a 64^3 base grid with four AMR
                levels, divided by the reference's default FDM = 0.845;
  seeded_15_128 -- 262,144 particles on a 128^3 base grid (the reference
                build's NMAX), four AMR levels, about 60 patches, FDM 0.845.

Usage: python tests/golden/make_golden.py [names]   (needs `make -C oracle ref`)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from asohf_amd import synth  # noqa: E402
from asohf_amd.geometry import Geometry, background, f32  # noqa: E402
from oracle import oracle as O  # noqa: E402


def save(name: str, snap) -> None:
    g = snap.geom
    out = dict(x=snap.x, y=snap.y, z=snap.z, m=snap.m, npart_esp=snap.npart_esp,
               nx=g.nx, ny=g.ny, nz=g.nz, lado0=g.lado0, rote=g.rote, rete=g.rete, fdm=g.fdm,
               interp_degree=g.interp_degree, npatch=g.npatch, pnx=g.pnx, pny=g.pny, pnz=g.pnz,
               prx=g.prx, pry=g.pry, prz=g.prz)
    for mode in "AB":
        dx, dy, dz, radx, u1, u11 = O.run_ref(snap, mode)
        out[f"u1_{mode}"] = u1
        out[f"u11_{mode}"] = u11
        out["dx"] = np.array([dx, dy, dz], np.float32)
        out["radx"] = radx
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(name, snap.n, "particles", g.npatch_total, "patches")


def nudge(v, k):
    """v moved by k ulps (k may be negative)."""
    v = np.float32(v)
    for _ in range(abs(k)):
        v = np.nextafter(v, np.float32(np.inf if k > 0 else -np.inf), dtype=np.float32)
    return v


def edges_case():
    nx = 16
    lado0 = f32(16.0)
    rote, rete = background(0.7, 0.3, 0.0)
    g0 = Geometry(nx, nx, nx, lado0, rote, rete, np.zeros(1, np.int32), [], [], [], [], [], [])
    radx, dx = g0.radx, g0.dx
    # level 1: two patches (one overlapping the other), level 2: two children
    l1 = [(4, 5, 6, 12, 10, 8), (7, 7, 4, 14, 12, 16)]           # (i1, j1, k1, n1, n2, n3)
    prx = [radx[i] for i, j, k, *_ in l1]
    pry = [radx[j] for i, j, k, *_ in l1]
    prz = [radx[k] for i, j, k, *_ in l1]
    pn = [(n1, n2, n3) for *_, n1, n2, n3 in l1]
    dxp1 = f32(dx / f32(2.0))
    l2 = [(0, 2, 2, 3, 8, 8, 6), (1, 5, 4, 6, 10, 8, 12)]        # (parent, i1, j1, k1, n1, n2, n3)
    for par, i1, j1, k1, n1, n2, n3 in l2:
        prx.append(f32(prx[par] + (f32(i1) - f32(1.5)) * dxp1))
        pry.append(f32(pry[par] + (f32(j1) - f32(1.5)) * dxp1))
        prz.append(f32(prz[par] + (f32(k1) - f32(1.5)) * dxp1))
        pn.append((n1, n2, n3))
    pn = np.array(pn)
    geom = Geometry(nx, nx, nx, lado0, rote, rete, np.array([0, 2, 2], np.int32),
                    pn[:, 0], pn[:, 1], pn[:, 2], prx, pry, prz,
                    pare=np.array([0, 0, 1, 2], np.int32))
    rng = np.random.default_rng(7)
    pts = []
    half = f32(lado0 / f32(2.0))
    xl = -half
    # base-grid faces and the box faces
    for k in range(0, nx + 1):
        face = f32(xl + f32(k) * dx)
        for d in (-1, 0, 1):
            pts.append((nudge(face, d), rng.uniform(-7, 7), rng.uniform(-7, 7)))
    pts += [(xl, xl, xl), (nudge(half, -1), nudge(half, -1), nudge(half, -1)),
            (xl, nudge(half, -1), f32(0.0))]
    # patch-box faces of both modes, and fine-cell faces inside each patch
    levels = geom.level_of_patch()
    for p in range(geom.npatch_total):
        ir = int(levels[p])
        dxpa = f32(dx / f32(2.0) ** ir)
        xlfix = f32(geom.prx[p] - dxpa)
        faces = [f32(xlfix - dx), f32(f32(xlfix + f32(geom.pnx[p]) * dxpa) + dx),       # mode A
                 f32(geom.prx[p] - f32(2.0) * dxpa),
                 f32(f32(geom.prx[p] - f32(2.0) * dxpa) + f32(geom.pnx[p] + 2) * dxpa)]  # mode B
        faces += [f32(xlfix + f32(i) * dxpa) for i in range(0, int(geom.pnx[p]) + 1, 3)]
        yc = f32(geom.pry[p] + f32(geom.pny[p] / 2) * dxpa)
        zc = f32(geom.prz[p] + f32(geom.pnz[p] / 2) * dxpa)
        for fce in faces:
            for d in (-1, 0, 1):
                pts.append((nudge(fce, d), yc, zc))
                pts.append((yc, nudge(fce, d), zc))
                pts.append((zc, yc, nudge(fce, d)))
    # a clustered cloud inside the level-2 patches
    for p in (2, 3):
        c = np.array([geom.prx[p], geom.pry[p], geom.prz[p]], np.float64) + 1.0
        pts += [tuple(v) for v in c + rng.normal(0, 0.6, size=(300, 3))]
    pts = np.array(pts, np.float32)
    pts = np.where(pts >= half, -half, pts).astype(np.float32)
    n = len(pts)
    m = (f32(1.0) + rng.random(n).astype(np.float32)) * f32(1e-6)
    # every species present: contiguous blocks of decreasing size
    esp = np.zeros(9, np.int32)
    left = n
    for s in range(9):
        take = left if s == 8 else max(left // 3, 1)
        esp[s] = take
        left -= take
    perm = rng.permutation(n)
    return synth.Snapshot(pts[perm, 0].copy(), pts[perm, 1].copy(), pts[perm, 2].copy(),
                          m.astype(np.float32), esp, geom)


def main() -> None:
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    only = set(sys.argv[1:])

    def want(name):
        return not only or name in only
    for seed in (11, 12):
        if want(f"seeded_{seed}"):
            snap = synth.make_snapshot(synth.WORKLOADS[1], n_side=20, nx=16, patches_per_level=(3, 3, 2),
                                       side_hi=32, seed=seed, fdm=1.0)
            save(f"seeded_{seed}", snap)
    if want("seeded_13_cic"):
        snap = synth.make_snapshot(synth.WORKLOADS[1], n_side=32, nx=32, patches_per_level=(5, 4, 3),
                                   side_hi=32, seed=13, interp_degree=1, fdm=1.0)
        save("seeded_13_cic", snap)
    if want("edges"):
        save("edges", edges_case())
    # RENORM_DENSITY at the reference's default FDM (asohf.dat:11), on a
    # larger case: 64,000 particles, a 64^3 base grid, four AMR levels
    if want("seeded_14_fdm"):
        snap = synth.make_snapshot(synth.WORKLOADS[1], n_side=40, nx=64, patches_per_level=(14, 10, 8, 5),
                                   side_hi=32, seed=14, fdm=0.845)
        save("seeded_14_fdm", snap)
    # the planned size of the reference's parity fixtures: 64^3 particles on a
    # 128^3 base grid (the reference build's NMAX), about 60 patches
    # (98% of the particles in haloes, so most base cells stay empty and the
    # fixture compresses to a few MB)
    if want("seeded_15_128"):
        import dataclasses
        spec = dataclasses.replace(synth.WORKLOADS[1], f_halo=0.98)
        snap = synth.make_snapshot(spec, n_side=64, nx=128, patches_per_level=(20, 16, 14, 10),
                                   side_hi=32, seed=15, fdm=0.845)
        save("seeded_15_128", snap)


if __name__ == "__main__":
    main()
