"""Random AMR hierarchies with many same-level overlaps, built by the
reference's placement rules: a patch of even sides starts on its parent's
cell PATCHX (base cell on level 1), PATCHRX = RADX(PATCHX) on level 1
(grids.f:300-302) and PATCHRX(parent) + (PATCHX - 1.5) * DXPA(parent) below
(grids.f:586-588), and its mother cells lie inside the parent.  Used by the
mask and dump tests."""
import types

import numpy as np

from asohf_amd.geometry import Geometry, background

f32 = np.float32


def random_hierarchy(rng, nx=64, counts=(12, 14, 10), side=(4, 16), lado0=40.0, n_particles=3000):
    """A snapshot-like namespace (geom, x, y, z, m, npart_esp) whose patches
    overlap heavily within each level."""
    rote, rete = background(0.7, 0.3, 0.0)
    g0 = Geometry(nx, nx, nx, f32(lado0), rote, rete, np.zeros(1, np.int32), *([np.zeros(0)] * 6))
    radx, dx = g0.radx, g0.dx
    pn, pr, pc, pare = [], [], [], []
    level_start = [0]
    for lev, cnt in enumerate(counts, start=1):
        lo = level_start[-1]
        for _ in range(cnt):
            half = rng.integers(side[0], side[1] + 1, 3)       # mother cells per axis
            if lev == 1:
                i1 = [int(rng.integers(2, nx - h)) for h in half]
                r = [radx[i] for i in i1]
                parent = 0
            else:
                prev = range(level_start[-2] if len(level_start) > 1 else 0, lo)
                q = int(rng.choice(list(prev)))
                half = np.minimum(half, np.array(pn[q]))
                i1 = [int(rng.integers(1, pn[q][a] - half[a] + 2)) for a in range(3)]
                dxp = f32(dx / f32(2.0) ** (lev - 1))
                r = [f32(pr[q][a] + (f32(i1[a]) - f32(1.5)) * dxp) for a in range(3)]
                parent = q + 1
            pn.append([int(2 * h) for h in half])
            pr.append([f32(v) for v in r])
            pc.append(i1)
            pare.append(parent)
        level_start.append(len(pn))
    pn = np.array(pn, np.int32)
    pr = np.array(pr, np.float32)
    pc = np.array(pc, np.int32)
    npatch = np.array([0] + list(counts), np.int32)
    geom = Geometry(nx, nx, nx, f32(lado0), rote, rete, npatch, pn[:, 0], pn[:, 1], pn[:, 2],
                    pr[:, 0], pr[:, 1], pr[:, 2], pare=np.array(pare, np.int32),
                    pcx=pc[:, 0], pcy=pc[:, 1], pcz=pc[:, 2])
    half_box = f32(lado0 / 2)
    pos = rng.uniform(-half_box, half_box, (n_particles, 3)).astype(np.float32)
    pos[pos >= half_box] = -half_box
    m = np.full(n_particles, f32(1e-3), np.float32)
    esp = np.zeros(9, np.int32)
    esp[0] = n_particles
    return types.SimpleNamespace(geom=geom, x=pos[:, 0].copy(), y=pos[:, 1].copy(), z=pos[:, 2].copy(),
                                 m=m, npart_esp=esp)
