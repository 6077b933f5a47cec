"""Species by local density (SORT_DM_PARTICLES_LOCALDENSITY, reader.f:140-336).

CPU: the oracle's FP32 restatement against the reference routine itself
(compiled from its sources into oracle/_ref), bit-exact: the class counts
NPART_ESP and the whole permutation.  GPU: asohf_hip_species against the
oracle, bit-exact (classes, counts, permutation) at sizes the oracle runs in
seconds, and the partition's invariants at the benchmark's full size.
"""
import numpy as np
import pytest

from asohf_amd import synth
from asohf_amd.geometry import background
from oracle import oracle as O

NMAX = 128   # the reference's compile-time NMAX = NMAY = NMAZ


def consts():
    rodo, re0 = background(0.7, 0.3, 0.0)     # ROTE, RETE at z = 0 are RODO, RE0
    return float(rodo), float(re0)


def snapshots():
    yield "seeded_c1", synth.make_snapshot(synth.WORKLOADS[1], n_side=48, nx=64, patches_per_level=(3, 2, 1), seed=5)
    yield "seeded_c1_b", synth.make_snapshot(synth.WORKLOADS[1], n_side=40, nx=128, patches_per_level=(2, 1), seed=9)


def edge_particles():
    """Particles on cell faces, at the box edges, beyond them (clamped), and
    piled into single cells so densities span several classes."""
    rng = np.random.default_rng(3)
    L = np.float32(40.0)
    dx = np.float32(L / 128)
    faces = (np.arange(-64, 64, dtype=np.float32) * dx)
    x = np.concatenate([faces, np.full(64, -L / 2, np.float32), np.full(64, L / 2 - dx / 4, np.float32),
                        np.full(32, L, np.float32), rng.uniform(-L / 2, L / 2, 4000).astype(np.float32),
                        np.full(3000, np.float32(0.01)), np.full(400, np.float32(5.003))])
    y = np.concatenate([faces[::-1], rng.uniform(-L / 2, L / 2, len(x) - 128).astype(np.float32)])
    z = rng.uniform(-L / 2, L / 2, len(x)).astype(np.float32)
    z[-3400:] = np.float32(-7.77)
    m = np.full(len(x), np.float32(1.09e-9), np.float32)
    m[::7] *= np.float32(3.0)
    return x, y.astype(np.float32), z, m, float(dx)


def threshold_cells(big_first=True, seed=4):
    """A cell whose FP32 density lands on either side of a class threshold
    depending on the summation order: one heavy particle at 0.9999 of 8^3*BAS
    and 10000 light ones, each below half an ulp of it.  Summed in particle
    order with the heavy one first, the light masses round away (class 2);
    summed in any order that adds light masses together first, the cell
    crosses 8^3 (class 3).  Further cells hold runs of 31..2049 particles of
    random masses (run lengths on both sides of the GPU's short/long and
    block boundaries), interleaved with uniform particles in index order."""
    rodo, re0 = consts()
    rng = np.random.default_rng(seed)
    L = np.float32(40.0)
    dx = np.float32(L / 128)
    bas = np.float32(np.float32(np.float32(dx * dx) * dx) * np.float32(rodo)) * np.float32(re0) ** 3
    centre = lambda i, j, k: [np.float32((c + 0.5) * dx - L / 2) for c in (i, j, k)]
    big = np.float32(512.0) * bas * np.float32(0.9999)
    eps = np.float32(np.spacing(big) * np.float32(0.4))
    xs, ms = [], []
    cx = centre(100, 20, 77)
    xs.append(np.tile(np.array(cx, np.float32), (10001, 1)))
    cell = np.full(10001, eps, np.float32)
    cell[0 if big_first else -1] = big
    ms.append(cell)
    for q, n_run in enumerate((31, 32, 33, 64, 1023, 1024, 1025, 2049)):
        xs.append(np.tile(np.array(centre(10 + 9 * q, 90, 40), np.float32), (n_run, 1)))
        ms.append((bas * np.float32(8.0) ** rng.uniform(0.5, 4.5) / np.float32(n_run)
                   * rng.uniform(0.5, 1.5, n_run)).astype(np.float32))
    xs.append(rng.uniform(-L / 2, L / 2, (5000, 3)).astype(np.float32))
    ms.append(np.full(5000, bas * np.float32(0.3), np.float32))
    pos = np.concatenate(xs)
    m = np.concatenate(ms)
    # interleave: random new indices, each cell keeping its own particles' order
    perm = rng.permutation(len(m))
    keys = [np.floor((pos[:, a].astype(np.float64) + L / 2) / dx).astype(np.int64) for a in range(3)]
    flat = keys[0] + 128 * (keys[1] + 128 * keys[2])
    new_index = np.empty(len(m), np.int64)
    for c in np.unique(flat):
        members = np.nonzero(flat == c)[0]                # original order
        new_index[members] = np.sort(perm[members])       # that many random slots, ascending
    out = np.empty((len(m), 3), np.float32)
    mo = np.empty(len(m), np.float32)
    out[new_index] = pos
    mo[new_index] = m
    return out[:, 0].copy(), out[:, 1].copy(), out[:, 2].copy(), mo, float(dx)


def test_oracle_threshold_cell_is_order_sensitive():
    """The threshold cell's class depends on the summation order, so the GPU
    comparison on it checks the order of the FP32 sums."""
    rodo, re0 = consts()
    for big_first, want in ((True, 2), (False, 3)):
        x, y, z, m, dx = threshold_cells(big_first)
        sp, esp, order = O.species32(x, y, z, m, NMAX, dx, rodo, re0)
        heavy = int(np.argmax(m))
        assert sp[heavy] == want, (big_first, sp[heavy])
        cell = (x == x[heavy]) & (y == y[heavy]) & (z == z[heavy])
        assert cell.sum() == 10001 and np.all(sp[cell] == want)
        # the light masses are interleaved with other particles in index order
        idx = np.nonzero(cell)[0]
        assert idx[-1] - idx[0] > 10001
    # FP64 (any order) puts the cell in class 3: only the sequential FP32 sum gives 2
    x, y, z, m, dx = threshold_cells(True)
    heavy = int(np.argmax(m))
    cell = (x == x[heavy]) & (y == y[heavy]) & (z == z[heavy])
    bas = np.float64(np.float32(np.float32(np.float32(np.float32(dx * dx) * dx) * np.float32(rodo)) * np.float32(re0) ** 3))
    assert np.floor(np.log(m[cell].astype(np.float64).sum() / bas) / np.log(8.0)) == 3


@pytest.mark.skipif(not O.ref_available(), reason="the reference build (oracle/_ref) is not present")
def test_oracle_matches_reference_routine():
    rodo, re0 = consts()
    for name, snap in snapshots():
        dx, esp_r, order_r = O.run_ref_species(snap.x, snap.y, snap.z, snap.m, snap.geom.nx,
                                               float(snap.geom.lado0), rodo, re0)
        assert dx == float(snap.geom.dx)
        sp, esp_o, order_o = O.species32(snap.x, snap.y, snap.z, snap.m, NMAX, dx, rodo, re0)
        assert np.array_equal(esp_o, esp_r), name
        assert np.array_equal(order_o, order_r), name
        assert (esp_o > 0).sum() >= 3, name    # several classes exercised


@pytest.mark.skipif(not O.ref_available(), reason="the reference build (oracle/_ref) is not present")
def test_oracle_matches_reference_on_edges():
    rodo, re0 = consts()
    x, y, z, m, dx = edge_particles()
    dxr, esp_r, order_r = O.run_ref_species(x, y, z, m, 128, 40.0, rodo, re0)
    assert dxr == np.float32(dx)
    _, esp_o, order_o = O.species32(x, y, z, m, NMAX, dxr, rodo, re0)
    assert np.array_equal(esp_o, esp_r)
    assert np.array_equal(order_o, order_r)


def test_oracle_partition_invariants():
    """Pins what the mathematics fixes: a stable partition (indices increase
    within a class), counts that add up, and classes that follow the density
    thresholds 8^k of a brute-force NGP count."""
    rodo, re0 = consts()
    snap = next(iter(snapshots()))[1]
    dx = float(snap.geom.dx)
    sp, esp, order = O.species32(snap.x, snap.y, snap.z, snap.m, NMAX, dx, rodo, re0)
    assert esp.sum() == snap.n and np.array_equal(np.bincount(sp, minlength=9), esp)
    assert np.array_equal(np.sort(order), np.arange(snap.n))
    bounds = np.concatenate([[0], np.cumsum(esp)])
    for s in range(9):
        seg = order[bounds[s]:bounds[s + 1]]
        assert np.all(np.diff(seg) > 0) and np.all(sp[seg] == s)
    # brute force in FP64: density of each particle's cell, class = floor(log8)
    xmin = -NMAX * dx / 2
    idx = [np.clip(((v.astype(np.float64) - xmin) / dx).astype(np.int64), 0, NMAX - 1) for v in (snap.x, snap.y, snap.z)]
    flat = idx[0] + NMAX * (idx[1] + NMAX * idx[2])
    bas = dx ** 3 * rodo * re0 ** 3
    d = np.bincount(flat, weights=snap.m.astype(np.float64), minlength=NMAX ** 3)[flat] / bas
    want = np.clip(np.floor(np.log(d) / np.log(8.0)), 0, 8).astype(np.int64)
    near = np.abs(np.log(d) / np.log(8.0) - np.round(np.log(d) / np.log(8.0))) < 1e-5
    assert np.array_equal(sp[~near], want[~near])


@pytest.mark.gpu
def test_gpu_species_matches_oracle():
    import torch
    from asohf_amd.density import DensityBuilder
    rodo, re0 = consts()
    b = DensityBuilder(0)
    cases = [(name, s.x, s.y, s.z, s.m, float(s.geom.dx)) for name, s in snapshots()]
    cases.append(("edges",) + edge_particles())
    cases.append(("threshold_cells",) + threshold_cells(True))
    cases.append(("threshold_cells_reversed",) + threshold_cells(False))
    big = synth.make_snapshot(synth.WORKLOADS[2])
    cases.append(("c2_full", big.x, big.y, big.z, big.m, float(big.geom.dx)))
    for name, x, y, z, m, dx in cases:
        print(f"case species {name} n={len(x)}", flush=True)
        t = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, z, m)]
        sp, order, esp = b.species(*t, nmax=(NMAX,) * 3, d=(dx,) * 3, rodo=rodo, re0=re0)
        torch.cuda.synchronize()
        sp_o, esp_o, order_o = O.species32(x, y, z, m, NMAX, dx, rodo, re0)
        assert np.array_equal(esp, esp_o), name
        assert np.array_equal(sp.cpu().numpy(), sp_o), name
        assert np.array_equal(order.cpu().numpy(), order_o), name
    b.close()
