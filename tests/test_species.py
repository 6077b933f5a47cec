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
    yield "seeded_c1", synth.make_snapshot(synth.CONFIGS[1], n_side=48, nx=64, patches_per_level=(3, 2, 1), seed=5)
    yield "seeded_c1_b", synth.make_snapshot(synth.CONFIGS[1], n_side=40, nx=128, patches_per_level=(2, 1), seed=9)


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
    big = synth.make_snapshot(synth.CONFIGS[2])
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
