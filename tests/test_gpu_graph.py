"""The sync-free and hipGraph forms of the batched build (GPU).

The first call on new sizes is exact (count pass, one host read, buffers grown,
deposit).  Later calls reuse those buffers with no host round trip before the
deposit; a gate kernel checks the capacities on the device and the host reruns
the exact form when they do not suffice.  With use_graph the whole sequence is
one hipGraph, captured on first use and replayed while every kernel argument
stays the same.

Bars: deterministic mode bit-identical across the exact, sync-free and graph
forms and across graph replays (with eager work between the replays);
every form within 1e-5 of the FP64 oracle on cells above the mean; atomic
mode within the FP32 accumulation tolerance of the parity tests.
"""
import numpy as np
import pytest
import torch

from asohf_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

MODES = {"table": "A", "tsc": "B"}


def small(seed=11, n_side=32):
    return synth.make_snapshot(synth.WORKLOADS[1], n_side=n_side, nx=32, patches_per_level=(4, 3, 2), seed=seed)


def to_dev(snap, k=None):
    d = torch.device("cuda:0")
    sl = slice(None) if k is None else slice(0, k)
    return tuple(torch.from_numpy(np.ascontiguousarray(a[sl])).to(d) for a in (snap.x, snap.y, snap.z, snap.m))


def rel_err_above_mean(got, ref):
    got = np.asarray(got, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    sel = ref > ref.mean()
    if not sel.any():
        return 0.0
    return float(np.max(np.abs(got[sel] - ref[sel]) / ref[sel]))


def oracle_err(snap, mode, u1, u11, k=None):
    if k is not None:
        import types
        esp = snap.npart_esp.copy()
        # the subset is a prefix: species counts of the first k particles
        bounds = np.concatenate([[0], np.cumsum(esp)])
        esp = np.array([max(0, min(k, bounds[i + 1]) - bounds[i]) for i in range(len(esp))], np.int32)
        snap = types.SimpleNamespace(x=snap.x[:k], y=snap.y[:k], z=snap.z[:k], m=snap.m[:k],
                                     npart_esp=esp, geom=snap.geom)
    d1, d11, _ = O.density64(snap, MODES[mode])
    return max(rel_err_above_mean(u1, d1), rel_err_above_mean(u11, d11[: u11.size]))


def fresh():
    from asohf_amd.density import DensityBuilder
    return DensityBuilder(0)


@pytest.mark.parametrize("mode", ["tsc", "table"])
def test_graph_replays_bit_identical(mode):
    snap = small()
    b = fresh()
    x, y, z, m = to_dev(snap)
    s1, n11 = b.out_shapes(snap.geom)
    dev = x.device
    u1 = torch.empty(s1, dtype=torch.float32, device=dev)
    u11 = torch.empty(n11, dtype=torch.float32, device=dev)
    gm = torch.empty(1 + snap.geom.npatch_total, dtype=torch.float64, device=dev)
    kw = dict(mode=mode, u1=u1, u11=u11, grid_mass=gm)
    print(f"case graph {mode} exact", flush=True)
    b.density(snap.geom, x, y, z, m, snap.npart_esp, **kw)
    torch.cuda.synchronize()
    ref = [t.cpu().numpy().copy() for t in (u1, u11, gm)]
    assert oracle_err(snap, mode, ref[0], ref[1]) <= 1e-5
    print(f"case graph {mode} sync-free", flush=True)
    b.density(snap.geom, x, y, z, m, snap.npart_esp, **kw)
    torch.cuda.synchronize()
    for a, r in zip((u1, u11, gm), ref):
        assert np.array_equal(a.cpu().numpy(), r)
    for rep in range(4):
        print(f"case graph {mode} replay {rep}", flush=True)
        u1.fill_(-1.0)
        u11.fill_(-1.0)
        gm.zero_()
        torch.cuda.synchronize()
        b.density(snap.geom, x, y, z, m, snap.npart_esp, use_graph=True, **kw)
        torch.cuda.synchronize()
        for a, r in zip((u1, u11, gm), ref):
            assert np.array_equal(a.cpu().numpy(), r), rep
        st = b.stats()
        assert st["ms_total"] > 0.0 and st["n_pairs"] > 0
    b.close()


def test_graph_atomic_and_side_stream():
    snap = small(seed=5)
    b = fresh()
    x, y, z, m = to_dev(snap)
    s1, n11 = b.out_shapes(snap.geom)
    u1 = torch.empty(s1, dtype=torch.float32, device=x.device)
    u11 = torch.empty(n11, dtype=torch.float32, device=x.device)
    gm = torch.empty(1 + snap.geom.npatch_total, dtype=torch.float64, device=x.device)
    side = torch.cuda.Stream()
    for rep in range(3):
        print(f"case graph atomic side-stream {rep}", flush=True)
        with torch.cuda.stream(side):
            u1.fill_(-1.0)
            b.density(snap.geom, x, y, z, m, mode="tsc", reduction="atomic", u1=u1, u11=u11,
                      grid_mass=gm, use_graph=rep > 0)
        torch.cuda.synchronize()
        assert oracle_err(snap, "tsc", u1.cpu().numpy(), u11.cpu().numpy()) <= 1e-4
    b.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_buffers_regrow_on_larger_input(use_graph):
    """A call larger than the one that sized the buffers overflows on the device
    and is rebuilt exactly; a smaller one after it fits."""
    snap = small(seed=3, n_side=40)
    n = len(snap.x)
    b = fresh()
    for k in (n // 4, n, n // 2, n):
        print(f"case regrow graph={use_graph} k={k}", flush=True)
        x, y, z, m = to_dev(snap, k)
        r = b.density(snap.geom, x, y, z, m, mode="tsc", use_graph=use_graph)
        torch.cuda.synchronize()
        u1, u11, gm = (t.cpu().numpy() for t in (r.u1, r.u11, r.grid_mass))
        assert oracle_err(snap, "tsc", u1, u11, k) <= 1e-5, k
        total = float(snap.m[:k].astype(np.float64).sum())
        assert abs(gm[0] - total) <= 1e-12 * total
        assert b.stats()["n_particles"] == k
    b.close()


def test_release_workspaces_then_rebuild():
    """Freeing the context's workspaces (and its graph) between calls: the
    next calls size them again and give the same bits."""
    snap = small(seed=9)
    b = fresh()
    x, y, z, m = to_dev(snap)
    r0 = b.density(snap.geom, x, y, z, m, mode="tsc")
    torch.cuda.synchronize()
    ref = [t.cpu().numpy().copy() for t in (r0.u1, r0.u11, r0.grid_mass)]
    for use_graph in (True, False, True):
        print(f"case release graph={use_graph}", flush=True)
        b.release_workspaces()
        for _ in range(2):
            r = b.density(snap.geom, x, y, z, m, mode="tsc", use_graph=use_graph,
                          u1=r0.u1, u11=r0.u11, grid_mass=r0.grid_mass)
            torch.cuda.synchronize()
            for a, want in zip((r.u1, r.u11, r.grid_mass), ref):
                assert np.array_equal(a.cpu().numpy(), want)
    b.close()
