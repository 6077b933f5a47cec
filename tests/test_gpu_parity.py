"""Parity of the HIP density build against the oracle (GPU).

Bars (the tolerance is written in each test):
  * cell / patch indices of the binning stage: bit-exact with the reference's
    FP32 expressions (oracle.pairs, itself bit-exact with the reference);
  * deterministic mode: max relative error <= 1e-5 against the FP64 oracle on
    cells above the field's mean; base-grid mass conserved to 1e-12; the
    fields bit-identical between runs and between tile shapes;
  * atomic mode: within the FP32 accumulation error of the reference itself
    (twice the reference restatement's own error against FP64, at least 1e-5).
"""
import numpy as np
import pytest
import torch

from asohf_amd import synth
from oracle import oracle as O

from conftest import FIXTURES, load_fixture

pytestmark = pytest.mark.gpu

MODES = {"table": "A", "tsc": "B"}


@pytest.fixture(scope="module")
def bld():
    from asohf_amd.density import DensityBuilder
    b = DensityBuilder(0)
    yield b
    b.close()


def dev(snap):
    d = torch.device("cuda:0")
    return tuple(torch.from_numpy(np.ascontiguousarray(a)).to(d) for a in (snap.x, snap.y, snap.z, snap.m))


def rel_err_above_mean(got, ref):
    got = np.asarray(got, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    if ref.size == 0:
        return 0.0, 0.0
    sel = ref > ref.mean()
    if not sel.any():
        return 0.0, 0.0
    r = np.abs(got[sel] - ref[sel]) / ref[sel]
    return float(r.max()), float(np.sqrt((r ** 2).mean()))


def run(bld, snap, mode, red="deterministic", path="batched", **kw):
    x, y, z, m = dev(snap)
    r = bld.density(snap.geom, x, y, z, m, snap.npart_esp, mode=mode, reduction=red, path=path, **kw)
    torch.cuda.synchronize()
    return r.u1.cpu().numpy(), r.u11.cpu().numpy(), r.grid_mass.cpu().numpy()


def atomic_tol(snap, mode):
    u1, u11 = O.density32(snap, MODES[mode])
    d1, d11, _ = O.density64(snap, MODES[mode])
    e = max(rel_err_above_mean(u1, d1)[0], rel_err_above_mean(u11, d11)[0])
    return max(1e-5, 2.0 * e)


@pytest.mark.parametrize("name", FIXTURES)
@pytest.mark.parametrize("mode", ["table", "tsc"])
@pytest.mark.parametrize("red", ["deterministic", "atomic"])
def test_fields_vs_fp64_oracle(bld, name, mode, red):
    snap, _ = load_fixture(name)
    print(f"case {name} {mode} {red}", flush=True)
    u1, u11, gm = run(bld, snap, mode, red)
    d1, d11, dgm = O.density64(snap, MODES[mode])
    tol = 1e-5 if red == "deterministic" else atomic_tol(snap, mode)
    e1, _ = rel_err_above_mean(u1, d1)
    e11, _ = rel_err_above_mean(u11, d11[: u11.size] if d11.size else d11)
    assert e1 <= tol and e11 <= tol, (e1, e11, tol)
    total = float(snap.m[: snap.npart_esp.sum()].astype(np.float64).sum()) if mode == "table" \
        else float(snap.m.astype(np.float64).sum())
    if red == "deterministic":
        assert abs(gm[0] - total) <= 1e-12 * total
    else:
        assert abs(gm[0] - total) <= 1e-5 * total
    # per-patch deposited mass against the FP64 oracle
    if len(gm) > 1:
        np.testing.assert_allclose(gm[1:], dgm[1:], rtol=1e-5, atol=1e-6 * total)


@pytest.mark.parametrize("mode", ["table", "tsc"])
def test_membership_indices_bitexact(bld, golden, mode):
    snap, _ = golden
    x, y, z, _ = dev(snap)
    got = bld.membership(snap.geom, x, y, z, snap.npart_esp, mode=mode).cpu().numpy()
    n_use = int(snap.npart_esp.sum()) if mode == "table" else snap.n
    want = O.pairs(snap, MODES[mode], all_members=True)
    want = want[want[:, 1] < n_use]
    key = lambda a: np.lexsort((a[:, 4], a[:, 3], a[:, 2], a[:, 1], a[:, 0]))
    got, want = got[key(got)], want[key(want)]
    assert got.shape == want.shape
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mode", ["table", "tsc"])
@pytest.mark.parametrize("name", ["seeded_13_cic", "edges"])
def test_deterministic_bit_identical_across_runs_and_tiles(bld, mode, name):
    """Deterministic mode: bit-identical across runs.  TSC (and the narrow
    table entries): also across tile shapes and item sizes, the int64 sums
    being exact.  The kernel-table mode's wide classes of a tile whose wide
    entries fit one work item are deposited in the histogram form (FP64,
    rounded once per cell), the others term by term as the reference forms
    them (FP32 products); which form a tile takes depends on the tile shape
    and item size, so across those the fields agree to the FP32 rounding of
    the terms."""
    snap, _ = load_fixture(name)
    ref = run(bld, snap, mode)
    ref2 = run(bld, snap, mode)
    assert all(np.array_equal(a, b) for a, b in zip(ref, ref2))
    old = bld.get_tune()
    try:
        for tile in ((8, 8, 8), (16, 16, 16)):
            for chunk in (64, 4096):
                bld.set_tune(tile=tile, chunk=chunk, waves_per_cu=8)
                got = run(bld, snap, mode)
                again = run(bld, snap, mode)
                assert all(np.array_equal(a, b) for a, b in zip(got, again)), (tile, chunk)
                if mode == "tsc":
                    assert all(np.array_equal(a, b) for a, b in zip(ref, got)), (tile, chunk)
                else:
                    assert np.array_equal(ref[0], got[0])          # the base grid is narrow only
                    assert rel_err_above_mean(got[1], ref[1])[0] <= 1e-6, (tile, chunk)
                    assert np.allclose(got[2], ref[2], rtol=1e-6, atol=0)
    finally:
        bld.set_tune(**old)


def test_clustered_stencils_merge_exactly(bld):
    """Clumps of particles inside single cells on both sides of tile borders
    (cells 7|8, 15|16, 23|24, 31|32 and the periodic wrap), many thousands per
    cell: most lanes of a wave share one stencil, among them stencils that start
    at a tile halo's first cell (centre in the cell just below the tile), so the
    deposit sums lanes in registers in all its merge steps.  Masses span 1 to
    1e-7, so some centre remainders are negative.  Every particle deposits
    exactly its quantised mass Q(m) = rn(m * scale) (scale = 2^(30-e), e the
    binary exponent of the largest mass), so the base-grid mass equals
    sum(Q) / scale exactly; the fields are bit-identical under a shuffled
    particle order and across tile shapes, and they match the FP64 oracle."""
    src = synth.make_snapshot(synth.WORKLOADS[1], n_side=16, nx=32, patches_per_level=(3, 2), seed=5)
    g = src.geom
    rng = np.random.default_rng(11)
    dx = float(g.dx)
    x0 = -0.5 * g.nx * dx
    cells = np.array([7, 8, 15, 16, 23, 24, 31, 32])   # 1-based base cells
    pos, mass = [], []
    for _ in range(40):
        c = rng.choice(cells, 3)
        n = int(rng.integers(500, 4000))
        pos.append(x0 + (c - 0.5) * dx + rng.uniform(-0.45, 0.45, (n, 3)) * dx)
        mass.append(np.full(n, rng.choice([1.0, 1e-3, 1e-7]), np.float32))
    pos = np.concatenate(pos).astype(np.float32)
    m = np.concatenate(mass)
    n = len(m)
    esp = np.zeros(9, np.int32)
    esp[0] = n
    snap = synth.Snapshot(pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(), m, esp, g)
    print(f"case clustered stencils: {n} particles", flush=True)
    ref = run(bld, snap, "tsc")
    scale = 2.0 ** (30 - int(np.frexp(np.float32(m.max()))[1]))
    q = np.rint(m.astype(np.float64) * scale)   # m * scale is exact; rint rounds half to even
    assert ref[2][0] == q.sum() / scale
    tot = float(m.astype(np.float64).sum())
    assert abs(ref[2][0] - tot) <= 0.5 * n / scale
    perm = rng.permutation(n)
    shuf = synth.Snapshot(snap.x[perm], snap.y[perm], snap.z[perm], m[perm], esp, g)
    got = run(bld, shuf, "tsc")
    assert all(np.array_equal(a, b) for a, b in zip(ref, got))
    old = bld.get_tune()
    try:
        for tile in ((8, 8, 8), (16, 16, 16)):
            bld.set_tune(tile=tile, block=512, chunk=8192, waves_per_cu=16)
            got = run(bld, snap, "tsc")
            assert all(np.array_equal(a, b) for a, b in zip(ref, got)), tile
    finally:
        bld.set_tune(**old)
    d1, d11, _ = O.density64(snap, "B")
    assert rel_err_above_mean(ref[0], d1)[0] <= 1e-5
    assert rel_err_above_mean(ref[1], d11)[0] <= 1e-5


@pytest.mark.parametrize("mode", ["table", "tsc"])
def test_plain_form_vs_oracle(bld, golden, mode):
    snap, _ = golden
    u1, u11, gm = run(bld, snap, mode, "atomic", "plain")
    d1, d11, _ = O.density64(snap, MODES[mode])
    tol = atomic_tol(snap, mode)
    assert rel_err_above_mean(u1, d1)[0] <= tol
    assert rel_err_above_mean(u11, d11)[0] <= tol


def test_padded_layout_matches_packed(bld):
    snap, _ = load_fixture("seeded_11")
    g = snap.geom
    u1, u11, _ = run(bld, snap, "table")
    namr = (int(g.pnx.max()) + 2, int(g.pny.max()) + 1, int(g.pnz.max()) + 3)
    ld = (g.nx + 3, g.ny + 1)
    x, y, z, m = dev(snap)
    s1, n11 = bld.out_shapes(g, ld, namr)
    pad1 = torch.full(s1, -1.0, device="cuda:0")
    pad11 = torch.full((n11,), -1.0, device="cuda:0")
    bld.density(g, x, y, z, m, snap.npart_esp, mode="table", u1=pad1, u11=pad11, u1_ld=ld, u11_ld=namr)
    p1, p11 = pad1.cpu().numpy(), pad11.cpu().numpy()
    assert np.array_equal(p1[:, :g.ny, :g.nx], u1)
    assert np.all(p1[:, g.ny:, :] == -1) and np.all(p1[:, :, g.nx:] == -1)
    off = g.patch_offsets()
    blk = p11.reshape(g.npatch_total, namr[2], namr[1], namr[0])
    for p in range(g.npatch_total):
        a, b, c = g.pnx[p], g.pny[p], g.pnz[p]
        want = u11[off[p]:off[p + 1]].reshape(c, b, a)
        assert np.array_equal(blk[p, :c, :b, :a], want)
        untouched = blk[p].copy()
        untouched[:c, :b, :a] = -1
        assert np.all(untouched == -1)


def test_host_entry_point_matches_device(bld):
    snap, _ = load_fixture("edges")
    u1, u11, gm = run(bld, snap, "table")
    h1, h11, hgm = bld.density_host(snap.geom, snap.x, snap.y, snap.z, snap.m, snap.npart_esp, mode="table")
    assert np.array_equal(h1, u1) and np.array_equal(h11, u11) and np.array_equal(hgm, gm)


def test_edge_cases_empty_and_no_levels(bld):
    from asohf_amd._lib import AsohfError
    snap, _ = load_fixture("seeded_12")
    empty = synth.Snapshot(snap.x[:0], snap.y[:0], snap.z[:0], snap.m[:0], np.zeros(9, np.int32), snap.geom)
    for mode in ("table", "tsc"):
        u1, u11, gm = run(bld, empty, mode)
        assert not u1.any() and not u11.any() and not gm.any()
    # NaN position and out-of-wrap position are rejected, not written
    bad = synth.Snapshot(snap.x.copy(), snap.y, snap.z, snap.m, snap.npart_esp, snap.geom)
    bad.x[5] = np.nan
    with pytest.raises(AsohfError):
        run(bld, bad, "tsc")
    far = synth.Snapshot(snap.x.copy(), snap.y, snap.z, snap.m, snap.npart_esp, snap.geom)
    far.x[3] = np.float32(10 * snap.geom.lado0)
    with pytest.raises(AsohfError):
        run(bld, far, "table")
    # TSC: the reference clamps the base index (grids.f:1184-1192) and its
    # weights at the clamped centre (grids.f:1194-1213) leave [0, 1] more than
    # half a cell past the outer centres; such positions are rejected
    with pytest.raises(AsohfError):
        run(bld, far, "tsc")
    with pytest.raises(AsohfError):
        bld.adjoint(snap.geom, *dev(far), torch.zeros((snap.geom.nz, snap.geom.ny, snap.geom.nx),
                                                         device="cuda:0"), None)
    # the box faces themselves (and a few ulps beyond) are accepted
    half = np.float32(snap.geom.lado0 / 2)
    edge = synth.Snapshot(snap.x.copy(), snap.y.copy(), snap.z, snap.m, snap.npart_esp, snap.geom)
    edge.x[3] = -half
    edge.y[4] = np.nextafter(half, np.float32(0))
    edge.x[6] = np.nextafter(np.nextafter(-half, np.float32(-1)), np.float32(-1))
    u1, _, gm = run(bld, edge, "tsc")
    tot = float(edge.m.astype(np.float64).sum())
    assert abs(gm[0] - tot) <= 1e-12 * tot


def test_larger_snapshot_properties(bld):
    snap = synth.make_snapshot(synth.WORKLOADS[1], n_side=64, nx=64, patches_per_level=(30, 25, 20), seed=4)
    for mode in ("table", "tsc"):
        print(f"case larger {mode}", flush=True)
        u1, u11, gm = run(bld, snap, mode)
        d1, d11, dgm = O.density64(snap, MODES[mode], nthreads=0)
        assert rel_err_above_mean(u1, d1)[0] <= 1e-5
        assert rel_err_above_mean(u11, d11)[0] <= 1e-5
        tot = float(snap.m.astype(np.float64).sum())
        assert abs(gm[0] - tot) <= 1e-12 * tot
        a1, a11, _ = run(bld, snap, mode, "atomic")
        assert rel_err_above_mean(a1, d1)[0] <= atomic_tol(snap, mode)


def test_full_size_workload_properties(bld):
    """The 16.8M-particle benchmark workload at full size: size-independent
    properties (exact mass conservation, bit-identical results under a
    reversed particle order, across runs and in the graph form) and FP64
    oracle parity on a seeded 2^17-particle sample with the full hierarchy."""
    snap = synth.make_snapshot(synth.WORKLOADS[2])
    g = snap.geom
    print(f"case full-size c2: {snap.n} particles, {g.npatch_total} patches", flush=True)
    d = torch.device("cuda:0")
    x, y, z, m = dev(snap)
    ref = run(bld, snap, "tsc")
    tot = float(snap.m.astype(np.float64).sum())
    assert abs(ref[2][0] - tot) <= 1e-12 * tot
    # every particle deposits exactly its quantised mass: per-grid FP64 sums add
    # up to the total counted over the (particle, grid) memberships
    assert np.all(ref[2] >= 0)
    rev = [t.flip(0).contiguous() for t in (x, y, z, m)]
    r = bld.density(g, *rev, mode="tsc")
    torch.cuda.synchronize()
    for a, b in zip((r.u1, r.u11, r.grid_mass), ref):
        assert np.array_equal(a.cpu().numpy(), b)
    s1, n11 = bld.out_shapes(g)
    u1 = torch.empty(s1, dtype=torch.float32, device=d)
    u11 = torch.empty(n11, dtype=torch.float32, device=d)
    gm = torch.empty(1 + g.npatch_total, dtype=torch.float64, device=d)
    for rep in range(2):
        u1.zero_()
        torch.cuda.synchronize()
        bld.density(g, x, y, z, m, mode="tsc", u1=u1, u11=u11, grid_mass=gm, use_graph=True)
        torch.cuda.synchronize()
        for a, b in zip((u1, u11, gm), ref):
            assert np.array_equal(a.cpu().numpy(), b), rep
    # atomic mode against the deterministic fields (FP32 accumulation order)
    a1, a11, _ = run(bld, snap, "tsc", "atomic")
    assert rel_err_above_mean(a1, ref[0])[0] <= 1e-4
    assert rel_err_above_mean(a11, ref[1])[0] <= 1e-4
    # oracle parity on a sample, full hierarchy
    rng = np.random.default_rng(2)
    idx = np.sort(rng.choice(snap.n, size=1 << 17, replace=False))
    sub = synth.Snapshot(snap.x[idx], snap.y[idx], snap.z[idx], snap.m[idx], snap.npart_esp, g)
    u1s, u11s, gms = run(bld, sub, "tsc")
    d1, d11, dgm = O.density64(sub, "B", nthreads=0)
    assert rel_err_above_mean(u1s, d1)[0] <= 1e-5
    assert rel_err_above_mean(u11s, d11)[0] <= 1e-5
    # per-grid FP64 mass: the rounded fixed-point terms of cells clipped at a
    # patch border differ from the exact ones by half a unit (2^-31 of the mass)
    assert np.allclose(gms, dgm, rtol=1e-6, atol=0)


@pytest.mark.parametrize("mode", ["tsc", "table"])
def test_overlapping_patches_overflow_emission_lists(bld, mode):
    """Twelve copies of every level-1 patch (same-level overlap is allowed):
    particles inside emit 13+ entries, more than a workgroup's emission list
    holds, so the scatter pass enumerates those workgroups again; the fields
    still match the FP64 oracle and every copy gets the same density (both
    modes keep emission lists)."""
    snap = synth.make_snapshot(synth.WORKLOADS[1], n_side=32, nx=32, patches_per_level=(4, 3, 2), seed=21)
    g = snap.geom
    k = int(g.npatch[1])
    rep = 12
    idx = np.concatenate([np.tile(np.arange(k), rep), np.arange(k, g.npatch_total)])
    npatch = g.npatch.copy()
    npatch[1] = k * rep
    from asohf_amd.geometry import Geometry
    g2 = Geometry(g.nx, g.ny, g.nz, g.lado0, g.rote, g.rete, npatch, g.pnx[idx], g.pny[idx], g.pnz[idx],
                  g.prx[idx], g.pry[idx], g.prz[idx], fdm=g.fdm, interp_degree=g.interp_degree)
    s2 = synth.Snapshot(snap.x, snap.y, snap.z, snap.m, snap.npart_esp, g2)
    for attempt in range(2):   # the second call reuses the buffers sized by the first
        print(f"case overlap {mode} call {attempt}", flush=True)
        u1, u11, gm = run(bld, s2, mode)
        d1, d11, dgm = O.density64(s2, MODES[mode], nthreads=0)
        assert rel_err_above_mean(u1, d1)[0] <= 1e-5
        assert rel_err_above_mean(u11, d11)[0] <= 1e-5
        tot = float(snap.m[: int(snap.npart_esp.sum())].astype(np.float64).sum())
        assert abs(gm[0] - tot) <= 1e-12 * tot
        off = g2.patch_offsets()
        for p in range(k):
            a = u11[off[p]:off[p + 1]]
            for r in range(1, rep):
                q = p + r * k
                assert np.array_equal(a, u11[off[q]:off[q + 1]]), (p, r)


def test_hierarchy_binning_matches_patch_hash_binning(bld, monkeypatch):
    """The TSC build bins along the patch hierarchy (each patch's members found
    from the home entries of a grid whose box holds the patch's box); the
    patch-hash binning enumerates every particle over every level instead.
    Both give the same (tile, particle) entries, so the exact fields, the
    member counts and the entry counts agree bit for bit: on every fixture, on
    a snapshot with overlapping copies of every level-1 patch (several
    emitters and many targets per tile), on a second call with other
    particles (capacities reused or re-planned) and on the 16.8M-particle
    workload."""
    from asohf_amd.density import DensityBuilder
    monkeypatch.setenv("ASOHF_TSC_BINNING", "hash")
    hb = DensityBuilder(0)
    monkeypatch.delenv("ASOHF_TSC_BINNING")
    cases = [(nm, load_fixture(nm)[0]) for nm in FIXTURES]
    snap = synth.make_snapshot(synth.WORKLOADS[1], n_side=32, nx=32, patches_per_level=(4, 3, 2), seed=21)
    g = snap.geom
    k = int(g.npatch[1])
    idx = np.concatenate([np.tile(np.arange(k), 5), np.arange(k, g.npatch_total)])
    npatch = g.npatch.copy()
    npatch[1] = k * 5
    from asohf_amd.geometry import Geometry
    g2 = Geometry(g.nx, g.ny, g.nz, g.lado0, g.rote, g.rete, npatch, g.pnx[idx], g.pny[idx], g.pnz[idx],
                  g.prx[idx], g.pry[idx], g.prz[idx], fdm=g.fdm, interp_degree=g.interp_degree)
    cases.append(("overlap", synth.Snapshot(snap.x, snap.y, snap.z, snap.m, snap.npart_esp, g2)))
    other = synth.make_snapshot(synth.WORKLOADS[1], n_side=32, nx=32, patches_per_level=(4, 3, 2), seed=22)
    cases.append(("other particles", synth.Snapshot(other.x, other.y, other.z, other.m, other.npart_esp, g2)))
    cases.append(("c2", synth.make_snapshot(synth.WORKLOADS[2])))
    try:
        for name, sn in cases:
            print(f"case hierarchy vs hash binning: {name}", flush=True)
            for red in ("deterministic", "atomic"):
                a = run(bld, sn, "tsc", red)
                sa = bld.stats()
                b = run(hb, sn, "tsc", red)
                sb = hb.stats()
                assert sa["n_pairs"] == sb["n_pairs"] and sa["n_entries"] == sb["n_entries"], (name, sa, sb)
                if red == "deterministic":
                    assert all(np.array_equal(u, v) for u, v in zip(a, b)), name
                else:
                    assert rel_err_above_mean(a[0], b[0])[0] <= 1e-4
                    assert rel_err_above_mean(a[1], b[1])[0] <= 1e-4
    finally:
        hb.close()
