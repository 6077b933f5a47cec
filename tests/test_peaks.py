"""Peak candidates of the grid halo finder (HALOFIND_GRID, haloes_grids.f:
316-404 on the base grid and 700-790 on AMR levels) and their order
(SORT_CELLS / SORT_CELLS_AMR, haloes_grids.f:1341-1425, ARGSORT num.f:150-200).

CPU: the oracle's sort against the reference's own SORT_CELLS_AMR (compiled
from its sources into oracle/_ref) on inputs full of ties; its tie rule
against a brute-force statement (key ascending, scan position descending);
its scan against an independent numpy formulation and hand-built fields with
known answers.  GPU: asohf_hip_peak_candidates against the oracle, exactly
(densities, indices, patch numbers, order), on both U11 layouts, with CONTA1
masks, on fields with ties, and on the 16.8M-particle workload's own field.
"""
import numpy as np
import pytest

from asohf_amd import synth
from oracle import oracle as O

CONTRASTEC = 200.0


def brute_peaks(fields, contrastec, masks=None):
    """Independent statement: interior cells (one-cell border off), value
    >= contrastec/6 (FP32), strictly above the 6 face neighbours, mask 1;
    order by (1/value ascending, scan position descending)."""
    peak = np.float32(contrastec) / np.float32(6.0)
    recs = []
    pos = 0
    for g, f in enumerate(fields):
        f = np.asarray(f, np.float32)
        c = f[1:-1, 1:-1, 1:-1]
        ok = c >= peak
        for ax in range(3):
            for sh in (-1, 1):
                ok &= c > np.roll(f, -sh, axis=ax)[1:-1, 1:-1, 1:-1]
        if masks is not None:
            ok &= np.asarray(masks[g])[1:-1, 1:-1, 1:-1] == 1
        kz, jy, ix = np.nonzero(ok)
        kz, jy, ix = kz + 1, jy + 1, ix + 1
        raster = pos + ix + f.shape[2] * (jy + f.shape[1] * kz)
        for a, b, cc, r in zip(ix, jy, kz, raster):
            recs.append((f[cc, b, a], a + 1, b + 1, cc + 1, g, r))
        pos += f.size
    if not recs:
        return tuple(np.zeros(0, t) for t in (np.float32, np.int32, np.int32, np.int32, np.int32))
    v = np.array([r[0] for r in recs], np.float32)
    key = np.float32(1.0) / v
    raster = np.array([r[5] for r in recs], np.int64)
    o = np.lexsort((-raster, key))
    out = [np.float32(1.0) / key[o]]
    for i in (1, 2, 3, 4):
        out.append(np.array([recs[j][i] for j in o], np.int32))
    return tuple(out)


def quantized_field(rng, shape, levels=12, scale=60.0):
    """Values on a coarse lattice: many plateaus (non-maxima) and many
    equal-valued maxima (ties in the sort)."""
    return (np.floor(rng.uniform(0, levels, shape)) * np.float32(scale / levels)).astype(np.float32)


def test_oracle_argsort_ties_descending_position():
    rng = np.random.default_rng(0)
    for n in (0, 1, 2, 3, 5, 8, 9, 100, 1023, 1024, 1025, 5000):
        a = rng.integers(0, max(1, n // 7), n).astype(np.float32)
        ind = O.argsort32(a)
        assert np.array_equal(ind, np.lexsort((-np.arange(n), a)) + 1), n


@pytest.mark.skipif(not O.ref_available(), reason="the reference build (oracle/_ref) is not present")
def test_oracle_sort_matches_reference_sort_cells():
    rng = np.random.default_rng(1)
    for n in (1, 2, 17, 640, 4099):
        d = rng.uniform(34.0, 900.0, n).astype(np.float32)
        d[rng.integers(0, n, n // 2)] = np.float32(40.0)       # ties
        d[rng.integers(0, n, n // 5)] = np.float32(123.456)
        x, y, z, p = (rng.integers(1, 64, n).astype(np.int32) for _ in range(4))
        ref = O.run_ref_sort_cells(d, x, y, z, p)
        # the oracle's ordering on the same list, as peaks32 applies it
        key = np.float32(1.0) / d
        ind = O.argsort32(key) - 1
        assert np.array_equal(ref[0], np.float32(1.0) / key[ind]), n
        for got, a in zip(ref[1:], (x, y, z, p)):
            assert np.array_equal(got, a[ind]), n
    # the reference returns 1/(1/d), which differs from d for some values:
    # returning d itself would fail the comparison above
    d = np.random.default_rng(5).uniform(34.0, 900.0, 10000).astype(np.float32)
    assert np.any(np.float32(1.0) / (np.float32(1.0) / d) != d)


def test_oracle_peaks_known_fields():
    # one smooth peak at (17, 9, 5) (1-based), value above CONTRASTEC/6
    z, y, x = np.meshgrid(np.arange(12), np.arange(20), np.arange(30), indexing="ij")
    f = (80.0 * np.exp(-((x - 16) ** 2 + (y - 8) ** 2 + (z - 4) ** 2) / 20.0)).astype(np.float32)
    d, ix, iy, iz, g = O.peaks32([f], CONTRASTEC)
    assert list(zip(ix, iy, iz)) == [(17, 9, 5)] and d[0] == np.float32(1.0) / (np.float32(1.0) / f[4, 8, 16])
    # below the threshold: none
    assert len(O.peaks32([f * np.float32(0.3)], CONTRASTEC)[0]) == 0
    # a peak on the border is not a candidate (BOR = 1)
    f2 = np.zeros((6, 6, 6), np.float32)
    f2[0, 3, 3] = 100.0
    f2[3, 3, 3] = 100.0
    d, ix, iy, iz, g = O.peaks32([f2], CONTRASTEC)
    assert list(zip(ix, iy, iz)) == [(4, 4, 4)]
    # a plateau of two equal cells: neither is a strict maximum
    f3 = np.zeros((6, 6, 6), np.float32)
    f3[3, 3, 2] = f3[3, 3, 3] = 100.0
    assert len(O.peaks32([f3], CONTRASTEC)[0]) == 0
    # two equal peaks in two grids: the later one in scan order comes first
    d, ix, iy, iz, g = O.peaks32([f2, f2], CONTRASTEC)
    assert list(g) == [1, 0]
    # the mask removes a candidate
    m = np.ones_like(f2, np.int32)
    m[3, 3, 3] = 0
    assert len(O.peaks32([f2], CONTRASTEC, masks=[m])[0]) == 0


def test_oracle_peaks_vs_brute_force():
    rng = np.random.default_rng(2)
    fields = [quantized_field(rng, (rng.integers(3, 12), rng.integers(3, 12), rng.integers(3, 12)))
              for _ in range(20)]
    masks = [(rng.uniform(size=f.shape) > 0.2).astype(np.int32) for f in fields]
    for mk in (None, masks):
        got = O.peaks32(fields, CONTRASTEC, masks=mk)
        want = brute_peaks(fields, CONTRASTEC, masks=mk)
        assert len(got[0]) > 20
        for a, b in zip(got, want):
            assert np.array_equal(a, b)


# ---------------------------------------------------------------- GPU
def _dev(a, dtype=None):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a if dtype is None else a.astype(dtype))).cuda()


def _check(got, want, label):
    got = [t.cpu().numpy() for t in got]
    assert len(got[0]) == len(want[0]), (label, len(got[0]), len(want[0]))
    for k, (a, b) in enumerate(zip(got, want)):
        assert np.array_equal(a, b), (label, k)


def _levels_from_packed(geom, u11):
    """Per level: the list of patch fields [z, y, x] and the global patch
    number of the level's first patch."""
    sizes = geom.patch_cells()
    off = np.concatenate([[0], np.cumsum(sizes)])
    out, p = [], 0
    for lev in range(1, geom.nl + 1):
        fields = []
        for q in range(p, p + int(geom.npatch[lev])):
            fields.append(u11[off[q]:off[q + 1]].reshape(geom.pnz[q], geom.pny[q], geom.pnx[q]))
        out.append((lev, fields, p))
        p += int(geom.npatch[lev])
    return out


@pytest.mark.gpu
def test_gpu_peaks_synthetic_fields_both_layouts():
    import torch
    from asohf_amd.density import DensityBuilder
    rng = np.random.default_rng(7)
    b = DensityBuilder(0)
    snap = synth.make_snapshot(synth.WORKLOADS[1], n_side=24, nx=32, patches_per_level=(5, 4, 3), seed=3)
    g = snap.geom
    # base grid: quantized values (ties, plateaus), packed and with padded leading dims
    u1 = quantized_field(rng, (g.nz, g.ny, g.nx))
    want = O.peaks32([u1], CONTRASTEC)
    print("case peaks base packed", flush=True)
    got = b.peak_candidates(g, 0, u1=_dev(u1), contrastec=CONTRASTEC)
    _check(got, (*want[:4], np.zeros(len(want[0]), np.int32)), "base")
    pad = np.full((g.nz, g.ny + 3, g.nx + 5), 1e6, np.float32)
    pad[:, :g.ny, :g.nx] = u1
    print("case peaks base padded", flush=True)
    got = b.peak_candidates(g, 0, u1=_dev(pad), contrastec=CONTRASTEC, u1_ld=(g.nx + 5, g.ny + 3))
    _check(got, (*want[:4], np.zeros(len(want[0]), np.int32)), "base padded")
    # AMR levels: packed u11 with masks, then the same in a padded U11
    u11 = quantized_field(rng, (int(g.patch_cells().sum()),))
    mask = (rng.uniform(size=u11.shape) > 0.15).astype(np.int32)
    na = (int(g.pnx.max()) + 2, int(g.pny.max()) + 1, int(g.pnz.max()) + 3)
    padded = np.full((g.npatch_total, na[2], na[1], na[0]), 1e6, np.float32)
    padmask = np.zeros(padded.shape, np.int32)
    sizes = g.patch_cells()
    off = np.concatenate([[0], np.cumsum(sizes)])
    for q in range(g.npatch_total):
        sh = (g.pnz[q], g.pny[q], g.pnx[q])
        padded[q, :sh[0], :sh[1], :sh[2]] = u11[off[q]:off[q + 1]].reshape(sh)
        padmask[q, :sh[0], :sh[1], :sh[2]] = mask[off[q]:off[q + 1]].reshape(sh)
    mlev = {lev: f for lev, f, _ in _levels_from_packed(g, mask.astype(np.float32))}
    for lev, fields, p0 in _levels_from_packed(g, u11):
        masks = [m.astype(np.int32) for m in mlev[lev]]
        for use_mask in (False, True):
            want = O.peaks32(fields, CONTRASTEC, masks=masks if use_mask else None)
            want = (*want[:4], (want[4] + p0 + 1).astype(np.int32))
            print(f"case peaks level {lev} mask={use_mask} packed", flush=True)
            got = b.peak_candidates(g, lev, u11=_dev(u11), conta=_dev(mask) if use_mask else None,
                                    contrastec=CONTRASTEC)
            _check(got, want, f"level {lev} packed")
            print(f"case peaks level {lev} mask={use_mask} padded", flush=True)
            got = b.peak_candidates(g, lev, u11=_dev(padded.ravel()),
                                    conta=_dev(padmask.ravel()) if use_mask else None, contrastec=CONTRASTEC,
                                    u11_ld=na)
            _check(got, want, f"level {lev} padded")
    # a small output capacity grows and gives the same list
    print("case peaks small cap", flush=True)
    got = b.peak_candidates(g, 0, u1=_dev(u1), contrastec=CONTRASTEC, cap=3)
    want = O.peaks32([u1], CONTRASTEC)
    _check(got, (*want[:4], np.zeros(len(want[0]), np.int32)), "cap")
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["golden", "c2_full"])
def test_gpu_peaks_on_density_fields(which):
    """The candidates of the build's own fields, every level."""
    import torch
    from asohf_amd.density import DensityBuilder
    b = DensityBuilder(0)
    if which == "golden":
        snap = synth.make_snapshot(synth.WORKLOADS[1], n_side=48, nx=64, patches_per_level=(6, 4, 2), seed=11)
    else:
        snap = synth.make_snapshot(synth.WORKLOADS[2])
    g = snap.geom
    x, y, z, m = (_dev(a) for a in (snap.x, snap.y, snap.z, snap.m))
    r = b.density(g, x, y, z, m, snap.npart_esp, mode="tsc")
    torch.cuda.synchronize()
    u1 = r.u1.cpu().numpy().reshape(g.nz, g.ny, g.nx)
    u11 = r.u11.cpu().numpy()
    print(f"case peaks {which} base", flush=True)
    want = O.peaks32([u1], CONTRASTEC)
    assert len(want[0]) > 10
    got = b.peak_candidates(g, 0, u1=r.u1, contrastec=CONTRASTEC)
    _check(got, (*want[:4], np.zeros(len(want[0]), np.int32)), f"{which} base")
    for lev, fields, p0 in _levels_from_packed(g, u11):
        print(f"case peaks {which} level {lev}", flush=True)
        want = O.peaks32(fields, CONTRASTEC)
        got = b.peak_candidates(g, lev, u11=r.u11, contrastec=CONTRASTEC)
        _check(got, (*want[:4], (want[4] + p0 + 1).astype(np.int32)), f"{which} level {lev}")
    b.close()
