"""Pins the CPU oracle (CPU only).

  * bit-exact against the reference's own Fortran output (golden fixtures, and
    fresh runs of oracle/_ref where it is built);
  * closed forms of the kernel table (grids.f:997-1067);
  * an independent cell-centric brute force in FP64 for both modes (no index
    shortcut: every cell of every grid evaluates the weight of every member);
  * invariants: single-particle stencil pattern, periodic translation, mass
    conservation on the periodic base grid, weights summing to one;
  * the FP64 adjoint against central finite differences of the FP64 forward.
A planted mistake in any of these formulas (a dropped term, a wrong sign, a
transposed operand or an off-by-one index) fails at least one test here.
"""
import numpy as np
import pytest

from asohf_amd import synth
from asohf_amd.geometry import Geometry, background, f32
from oracle import oracle as O

from conftest import load_fixture


# ---------------------------------------------------------------------------
# bit-exact against the reference
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", ["A", "B"])
def test_fp32_bitexact_vs_reference_fixtures(golden, mode):
    snap, ref = golden
    assert np.array_equal(snap.geom.radx, ref["radx"])
    assert snap.geom.dx == ref["dx"][0]
    u1, u11 = O.density32(snap, mode)
    assert np.array_equal(u1, ref[f"u1_{mode}"])
    assert np.array_equal(u11, ref[f"u11_{mode}"])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("seed", [21, 22])
@pytest.mark.parametrize("mode", ["A", "B"])
def test_fp32_bitexact_vs_reference_fresh(seed, mode):
    snap = synth.make_snapshot(synth.WORKLOADS[1], n_side=24, nx=24, patches_per_level=(4, 3, 3),
                               side_hi=32, seed=seed)
    _, _, _, radx, r1, r11 = O.run_ref(snap, mode)
    u1, u11 = O.density32(snap, mode)
    assert np.array_equal(u1, r1) and np.array_equal(u11, r11)


def test_malla_matches_python_geometry():
    for n, L in ((16, 16.0), (128, 40.0), (256, 40.0), (1024, 500.0)):
        lado = f32(f32(L) - f32(L) / f32(n))
        rad, d = O.malla(n, lado)
        g = Geometry(n, n, n, L, 1.0, 1.0, np.zeros(1), [], [], [], [], [], [])
        assert np.array_equal(rad, g.radx) and d == g.dx


# ---------------------------------------------------------------------------
# kernel table closed forms
# ---------------------------------------------------------------------------
def test_kernel_table_closed_forms():
    K = O.kernels32(4, 2)
    kmax = 16
    np.testing.assert_array_equal(K[0, kmax - 1:kmax + 2], np.float32([0.125, 0.75, 0.125]))
    # min/max per level printed by the reference's COMPUTE_KERNELS (grids.f:1062) in a
    # multi-threaded run, whose OpenMP normalisation sum (grids.f:1042-1051) is
    # ordered differently.  The sequential FP32 sum over (2N+1)**3 terms carries
    # a relative error up to ~1e-5 at N=16, so agreement is to that level.
    probe = {1: (0.0272727, 0.4545455), 2: (0.00462963, 0.25), 3: (6.871273e-4, 0.13238662),
             4: (9.391421e-5, 0.06818186)}
    for ir, (lo, hi) in probe.items():
        n = 1 << ir
        row = K[ir, kmax - n:kmax + n + 1]
        assert abs(row.min() - lo) <= 2e-5 * lo + 1e-9
        assert abs(row.max() - hi) <= 2e-5 * hi
        assert np.array_equal(row, row[::-1])
        assert abs(row.astype(np.float64).sum() - 1.0) < 2e-5   # FP32 sequential normalisation sum
    K64 = O.kernels64(4, 2)
    for ir in range(5):
        n = 1 << ir
        row = K64[ir, kmax - n:kmax + n + 1]
        assert abs(row.sum() - 1.0) < 1e-14
        np.testing.assert_allclose(row, K[ir, kmax - n:kmax + n + 1], rtol=2e-5)
    # CIC-like table: 1 - |i|/(N+1/2), normalised
    C = O.kernels64(3, 1)
    for ir in range(4):
        n = 1 << ir
        i = np.arange(-n, n + 1)
        raw = 1.0 - np.abs(i) / (n + 0.5)
        np.testing.assert_allclose(C[ir, 8 - n:8 + n + 1], raw / raw.sum(), rtol=1e-14)


# ---------------------------------------------------------------------------
# helpers for hand-made geometries
# ---------------------------------------------------------------------------
def _geom(nx=8, L=8.0, npatch=(0,), pn=(), pr=()):
    rote, rete = background(0.7, 0.3, 0.0)
    pn = np.array(pn, np.int32).reshape(-1, 3)
    pr = np.array(pr, np.float32).reshape(-1, 3)
    return Geometry(nx, nx, nx, L, rote, rete, np.array(npatch, np.int32),
                    pn[:, 0], pn[:, 1], pn[:, 2], pr[:, 0], pr[:, 1], pr[:, 2])


def _snap(geom, pos, m, esp=None):
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = len(pos)
    if esp is None:
        esp = np.zeros(9, np.int32)
        esp[0] = n
    return synth.Snapshot(pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(),
                          np.asarray(m, np.float32).reshape(-1), np.asarray(esp, np.int32), geom)


@pytest.mark.parametrize("mode", ["A", "B"])
def test_single_particle_at_cell_centre(mode):
    g = _geom()
    c = g.radx[4]
    snap = _snap(g, [(c, c, c)], [2.0])
    u1, _, gm = O.density64(snap, mode)
    w = np.array([0.125, 0.75, 0.125])
    expect = np.zeros((8, 8, 8))
    expect[2:5, 2:5, 2:5] = 2.0 * w[:, None, None] * w[None, :, None] * w[None, None, :]
    den = float(g.dx) ** 3 * float(g.rote) * float(g.rete) ** 3
    np.testing.assert_allclose(u1 * den, expect, rtol=1e-12, atol=1e-15)
    assert abs(gm[0] - 2.0) < 1e-12


@pytest.mark.parametrize("mode", ["A", "B"])
def test_periodic_translation_by_one_cell(mode):
    rng = np.random.default_rng(3)
    g = _geom(nx=8, L=8.0)
    pos = rng.uniform(-3.9, 3.9, size=(50, 3)).astype(np.float32)
    m = rng.uniform(1, 2, 50).astype(np.float32)
    u1, _, _ = O.density64(_snap(g, pos, m), mode)
    shifted = pos.copy()
    shifted[:, 0] = np.where(pos[:, 0] + g.dx >= 4.0, pos[:, 0] + g.dx - 8.0, pos[:, 0] + g.dx)
    v1, _, _ = O.density64(_snap(g, shifted, m), mode)
    # mode A is index-only (exact); mode B sees the FP32 rounding of x+DX
    rtol = 1e-12 if mode == "A" else 1e-5
    np.testing.assert_allclose(v1, np.roll(u1, 1, axis=2), rtol=rtol, atol=rtol * u1.max())


@pytest.mark.parametrize("mode", ["A", "B"])
def test_base_grid_conserves_mass(golden, mode):
    snap, _ = golden
    _, _, gm = O.density64(snap, mode)
    total = snap.m.astype(np.float64).sum()
    assert abs(gm[0] - total) <= 1e-12 * total


def test_tsc_weights_sum_to_one():
    rng = np.random.default_rng(5)
    g = _geom(nx=8, L=8.0)
    for _ in range(20):
        p = rng.uniform(-3.5, 3.5, 3).astype(np.float32)
        u1, _, gm = O.density64(_snap(g, [p], [1.0]), "B")
        assert abs(gm[0] - 1.0) < 1e-13


# ---------------------------------------------------------------------------
# independent FP64 brute force, cell-centric
# ---------------------------------------------------------------------------
def _bspline(t):
    a = np.abs(t)
    return np.where(a <= 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def _brute_b(snap):
    g = snap.geom
    x, y, z, m = (v.astype(np.float64) for v in (snap.x, snap.y, snap.z, snap.m))
    L = float(g.lado0)
    den0 = float(g.dx) * float(g.dy) * float(g.dz) * float(g.rote) * float(g.rete) ** 3
    cen = g.radx[1:g.nx + 1].astype(np.float64)
    def per(t):   # periodic minimum image in cells
        return (t + g.nx / 2) % g.nx - g.nx / 2
    wx = _bspline(per((x[:, None] - cen[None, :]) / float(g.dx)))
    wy = _bspline(per((y[:, None] - cen[None, :]) / float(g.dy)))
    wz = _bspline(per((z[:, None] - cen[None, :]) / float(g.dz)))
    u1 = np.einsum("p,pk,pj,pi->kji", m, wz, wy, wx) / den0 / float(g.fdm)
    u11 = []
    lev = g.level_of_patch()
    for p in range(g.npatch_total):
        ir = int(lev[p])
        d = [f32(g.dx / f32(2.0) ** ir), f32(g.dy / f32(2.0) ** ir), f32(g.dz / f32(2.0) ** ir)]
        pr = [g.prx[p], g.pry[p], g.prz[p]]
        pn = [g.pnx[p], g.pny[p], g.pnz[p]]
        lo = [f32(pr[a] - f32(2.0) * d[a]) for a in range(3)]
        hi = [f32(lo[a] + f32(pn[a] + 2) * d[a]) for a in range(3)]
        pos = [snap.x, snap.y, snap.z]
        inside = np.ones(snap.n, bool)
        for a in range(3):
            inside &= (lo[a] < pos[a]) & (pos[a] < hi[a])
        ws = []
        for a in range(3):
            c = np.array([f32(pr[a] + (f32(i) - f32(1.5)) * d[a]) for i in range(1, pn[a] + 1)],
                         np.float64)
            ws.append(_bspline((pos[a][inside].astype(np.float64)[:, None] - c[None, :]) / float(d[a])))
        den = float(d[0]) * float(d[1]) * float(d[2]) * float(g.rote) * float(g.rete) ** 3
        f = np.einsum("p,pk,pj,pi->kji", m[inside], ws[2], ws[1], ws[0]) / den / float(g.fdm)
        u11.append(f.ravel())
    return u1, np.concatenate(u11) if u11 else np.zeros(0)


def _brute_a(snap):
    g = snap.geom
    K = O.kernels64(g.nl, g.interp_degree)
    kmax = 1 << g.nl
    den0 = float(g.dx) * float(g.dy) * float(g.dz) * float(g.rote) * float(g.rete) ** 3
    xl = f32(-g.lado0 / f32(2.0))
    def base_idx(v, d):
        return ((v - xl) / d).astype(np.int64) + 1     # FP32 INT((XP-XL)/DX)+1
    ix, iy, iz = base_idx(snap.x, g.dx), base_idx(snap.y, g.dy), base_idx(snap.z, g.dz)
    c = np.arange(1, g.nx + 1)
    def kw(i):
        off = (c[None, :] - i[:, None] + g.nx // 2) % g.nx - g.nx // 2
        return np.where(np.abs(off) <= 1, K[0][np.clip(off, -1, 1) + kmax], 0.0)
    u1 = np.einsum("p,pk,pj,pi->kji", snap.m.astype(np.float64), kw(iz), kw(iy), kw(ix)) / den0 / float(g.fdm)
    species = np.repeat(np.arange(9), snap.npart_esp)
    lev = g.level_of_patch()
    u11 = []
    for p in range(g.npatch_total):
        ir = int(lev[p])
        d = [f32(g.dx / f32(2.0) ** ir), f32(g.dy / f32(2.0) ** ir), f32(g.dz / f32(2.0) ** ir)]
        pr = [g.prx[p], g.pry[p], g.prz[p]]
        pn = [int(g.pnx[p]), int(g.pny[p]), int(g.pnz[p])]
        fix = [f32(pr[a] - d[a]) for a in range(3)]
        lo = [f32(fix[a] - g.dx) for a in range(3)]
        hi = [f32(f32(fix[a] + f32(pn[a]) * d[a]) + g.dx) for a in range(3)]
        pos = [snap.x, snap.y, snap.z]
        inside = np.ones(snap.n, bool)
        for a in range(3):
            inside &= (lo[a] < pos[a]) & (pos[a] < hi[a])
        irk = np.maximum(ir - species[inside], 0)
        nk = 1 << irk
        ws = []
        for a in range(3):
            idx = np.floor((pos[a][inside] - fix[a]) / d[a]).astype(np.int64) + 1
            cc = np.arange(1, pn[a] + 1)
            off = cc[None, :] - idx[:, None]
            ok = np.abs(off) <= nk[:, None]
            ws.append(np.where(ok, K[irk[:, None], np.clip(off, -kmax, kmax) + kmax], 0.0))
        den = float(d[0]) * float(d[1]) * float(d[2]) * float(g.rote) * float(g.rete) ** 3
        f = np.einsum("p,pk,pj,pi->kji", snap.m[inside].astype(np.float64), ws[2], ws[1], ws[0]) / den / float(g.fdm)
        u11.append(f.ravel())
    return u1, np.concatenate(u11) if u11 else np.zeros(0)


@pytest.mark.parametrize("name", ["edges", "seeded_11"])
@pytest.mark.parametrize("mode", ["A", "B"])
def test_fp64_oracle_vs_brute_force(name, mode):
    snap, _ = load_fixture(name)
    u1, u11, _ = O.density64(snap, mode)
    b1, b11 = (_brute_a if mode == "A" else _brute_b)(snap)
    scale = max(np.abs(b1).max(), 1e-300)
    np.testing.assert_allclose(u1, b1, rtol=1e-10, atol=1e-12 * scale)
    scale11 = max(np.abs(b11).max(), 1e-300) if b11.size else 1.0
    np.testing.assert_allclose(u11, b11, rtol=1e-10, atol=1e-12 * scale11)


def test_fp32_close_to_fp64(golden):
    snap, _ = golden
    for mode in "AB":
        u1, u11 = O.density32(snap, mode)
        d1, d11, _ = O.density64(snap, mode)
        sel = d1 > d1.mean()
        assert np.max(np.abs(u1[sel] - d1[sel]) / d1[sel]) < 1e-3


# ---------------------------------------------------------------------------
# membership records
# ---------------------------------------------------------------------------
def test_pairs_consistent_with_fields(golden):
    snap, _ = golden
    pr = O.pairs(snap, "B")
    assert (pr[:, 0] == 0).sum() == snap.n
    # every patch with members has mass and vice versa
    _, _, gm = O.density64(snap, "B")
    counts = np.bincount(pr[:, 0], minlength=1 + snap.geom.npatch_total)
    assert np.all((counts[1:] > 0) == (gm[1:] > 0))
    # mode B patch indices lie in 0..N+1 (the IX range of grids.f:1321-1323)
    lev_n = np.concatenate([[snap.geom.nx], snap.geom.pnx])
    assert np.all(pr[:, 2] >= 0) and np.all(pr[:, 2] <= lev_n[pr[:, 0]] + 2)


# ---------------------------------------------------------------------------
# adjoint vs finite differences
# ---------------------------------------------------------------------------
def test_adjoint64_vs_central_differences():
    snap, _ = load_fixture("seeded_11")
    g = snap.geom
    rng = np.random.default_rng(0)
    g1 = rng.standard_normal((g.nz, g.ny, g.nx))
    g11 = rng.standard_normal(int(g.patch_cells().sum()))
    gm, gx, gy, gz = O.adjoint64(snap, g1, g11, nthreads=1)

    def loss(s):
        u1, u11, _ = O.density64(s, "B")
        return float((g1 * u1).sum() + (g11 * u11).sum())

    base_pairs = O.pairs(snap, "B")
    checked = 0
    for i in rng.choice(snap.n, 40, replace=False):
        h = float(g.dx) / (2.0 ** g.nl) * 1e-3
        for comp, grad in (("x", gx), ("y", gy), ("z", gz)):
            s_p = synth.Snapshot(snap.x.copy(), snap.y.copy(), snap.z.copy(), snap.m, snap.npart_esp, g)
            s_m = synth.Snapshot(snap.x.copy(), snap.y.copy(), snap.z.copy(), snap.m, snap.npart_esp, g)
            v = getattr(snap, comp)[i]
            vp, vm = np.float32(v + h), np.float32(v - h)
            getattr(s_p, comp)[i] = vp
            getattr(s_m, comp)[i] = vm
            # skip particles whose patch membership changes inside the stencil of h
            same = lambda a, b: np.array_equal(a[a[:, 1] == i][:, 0], b[b[:, 1] == i][:, 0])
            if not (same(O.pairs(s_p, "B"), base_pairs) and same(O.pairs(s_m, "B"), base_pairs)):
                continue
            fd = (loss(s_p) - loss(s_m)) / (float(vp) - float(vm))
            assert abs(fd - grad[i]) <= 2e-4 * max(abs(grad[i]), 1e-3 * np.abs(grad).max()), (i, comp)
            checked += 1
        # mass derivative: L is linear in m
        s_p = synth.Snapshot(snap.x, snap.y, snap.z, snap.m.copy(), snap.npart_esp, g)
        s_p.m[i] = np.float32(snap.m[i] * 2)
        fd_m = (loss(s_p) - loss(snap)) / float(snap.m[i])
        assert abs(fd_m - gm[i]) <= 1e-8 * max(abs(gm[i]), 1e-30) + 1e-12 * np.abs(gm).max()
    assert checked >= 60
