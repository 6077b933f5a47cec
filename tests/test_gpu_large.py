"""The benchmark workloads at full size on the GPU, against the oracle.

  * 256^3 particles, 4 AMR levels (924 patches): the cell and patch indices of
    every (particle, grid) member bit-exact against the oracle's restatement of
    the reference's expressions (grids.f:1180-1192, 1294-1323) on 2^20 sampled
    particles with the whole hierarchy; the whole field (all 16.8M particles)
    against the FP64 oracle.  The same particles on the hierarchy CREATE_MESH
    builds for them: membership, the whole TSC field and a kernel-table sample.
  * 512^3 particles, 6 levels (about 20,000 patches), both reductions, and
    1024^3 particles, 7 levels (about 150,000 patches), forward and adjoint:
    size-independent properties of the whole build (base mass exact,
    bit-identical under a reversed particle order and in the graph form) and
    FP64 oracle parity on a seeded 2^17-particle sample built with the whole
    hierarchy (fields, per-grid masses); the adjoint against the FP64 oracle
    on that sample and against central differences on 1,000 particles.
"""
import numpy as np
import pytest
import torch

from asohf_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bld():
    from asohf_amd.density import DensityBuilder
    b = DensityBuilder(0)
    yield b
    b.close()


def rel_err_above_mean(got, ref):
    got = np.asarray(got, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    sel = ref > ref.mean()
    if not sel.any():
        return 0.0
    return float(np.max(np.abs(got[sel] - ref[sel]) / ref[sel]))


def host_sample(xyzm, esp_full, geom, k, seed):
    """k seeded particles (reference order kept), species counts recomputed."""
    n = int(xyzm[0].numel())
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(n, size=k, replace=False))
    ti = torch.from_numpy(idx).to(xyzm[0].device)
    dv = [a[ti].contiguous() for a in xyzm]
    bounds = np.concatenate([[0], np.cumsum(np.asarray(esp_full, np.int64))])
    esp = np.bincount(np.searchsorted(bounds, idx, side="right") - 1, minlength=9)[:9].astype(np.int32)
    hs = synth.Snapshot(*[a.cpu().numpy() for a in dv], esp, geom)
    return hs, dv, idx


def cell_unit(geom, mmax):
    """Output value of one fixed-point unit (at most 2^-29 of the largest mass,
    the deterministic deposit's quantum) per cell: u1, then u11 cells."""
    u0 = mmax * 2.0 ** -29 / (float(geom.denbas(0)) * float(geom.fdm))
    per = np.array([mmax * 2.0 ** -29 / (float(geom.denbas(int(l))) * float(geom.fdm))
                    for l in range(geom.nl + 1)])
    return u0, np.repeat(per[geom.level_of_patch()], geom.patch_cells())


def rel_err_above(got, ref, floor=None):
    got = np.asarray(got, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    sel = ref > ref.mean()
    if floor is not None:
        sel &= ref >= floor
    if not sel.any():
        return 0.0
    return float(np.max(np.abs(got[sel] - ref[sel]) / ref[sel]))


def sample_parity(bld, snap, xyzm, reduction, seed, tol):
    """GPU build of a 2^17-particle sample with the whole hierarchy against the
    FP64 oracle: fields and per-grid masses.  The sample thins the density by
    the sampling factor, so many cells above the field's mean hold a single
    small corner term.  The deterministic deposit rounds every term to a
    fixed-point unit (at most max(m) * 2^-29), so it is held to the bar on
    cells of at least 2^21 units (2^-8 of a particle's mass; a full build's
    cells above the mean hold about a particle's mass or more); the
    FP32-atomic deposit, whose terms are not quantised, on every cell above
    the mean."""
    hs, dv, _ = host_sample(xyzm, snap.npart_esp, snap.geom, 1 << 17, seed)
    r = bld.density(snap.geom, *dv, hs.npart_esp, mode="tsc", reduction=reduction)
    u1, u11, gm = r.u1.cpu().numpy(), r.u11.cpu().numpy(), r.grid_mass.cpu().numpy()
    d1, d11, dgm = O.density64(hs, "B", nthreads=0)
    if reduction == "deterministic":
        q1, q11 = cell_unit(snap.geom, float(hs.m.max()))
        e1 = rel_err_above(u1, d1, 2.0 ** 21 * q1)
        e11 = rel_err_above(u11, d11, 2.0 ** 21 * q11)
        # every cell above the mean: within the quantum of its terms (27 per
        # member at most, half a unit each)
        for got, ref, q in ((u1.ravel(), d1.ravel(), q1), (u11, d11, q11)):
            sel = ref > ref.mean()
            assert np.all(np.abs(got[sel] - ref[sel]) <= 1e-5 * ref[sel] + 64 * np.broadcast_to(q, ref.shape)[sel])
    else:
        e1, e11 = rel_err_above(u1, d1), rel_err_above(u11, d11)
    print(f"  sample {reduction}: base {e1:.2e} patches {e11:.2e}", flush=True)
    assert e1 <= tol and e11 <= tol
    # no grid gets mass that the oracle does not give it (the converse holds
    # up to the quantum: a grid reached only by corner terms below half a
    # unit keeps none in the deterministic deposit)
    assert np.all((dgm > 0) | (gm == 0))
    # per-grid FP64 masses: a grid's deposit is the sum of its members'
    # terms, computed in FP32 as the reference does (a few ulps each) and,
    # deterministic, rounded to the fixed-point unit (at most max(m) * 2^-29,
    # half of it per term, 27 terms per member); a member whose whole stencil
    # lies in the grid deposits exactly its quantised mass, so only members
    # clipped at a patch border carry the FP32 error; atomic terms add FP32
    # accumulation
    memb = np.bincount(bld.membership(snap.geom, dv[0], dv[1], dv[2], hs.npart_esp, mode="tsc")
                       .cpu().numpy()[:, 0], minlength=snap.geom.npatch_total + 1)
    mmax = float(hs.m.max())
    if reduction == "deterministic":
        tol = 1e-6 * np.abs(dgm) + 13.5 * mmax * 2.0 ** -29 * memb
        tot = float(hs.m.astype(np.float64).sum())
        assert abs(gm[0] - tot) <= 1e-12 * tot     # the base grid holds every stencil whole
    else:
        tol = 1e-4 * np.abs(dgm) + 1e-6 * mmax * memb
    bad = np.abs(gm - dgm) > tol
    assert not bad.any(), (np.nonzero(bad)[0][:10], gm[bad][:10], dgm[bad][:10], memb[bad][:10])
    return hs


def whole_build_properties(bld, snap, xyzm, reduction="deterministic"):
    """Base mass, bit identity under a reversed order and in the graph form."""
    g = snap.geom
    d = xyzm[0].device
    s1, n11 = bld.out_shapes(g)
    u1 = torch.empty(s1, dtype=torch.float32, device=d)
    u11 = torch.empty(n11, dtype=torch.float32, device=d)
    gm = torch.empty(1 + g.npatch_total, dtype=torch.float64, device=d)
    bld.density(g, *xyzm, snap.npart_esp, mode="tsc", reduction=reduction, u1=u1, u11=u11, grid_mass=gm)
    torch.cuda.synchronize()
    tot = float(xyzm[3].double().sum())
    # deterministic: every particle deposits exactly its quantised mass on
    # the base grid; atomic: FP32 cell sums
    assert abs(float(gm[0]) - tot) <= (1e-12 if reduction == "deterministic" else 1e-6) * tot
    ref = [t.clone() for t in (u1, u11, gm)]
    if reduction != "deterministic":
        return ref
    # TSC weights do not depend on the species order, so any permutation of
    # the particles gives the same exact sums
    rev = [t.flip(0).contiguous() for t in xyzm]
    r = bld.density(g, *rev, None, mode="tsc")
    assert all(torch.equal(a, b) for a, b in zip((r.u1, r.u11, r.grid_mass), ref))
    del rev, r
    for rep in range(2):
        u1.zero_()
        bld.density(g, *xyzm, snap.npart_esp, mode="tsc", u1=u1, u11=u11, grid_mass=gm, use_graph=True)
        torch.cuda.synchronize()
        assert all(torch.equal(a, b) for a, b in zip((u1, u11, gm), ref)), rep
    return ref


def test_c2_membership_bitexact_and_whole_field(bld):
    spec = synth.WORKLOADS[2]
    snap = synth.make_snapshot(spec)
    g = snap.geom
    print(f"case c2 membership: {snap.n} particles, {g.npatch_total} patches", flush=True)
    d = torch.device("cuda:0")
    xyzm = tuple(torch.from_numpy(a).to(d) for a in (snap.x, snap.y, snap.z, snap.m))
    hs, dv, idx = host_sample(xyzm, snap.npart_esp, g, 1 << 20, 5)
    got = bld.membership(g, dv[0], dv[1], dv[2], hs.npart_esp, mode="tsc").cpu().numpy()
    want = O.pairs(hs, "B", all_members=True)
    key = lambda a: np.lexsort((a[:, 4], a[:, 3], a[:, 2], a[:, 1], a[:, 0]))
    got, want = got[key(got)], want[key(want)]
    assert got.shape == want.shape and len(got) > (1 << 20)
    assert np.array_equal(got, want)
    # every level of the hierarchy takes part
    lev = np.concatenate([[0], g.level_of_patch()])
    assert set(np.unique(lev[got[:, 0]])) == set(range(g.nl + 1))
    print("case c2 whole field vs FP64", flush=True)
    r = bld.density(g, *xyzm, snap.npart_esp, mode="tsc")
    u1, u11, gm = r.u1.cpu().numpy(), r.u11.cpu().numpy(), r.grid_mass.cpu().numpy()
    d1, d11, dgm = O.density64(snap, "B", nthreads=0)
    assert rel_err_above_mean(u1, d1) <= 1e-5
    assert rel_err_above_mean(u11, d11) <= 1e-5
    assert np.allclose(gm, dgm, rtol=1e-6, atol=0)
    assert float(g.fdm) != 1.0        # RENORM_DENSITY's division is exercised


def test_c2_on_create_mesh_hierarchy(bld):
    """The 256^3-particle workload on the hierarchy CREATE_MESH (grids.f:2-707)
    builds for its particles (asohf_hip_create_mesh, asohf.dat's mesh block,
    the overdensity rule of grids.f:107-129, 357-405) instead of the synthetic
    one: TSC membership bit-exact and the whole TSC field against the FP64
    oracle; the kernel-table DENSITY (the live path on such hierarchies) on a
    2^18-particle sample against the FP64 oracle."""
    spec = synth.WORKLOADS[2]
    snap = synth.make_snapshot(spec)
    d = torch.device("cuda:0")
    xyzm = tuple(torch.from_numpy(a).to(d) for a in (snap.x, snap.y, snap.z, snap.m))
    g = bld.create_mesh(snap.geom.truncated(0), *xyzm[:3], snap.npart_esp, nl_mesh=spec.nl, npalev=1 << 20)
    assert g.nl == spec.nl and g.npatch_total > 100 and float(g.fdm) == float(snap.geom.fdm)
    full = synth.Snapshot(snap.x, snap.y, snap.z, snap.m, snap.npart_esp, g)
    print(f"case c2 on CREATE_MESH: {snap.n} particles, npatch {g.npatch.tolist()}", flush=True)
    hs, dv, _ = host_sample(xyzm, snap.npart_esp, g, 1 << 18, 7)
    got = bld.membership(g, dv[0], dv[1], dv[2], hs.npart_esp, mode="tsc").cpu().numpy()
    want = O.pairs(hs, "B", all_members=True)
    key = lambda a: np.lexsort((a[:, 4], a[:, 3], a[:, 2], a[:, 1], a[:, 0]))
    assert np.array_equal(got[key(got)], want[key(want)])
    lev = np.concatenate([[0], g.level_of_patch()])
    assert set(np.unique(lev[got[:, 0]])) == set(range(g.nl + 1))
    print("case c2 on CREATE_MESH: whole TSC field vs FP64", flush=True)
    r = bld.density(g, *xyzm, snap.npart_esp, mode="tsc")
    u1, u11, gm = r.u1.cpu().numpy(), r.u11.cpu().numpy(), r.grid_mass.cpu().numpy()
    d1, d11, dgm = O.density64(full, "B", nthreads=0)
    assert rel_err_above_mean(u1, d1) <= 1e-5
    assert rel_err_above_mean(u11, d11) <= 1e-5
    assert np.allclose(gm, dgm, rtol=1e-6, atol=0)
    tot = float(snap.m.astype(np.float64).sum())
    assert abs(gm[0] - tot) <= 1e-12 * tot
    del r, d1, d11
    print("case c2 on CREATE_MESH: kernel-table sample vs FP64", flush=True)
    ra = bld.density(g, *dv, hs.npart_esp, mode="table")
    a1, a11, agm = O.density64(hs, "A", nthreads=0)
    q1, q11 = cell_unit(g, float(hs.m.max()))
    assert rel_err_above(ra.u1.cpu().numpy(), a1, 2.0 ** 21 * q1) <= 1e-5
    assert rel_err_above(ra.u11.cpu().numpy(), a11, 2.0 ** 21 * q11) <= 1e-5
    assert abs(float(ra.grid_mass[0]) - float(hs.m.astype(np.float64).sum())) <= 1e-12 * float(hs.m.sum())
    bld.release_workspaces()


@pytest.mark.parametrize("reduction", ["deterministic", "atomic"])
def test_c3_full_size(bld, reduction):
    spec = synth.WORKLOADS[3]
    snap = synth.make_snapshot_device(spec)
    xyzm = (snap.x, snap.y, snap.z, snap.m)
    print(f"case c3 {reduction}: {snap.n} particles, {snap.geom.npatch_total} patches", flush=True)
    ref = whole_build_properties(bld, snap, xyzm, reduction)
    if reduction == "atomic":
        # the FP32-atomic fields against the exact deterministic ones
        det = bld.density(snap.geom, *xyzm, snap.npart_esp, mode="tsc")
        assert rel_err_above_mean(ref[0].cpu().numpy(), det.u1.cpu().numpy()) <= 1e-4
        assert rel_err_above_mean(ref[1].cpu().numpy(), det.u11.cpu().numpy()) <= 1e-4
        del det
    del ref
    bld.release_workspaces()
    sample_parity(bld, snap, xyzm, reduction, 3, 1e-5)
    bld.release_workspaces()


def test_c5_full_size_forward_and_adjoint(bld):
    spec = synth.WORKLOADS[5]
    snap = synth.make_snapshot_device(spec)
    g = snap.geom
    xyzm = (snap.x, snap.y, snap.z, snap.m)
    print(f"case c5 forward: {snap.n} particles, {g.npatch_total} patches", flush=True)
    ref = whole_build_properties(bld, snap, xyzm)
    del ref
    bld.release_workspaces()
    hs = sample_parity(bld, snap, xyzm, "deterministic", 5, 1e-5)
    sample_parity(bld, snap, xyzm, "atomic", 5, 1e-5)
    bld.release_workspaces()
    # the adjoint of L = sum(w1 * u1) + sum(w11 * u11) over the whole build
    print("case c5 adjoint", flush=True)
    gen = torch.Generator(device="cuda:0").manual_seed(55)
    s1, n11 = bld.out_shapes(g)
    w1 = torch.randn(s1, generator=gen, device="cuda:0")
    w11 = torch.randn(n11, generator=gen, device="cuda:0")
    dm, dx, dy, dz = bld.adjoint(g, *xyzm, w1, w11)
    torch.cuda.synchronize()
    # central differences move the FP32 position by h = 1e-3 of the finest
    # cell holding the particle; particles where that is below a few ulps of
    # the coordinate (the deepest levels far from the box centre) cannot be
    # differenced in FP32 and are not picked
    rng = np.random.default_rng(56)
    cand = np.sort(rng.choice(snap.n, 20000, replace=False))
    tc = torch.from_numpy(cand).to("cuda:0")
    rec = bld.membership(g, *[t[tc].contiguous() for t in xyzm[:3]], None, mode="tsc").cpu().numpy()
    lev_of_grid = np.concatenate([[0], g.level_of_patch()])
    deep = np.zeros(len(cand), np.int64)
    np.maximum.at(deep, rec[:, 1], lev_of_grid[rec[:, 0]])
    hcand = float(g.dx) / 2.0 ** deep * 1e-3
    pc = np.stack([t[tc].cpu().numpy() for t in xyzm[:3]])
    ok = np.all(hcand[None, :] >= 8 * np.spacing(np.abs(pc).astype(np.float32)).astype(np.float64), axis=0)
    picks = cand[ok][:1000]
    hpick = hcand[ok][:1000]
    assert len(picks) == 1000
    ti = torch.from_numpy(picks).to("cuda:0")
    got = [t[ti].cpu().numpy().astype(np.float64) for t in (dm, dx, dy, dz)]
    pos = [t[ti].cpu().numpy() for t in xyzm]
    del dm, dx, dy, dz
    bld.release_workspaces()
    w1h = w1.cpu().numpy().astype(np.float64)
    w11h = w11.cpu().numpy().astype(np.float64)
    del w1, w11
    # the adjoint on the 2^17 sample against the FP64 oracle adjoint
    ts = [torch.from_numpy(a).to("cuda:0") for a in (hs.x, hs.y, hs.z, hs.m)]
    sg = [t.cpu().numpy().astype(np.float64)
          for t in bld.adjoint(g, *ts, torch.from_numpy(w1h.astype(np.float32)).to("cuda:0"),
                               torch.from_numpy(w11h.astype(np.float32)).to("cuda:0"))]
    want = O.adjoint64(hs, w1h, w11h)
    for a, b, nm in zip(sg, want, ("dm", "dx", "dy", "dz")):
        assert np.max(np.abs(a - b)) <= 1e-4 * np.abs(b).max(), nm
    # central differences of each picked particle's own loss (linearity),
    # positions moved in FP32
    one = synth.Snapshot(*[np.zeros(1, np.float32) for _ in range(4)], np.zeros(9, np.int32), g)
    gmax = [np.abs(v).max() for v in got]
    bad = 0
    for j in range(len(picks)):
        one.m[0] = pos[3][j]
        h = hpick[j]
        p = [pos[0][j], pos[1][j], pos[2][j]]
        for comp in range(3):
            pp, pm = list(p), list(p)
            pp[comp] = np.float32(p[comp] + h)
            pm[comp] = np.float32(p[comp] - h)
            fd = (O.loss_one(one, 0, w1h, w11h, pp) - O.loss_one(one, 0, w1h, w11h, pm)) / \
                 (float(pp[comp]) - float(pm[comp]))
            gv = got[1 + comp][j]
            # TSC weights are piecewise quadratic with C1 joins: central
            # differences are exact except within h of a join, or of a patch
            # box face (where the particle starts or stops counting there)
            if abs(fd - gv) > 2e-3 * max(abs(fd), 1e-2 * gmax[1 + comp]):
                bad += 1
        # the mass derivative: L is linear in m
        fdm = O.loss_one(one, 0, w1h, w11h, p) / float(pos[3][j])
        assert abs(fdm - got[0][j]) <= 1e-4 * max(abs(fdm), 1e-2 * gmax[0])
    assert bad <= 10, bad
