"""The hand-written TSC adjoint (GPU) and its autograd wrapper.

  * against the FP64 oracle adjoint (itself pinned by central differences of
    the FP64 forward in test_oracle.py): FP32 accumulation, tolerance 1e-4 of
    the gradient scale;
  * against FP64 central differences (h = 1e-3 of the finest cell holding the
    particle) on 1,000 random particles of a 256^3-particle snapshot on a
    256^3 grid, via the loss of each particle's own deposit (linearity);
  * against autograd through a dense PyTorch deposit (FP64, CPU) of 32^3
    particles on a 32^3 grid.
"""
import numpy as np
import pytest
import torch

from asohf_amd import synth
from asohf_amd.geometry import Geometry, background, f32
from oracle import oracle as O

from conftest import FIXTURES, load_fixture

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bld():
    from asohf_amd.density import DensityBuilder
    b = DensityBuilder(0)
    yield b
    b.close()


def dev(*arrs):
    return tuple(torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0") for a in arrs)


@pytest.mark.parametrize("name", FIXTURES)
def test_adjoint_vs_fp64_oracle(bld, name):
    snap, _ = load_fixture(name)
    g = snap.geom
    rng = np.random.default_rng(1)
    g1 = rng.standard_normal((g.nz, g.ny, g.nx)).astype(np.float32)
    g11 = rng.standard_normal(max(int(g.patch_cells().sum()), 1)).astype(np.float32)
    x, y, z, m = dev(snap.x, snap.y, snap.z, snap.m)
    t1, t11 = dev(g1, g11)
    got = [t.cpu().numpy().astype(np.float64) for t in bld.adjoint(g, x, y, z, m, t1, t11)]
    want = O.adjoint64(snap, g1.astype(np.float64), g11.astype(np.float64))
    for a, b, nm in zip(got, want, ("dm", "dx", "dy", "dz")):
        scale = np.abs(b).max()
        assert np.max(np.abs(a - b)) <= 1e-4 * scale, nm


def test_adjoint_without_patch_gradient(bld):
    """g11=None is a zero gradient on every patch cell, on a fixture whose
    particles sit in patches (the gather must not read past a short buffer)."""
    snap, _ = load_fixture("seeded_12")
    g = snap.geom
    assert g.npatch_total > 0
    rng = np.random.default_rng(2)
    g1 = rng.standard_normal((g.nz, g.ny, g.nx)).astype(np.float32)
    x, y, z, m = dev(snap.x, snap.y, snap.z, snap.m)
    t1, = dev(g1)
    got = [t.cpu().numpy().astype(np.float64) for t in bld.adjoint(g, x, y, z, m, t1, None)]
    want = O.adjoint64(snap, g1.astype(np.float64), np.zeros(int(g.patch_cells().sum())))
    for a, b, nm in zip(got, want, ("dm", "dx", "dy", "dz")):
        assert np.max(np.abs(a - b)) <= 1e-4 * np.abs(b).max(), nm
    # the C entry point itself refuses a NULL patch gradient when there are patches
    import ctypes as C
    from asohf_amd import _lib
    from asohf_amd.density import _GeomArgs, _ptr
    outs = [torch.empty(snap.n, dtype=torch.float32, device="cuda:0") for _ in range(4)]
    o = bld._opts("tsc", "deterministic", "batched", None, None)
    st = _lib.load().asohf_hip_density_adjoint(bld._h, _GeomArgs(g).ref, snap.n, _ptr(x), _ptr(y), _ptr(z),
                                               _ptr(m), _ptr(t1), None, *[_ptr(t) for t in outs], C.byref(o))
    assert st != 0


def test_adjoint_vs_central_differences_large(bld):
    spec = synth.WORKLOADS[4]
    snap = synth.make_snapshot(spec)
    g = snap.geom
    rng = np.random.default_rng(spec.seed)
    w = rng.standard_normal((g.nz, g.ny, g.nx)).astype(np.float32)
    x, y, z, m = dev(snap.x, snap.y, snap.z, snap.m)
    tw, = dev(w)
    dm, gx, gy, gz = (t.cpu().numpy() for t in bld.adjoint(g, x, y, z, m, tw, None))
    # the forward it differentiates: L = sum(w * rho)
    r = bld.density(g, x, y, z, m, None, mode="tsc", want_mass=False)
    L = float((tw.double() * r.u1.double()).sum())
    w64 = w.astype(np.float64)
    empty = np.zeros(1)
    h = float(g.dx) * 1e-3
    picks = rng.choice(snap.n, 1000, replace=False)
    lin = sum(float(dm[i]) * float(snap.m[i]) for i in picks[:50])
    assert np.isfinite(L) and np.isfinite(lin)
    for i in picks:
        for comp, grad in ((0, gx), (1, gy), (2, gz)):
            p = [snap.x[i], snap.y[i], snap.z[i]]
            pp, pm = list(p), list(p)
            pp[comp] = np.float32(p[comp] + h)
            pm[comp] = np.float32(p[comp] - h)
            fd = (O.loss_one(snap, i, w64, empty, pp) - O.loss_one(snap, i, w64, empty, pm)) / \
                 (float(pp[comp]) - float(pm[comp]))
            # TSC weights are piecewise quadratic with C1 joins: central
            # differences are exact except within h of a join
            assert abs(fd - grad[i]) <= 2e-3 * max(abs(fd), 1e-2 * np.abs(grad).max()), (i, comp, fd, grad[i])
        # mass: L is linear in m
        fdm = O.loss_one(snap, i, w64, empty) / float(snap.m[i])
        assert abs(fdm - dm[i]) <= 1e-4 * max(abs(fdm), 1e-2 * np.abs(dm).max())


def _dense_torch_deposit(x, y, z, m, geom):
    """Differentiable TSC base-grid deposit (the formulas of grids.f:1180-1228),
    FP64 on the CPU, via scatter_add."""
    nx = geom.nx
    xl = float(f32(-geom.lado0 / f32(2.0)))
    rad = torch.from_numpy(geom.radx.astype(np.float64))
    d = float(geom.dx)
    ws, idx = [], []
    for v in (x, y, z):
        i = torch.clamp(((v.detach().float() - xl) / float(geom.dx)).to(torch.int64) + 1, 1, nx)
        c = torch.stack([rad[i - 1], rad[i], rad[i + 1]], 1)
        b = (v[:, None] - c).abs() / d
        w = torch.stack([0.5 * (1.5 - b[:, 0]) ** 2, 0.75 - b[:, 1] ** 2, 0.5 * (1.5 - b[:, 2]) ** 2], 1)
        cells = torch.stack([i - 1, i, i + 1], 1)
        cells = torch.where(cells == nx + 1, torch.ones_like(cells), cells)
        cells = torch.where(cells == 0, torch.full_like(cells, nx), cells)
        ws.append(w)
        idx.append(cells - 1)
    den = float(geom.dx) ** 3 * float(geom.rote) * float(geom.rete) ** 3 * float(geom.fdm)
    out = torch.zeros(nx * nx * nx, dtype=torch.float64)
    for c in range(3):
        for b in range(3):
            for a in range(3):
                flat = idx[0][:, a] + nx * (idx[1][:, b] + nx * idx[2][:, c])
                out = out.scatter_add(0, flat, m * ws[0][:, a] * ws[1][:, b] * ws[2][:, c])
    return (out / den).reshape(nx, nx, nx)


def test_autograd_vs_dense_torch(bld):
    from asohf_amd.density import tsc_deposit
    snap = synth.make_snapshot(synth.WORKLOADS[4], n_side=32, nx=32)
    g = snap.geom
    rng = np.random.default_rng(4)
    w = rng.standard_normal((g.nz, g.ny, g.nx))
    x, y, z, m = (t.clone().requires_grad_(True) for t in dev(snap.x, snap.y, snap.z, snap.m))
    u1, _ = tsc_deposit(x, y, z, m, g, bld=bld)
    loss = (u1.double() * torch.from_numpy(w).to("cuda:0")).sum()
    loss.backward()
    xc, yc, zc, mc = (torch.from_numpy(a.astype(np.float64)).requires_grad_(True)
                      for a in (snap.x, snap.y, snap.z, snap.m))
    ref = _dense_torch_deposit(xc, yc, zc, mc, g)
    np.testing.assert_allclose(u1.detach().cpu().numpy(), ref.detach().numpy(), rtol=2e-5,
                               atol=2e-5 * ref.detach().abs().max().item())
    (ref * torch.from_numpy(w)).sum().backward()
    for got, want in ((x.grad, xc.grad), (y.grad, yc.grad), (z.grad, zc.grad), (m.grad, mc.grad)):
        got, want = got.cpu().numpy(), want.numpy()
        assert np.max(np.abs(got - want)) <= 1e-4 * np.abs(want).max()


def test_adjoint_rejects_invalid_input_in_gather_order(bld):
    """From 2^20 particles the adjoint gathers in base-cell order and its key
    pass makes the input checks: a NaN position, a negative mass and a
    position far outside the base grid are each rejected; the same particles
    without them pass with finite gradients."""
    from asohf_amd._lib import AsohfError
    snap, _ = load_fixture("seeded_12")
    g = snap.geom
    n = (1 << 20) + 7
    rng = np.random.default_rng(5)
    L = float(g.lado0)
    x0, y0, z0 = (((rng.random(n) * 0.98 - 0.49) * L).astype(np.float32) for _ in range(3))
    m0 = np.full(n, snap.m[0], np.float32)
    w = torch.zeros((g.nz, g.ny, g.nx), device="cuda:0")
    for case in ("nan", "negative mass", "far"):
        print(f"case adjoint invalid input {case} n={n}", flush=True)
        x, y, z, m = x0.copy(), y0.copy(), z0.copy(), m0.copy()
        if case == "nan":
            x[n // 3] = np.nan
        elif case == "negative mass":
            m[n // 2] = -1.0
        else:
            y[n - 1] = np.float32(10 * L)
        with pytest.raises(AsohfError):
            bld.adjoint(g, *dev(x, y, z, m), w, None)
    outs = bld.adjoint(g, *dev(x0, y0, z0, m0), w + 1.0, None)
    for t in outs:
        assert torch.isfinite(t).all()
