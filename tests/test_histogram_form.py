"""The kernel-table density (DENSITY, grids.f:710-994) as an NGP mass
histogram convolved with a separable kernel, against the direct per-particle
form, both in FP64 on the CPU.

The reference deposits every member of a patch with its table weights
K(i - IX) K(j - JY) K(k - KZ) over the clipped box [IX - N, IX + N]
(grids.f:888-951; table grids.f:997-1067), N = 2^max(IR - s, 0) for species
s.  The weights depend only on the integer offset, so per patch and per
species class the field is (histogram of the members' masses at their cells
IX, JY, KZ on the box [1 - N, N1 + N]) convolved with K x K x K, restricted
to the patch's cells; on the base grid the same with N = 1 and the periodic
wrap (grids.f:773-815).  The GPU's wide-stencil deposit evaluates exactly
this form per tile; here numpy builds it from the oracle's membership records
(the reference's index expressions) and compares it with the oracle's direct
FP64 deposit on every fixture, including the one with every species width."""
import numpy as np
import pytest

from oracle import oracle as O

from conftest import FIXTURES, load_fixture


def conv1(h, k, axis, n_out, pad):
    """out[i] = sum_d h[i + pad - d] k[d + pad] for i in 0..n_out-1 along axis."""
    h = np.moveaxis(h, axis, 0)
    out = np.zeros((n_out,) + h.shape[1:])
    for d in range(-pad, pad + 1):
        out += k[d + pad] * h[pad - d: pad - d + n_out]
    return np.moveaxis(out, 0, axis)


def den64(g, ir):
    """DENBAS of level ir in FP64 on the FP32 cell size, as the FP64 oracle
    forms it."""
    d = [float(np.float32(v / np.float32(2.0) ** ir)) for v in (g.dx, g.dy, g.dz)]
    return d[0] * d[1] * d[2] * float(g.rote) * float(g.rete) ** 3


def histogram_form(snap, row_shift=0):
    g = snap.geom
    nl = g.nl
    K = O.kernels64(nl, g.interp_degree)
    kmax = 1 << nl
    rec = O.pairs(snap, "A", all_members=True)          # (grid, particle, ix, jy, kz)
    bounds = np.concatenate([[0], np.cumsum(snap.npart_esp.astype(np.int64))])
    species = np.searchsorted(bounds, rec[:, 1], side="right") - 1
    m = snap.m.astype(np.float64)
    # base grid: N = 1 for every particle, single periodic wrap
    base = rec[rec[:, 0] == 0]
    h = np.zeros((g.nz, g.ny, g.nx))
    np.add.at(h, ((base[:, 4] - 1) % g.nz, (base[:, 3] - 1) % g.ny, (base[:, 2] - 1) % g.nx), m[base[:, 1]])
    k1 = K[0, kmax - 1: kmax + 2]
    u1 = np.zeros_like(h)
    for c in range(3):
        for b in range(3):
            for a in range(3):
                u1 += k1[a] * k1[b] * k1[c] * np.roll(h, (c - 1, b - 1, a - 1), axis=(0, 1, 2))
    u1 = u1 / den64(g, 0) / float(g.fdm)
    # patches: one histogram per (patch, class)
    off = g.patch_offsets()
    lev = g.level_of_patch()
    u11 = np.zeros(int(off[-1]))
    for p in range(g.npatch_total):
        ir = int(lev[p])
        if ir == 0:
            continue
        sel = rec[:, 0] == 1 + p
        r = rec[sel]
        if len(r) == 0:
            continue
        irk = np.maximum(ir - species[sel], 0)
        n1, n2, n3 = int(g.pnx[p]), int(g.pny[p]), int(g.pnz[p])
        u = np.zeros((n3, n2, n1))
        for q in np.unique(irk):
            nk = 1 << int(q)
            rq = r[irk == q]
            h = np.zeros((n3 + 2 * nk, n2 + 2 * nk, n1 + 2 * nk))
            # the cells whose stencil meets the patch: IX in [1 - N, N1 + N]
            ix, jy, kz = rq[:, 2] - 1 + nk, rq[:, 3] - 1 + nk, rq[:, 4] - 1 + nk
            ok = (ix >= 0) & (ix < n1 + 2 * nk) & (jy >= 0) & (jy < n2 + 2 * nk) & (kz >= 0) & (kz < n3 + 2 * nk)
            np.add.at(h, (kz[ok], jy[ok], ix[ok]), m[rq[ok, 1]])
            k = K[int(q) - row_shift if q >= row_shift else int(q), kmax - nk: kmax + nk + 1]
            t = conv1(h, k, 2, n1, nk)
            t = conv1(t, k, 1, n2, nk)
            u += conv1(t, k, 0, n3, nk)
        u11[off[p]:off[p + 1]] = (u / den64(g, ir) / float(g.fdm)).ravel()
    return u1, u11


@pytest.mark.parametrize("name", FIXTURES)
def test_histogram_form_equals_direct_form(name):
    snap, _ = load_fixture(name)
    h1, h11 = histogram_form(snap)
    d1, d11, _ = O.density64(snap, "A")
    scale1 = np.abs(d1).max()
    assert np.max(np.abs(h1 - d1)) <= 1e-12 * scale1
    if d11.size:
        assert np.max(np.abs(h11 - d11)) <= 1e-12 * np.abs(d11).max()


def test_histogram_form_sees_every_species_width():
    """The edges fixture puts members of every class into the patches, so the
    comparison above covers N = 1 up to the largest half-width."""
    snap, _ = load_fixture("edges")
    rec = O.pairs(snap, "A", all_members=True)
    bounds = np.concatenate([[0], np.cumsum(snap.npart_esp.astype(np.int64))])
    species = np.searchsorted(bounds, rec[:, 1], side="right") - 1
    lev = np.concatenate([[0], snap.geom.level_of_patch()])[rec[:, 0]]
    widths = set(np.maximum(lev - species, 0).tolist())
    assert widths == set(range(int(lev.max()) + 1))


def test_histogram_form_catches_a_wrong_kernel_row():
    """A planted mistake in the form (each wide class convolved with the row of
    the level below, same width) differs from the direct form: the
    comparison is not vacuous."""
    snap, _ = load_fixture("edges")
    h1, h11 = histogram_form(snap, row_shift=1)
    d1, d11, _ = O.density64(snap, "A")
    assert np.max(np.abs(h11 - d11)) > 1e-6 * np.abs(d11).max()
