"""Grid masks and the density dump (asohf.f:515-545): COMPUTE_CR0AMR
(grids.f:2076-2226), VEINSGRID (grids.f:2229-2506) and the FW2 records.

CPU: the oracle's restatements against the reference's own routines and
WRITE statements (compiled from its sources into oracle/_ref, driver mode 5):
masks bit for bit and the dump file byte for byte, on hierarchies with heavy
same-level overlaps; closed-form cases (one patch's mother cells, two
coincident patches, disjoint patches).  GPU: the entry points against the
oracle, exactly, in the packed and padded U11 layouts (padding included), and
the dump writer's file against the oracle's bytes.
"""
import os
import sys
import tempfile

import numpy as np
import pytest

from asohf_amd import synth
from asohf_amd.geometry import Geometry, background
from oracle import oracle as O

sys.path.insert(0, os.path.dirname(__file__))
from hierarchies import random_hierarchy  # noqa: E402

f32 = np.float32


def cases():
    for seed in range(4):
        yield f"overlaps_{seed}", random_hierarchy(np.random.default_rng(seed))
    yield "synth_c1", synth.make_snapshot(synth.WORKLOADS[1], n_side=32, nx=64, patches_per_level=(8, 6, 4), seed=8,
                                          side_hi=32)


@pytest.mark.skipif(not O.ref_available(), reason="the reference build (oracle/_ref) is not present")
def test_oracle_masks_and_dump_match_reference():
    for name, snap in cases():
        g = snap.geom
        raw = O.run_ref_dump(snap)
        u1, c0, u11, c11, so = O.parse_dump(raw, g)
        mc0, mc11 = O.cr0amr(g)
        ms = O.veinsgrid(g)
        assert np.array_equal(c0, mc0), name
        assert np.array_equal(c11, mc11), name
        assert np.array_equal(so, ms), name
        assert O.dump_bytes(g, u1, mc0, u11, mc11, ms) == raw, name
        if name.startswith("overlaps"):
            assert (so == 0).sum() > 1000 and (c11 == 0).sum() > 1000, name


def one_level(patches, nx=32, lado0=20.0):
    """Level-1 patches given as (PATCHX, PATCHY, PATCHZ, NX, NY, NZ)."""
    rote, rete = background(0.7, 0.3, 0.0)
    g0 = Geometry(nx, nx, nx, f32(lado0), rote, rete, np.zeros(1, np.int32), *([np.zeros(0)] * 6))
    pc = np.array([p[:3] for p in patches], np.int32)
    pn = np.array([p[3:] for p in patches], np.int32)
    pr = g0.radx[pc]
    return Geometry(nx, nx, nx, f32(lado0), rote, rete, np.array([0, len(patches)], np.int32), pn[:, 0], pn[:, 1],
                    pn[:, 2], pr[:, 0], pr[:, 1], pr[:, 2], pare=np.zeros(len(patches), np.int32),
                    pcx=pc[:, 0], pcy=pc[:, 1], pcz=pc[:, 2])


def test_oracle_masks_closed_forms():
    # one patch: its mother cells (PATCHX .. PATCHX + N/2 - 1) are the only zeros
    g = one_level([(5, 7, 9, 8, 6, 4)])
    c0, c11 = O.cr0amr(g)
    want = np.ones((32, 32, 32), np.int32)
    want[9 - 1:9 - 1 + 2, 7 - 1:7 - 1 + 3, 5 - 1:5 - 1 + 4] = 0
    assert np.array_equal(c0, want) and np.all(c11 == 1)
    assert np.all(O.veinsgrid(g) == 1)
    # two coincident patches: the first keeps its interior and drops its
    # border, the second keeps exactly the first one's border
    g = one_level([(5, 5, 5, 8, 8, 8), (5, 5, 5, 8, 8, 8)])
    so = O.veinsgrid(g).reshape(2, 8, 8, 8)
    border = np.zeros((8, 8, 8), bool)
    border[[0, -1], :, :] = border[:, [0, -1], :] = border[:, :, [0, -1]] = True
    assert np.array_equal(so[0], (~border).astype(np.int32))
    assert np.array_equal(so[1], border.astype(np.int32))
    # disjoint patches: nothing overlaps
    g = one_level([(3, 3, 3, 8, 8, 8), (12, 3, 3, 8, 8, 8)])
    assert np.all(O.veinsgrid(g) == 1)


# ---------------------------------------------------------------- GPU
def _padded(g, packed, na, fill):
    out = np.full((g.npatch_total, na[2], na[1], na[0]), fill, packed.dtype)
    off = np.concatenate([[0], np.cumsum(g.patch_cells())])
    for p in range(g.npatch_total):
        sh = (g.pnz[p], g.pny[p], g.pnx[p])
        out[p, :sh[0], :sh[1], :sh[2]] = packed[off[p]:off[p + 1]].reshape(sh)
    return out


def _interiors(g, padded):
    return np.concatenate([padded[p, :g.pnz[p], :g.pny[p], :g.pnx[p]].ravel() for p in range(g.npatch_total)])


@pytest.mark.gpu
def test_gpu_masks_match_oracle():
    from asohf_amd.density import DensityBuilder
    b = DensityBuilder(0)
    all_cases = list(cases()) + [("c2_full", synth.make_snapshot(synth.WORKLOADS[2]))]
    for name, snap in all_cases:
        g = snap.geom
        mc0, mc11 = O.cr0amr(g)
        ms = O.veinsgrid(g)
        print(f"case masks {name} packed", flush=True)
        c0, c11 = b.refinement_masks(g)
        so = b.overlap_mask(g)
        assert np.array_equal(c0.cpu().numpy(), mc0), name
        assert np.array_equal(c11.cpu().numpy(), mc11), name
        assert np.array_equal(so.cpu().numpy(), ms), name
        if g.npatch_total and max(g.pnx.max(), g.pny.max(), g.pnz.max()) <= 32:
            na = (int(g.pnx.max()) + 1, int(g.pny.max()) + 2, int(g.pnz.max()) + 3)
            print(f"case masks {name} padded", flush=True)
            c0p, c11p = b.refinement_masks(g, u1_ld=(g.nx + 3, g.ny + 1), u11_ld=na)
            sop = b.overlap_mask(g, u11_ld=na)
            c0p = c0p.cpu().numpy()[:, :g.ny, :g.nx]
            c11p = c11p.cpu().numpy().reshape(g.npatch_total, na[2], na[1], na[0])
            sop = sop.cpu().numpy().reshape(g.npatch_total, na[2], na[1], na[0])
            assert np.array_equal(c0p, mc0), name
            # the reference sets whole padded slabs to 1 before marking
            assert np.array_equal(c11p, _padded(g, mc11, na, 1)), name
            assert np.array_equal(sop, _padded(g, ms, na, 1)), name
    b.close()


@pytest.mark.gpu
def test_gpu_density_dump_bytes():
    import torch
    from asohf_amd.density import DensityBuilder
    b = DensityBuilder(0)
    for name, snap in list(cases())[:3]:
        g = snap.geom
        x, y, z, m = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (snap.x, snap.y, snap.z, snap.m))
        r = b.density(g, x, y, z, m, snap.npart_esp, mode="tsc")
        c0, c11 = b.refinement_masks(g)
        so = b.overlap_mask(g)
        torch.cuda.synchronize()
        want = O.dump_bytes(g, r.u1.cpu().numpy(), c0.cpu().numpy(), r.u11.cpu().numpy(), c11.cpu().numpy(),
                            so.cpu().numpy())
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "density_asohf00001.res")
            print(f"case dump {name} packed", flush=True)
            b.write_density_dump(path, g, r.u1, c0, r.u11, c11, so)
            with open(path, "rb") as f:
                assert f.read() == want, name
            # the same from padded arrays
            na = (int(g.pnx.max()) + 2, int(g.pny.max()) + 1, int(g.pnz.max()) + 1)
            lx, ly = g.nx + 2, g.ny + 3
            u1p = torch.full((g.nz, ly, lx), -1.0, device="cuda")
            u1p[:, :g.ny, :g.nx] = r.u1.reshape(g.nz, g.ny, g.nx)
            c0p = torch.full((g.nz, ly, lx), 7, dtype=torch.int32, device="cuda")
            c0p[:, :g.ny, :g.nx] = c0.reshape(g.nz, g.ny, g.nx)
            pads = [torch.from_numpy(_padded(g, a.cpu().numpy(), na, v)).cuda()
                    for a, v in ((r.u11, -1.0), (c11, 7), (so, 7))]
            print(f"case dump {name} padded", flush=True)
            b.write_density_dump(path, g, u1p, c0p, *pads, u1_ld=(lx, ly), u11_ld=na)
            with open(path, "rb") as f:
                assert f.read() == want, name
    b.close()
