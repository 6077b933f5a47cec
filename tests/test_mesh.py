"""The mesh builder CREATE_MESH (grids.f:2-707).

Parity is pinned directly on the reference: tests/golden/create_mesh.npz
holds the hierarchies its own CREATE_MESH builds (oracle/_ref, driver mode 6)
for seeded snapshots that asohf_amd.synth regenerates bit for bit.  CPU: the
generator still produces those particles; the fixtures are what the reference
produces now (where its build is present); and they obey CREATE_MESH's
placement rules (closed forms: even sides within MIN_PATCHSIZE..NAMRX, mother
cells inside the parent and the base border, PATCHRX from RADX or from the
parent's cell centres).  GPU: asohf_hip_create_mesh equals the reference's
hierarchy exactly (sizes, PATCHX..Z, PARE, PATCHRX..Z bits, NPATCH, NL) on
every case, including per-level caps, BOR_OVLP and an input that refines
nothing.
"""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_mesh_golden as G  # noqa: E402
from oracle import oracle as O  # noqa: E402

f32 = np.float32
KEYS = ("pnx", "pny", "pnz", "px", "py", "pz", "pare", "prx", "pry", "prz")


def fixtures():
    z = np.load(os.path.join(HERE, "golden", "create_mesh.npz"))
    for c in G.CASES:
        meta = json.loads(bytes(z[c["name"] + "__meta"]).decode())
        want = {k: z[c["name"] + "__" + k] for k in ("nl_mesh", "npatch") + KEYS}
        yield c, meta, want


def test_generator_reproduces_fixture_particles():
    for c, meta, _ in fixtures():
        assert G.digest(G.snapshot(c)) == meta["digest"], c["name"]


@pytest.mark.skipif(not O.ref_available(), reason="the reference build (oracle/_ref) is not present")
def test_fixtures_are_the_reference_output():
    for c, _, want in fixtures():
        snap = G.snapshot(c)
        h = O.run_ref_create_mesh(snap.x, snap.y, snap.z, snap.m, snap.npart_esp, c["nx"], float(snap.geom.lado0),
                                  c["nl"], mpapolev=c.get("mpapolev"), **c["prm"])
        for k, v in h.items():
            assert np.array_equal(np.asarray(v), want[k]), (c["name"], k)


def test_fixtures_obey_placement_rules():
    for c, _, w in fixtures():
        prm = dict(O.MESH_DEFAULTS, **c["prm"])
        snap = G.snapshot(c)
        g = snap.geom
        npatch = w["npatch"]
        nl = int(w["nl_mesh"])
        P = int(npatch.sum())
        assert P == len(w["pnx"]) and all(npatch[1:nl + 1] > 0) and npatch[nl + 1:].sum() == 0
        starts = np.concatenate([[0], np.cumsum(npatch[1:])])
        for p in range(P):
            lev = int(np.searchsorted(starts, p, side="right"))
            n = (w["pnx"][p], w["pny"][p], w["pnz"][p])
            pc = (w["px"][p], w["py"][p], w["pz"][p])
            pr = (w["prx"][p], w["pry"][p], w["prz"][p])
            assert all(v % 2 == 0 and prm["min_patchsize"] <= v <= 32 for v in n), (c["name"], p)
            if lev == 1:
                assert w["pare"][p] == 0
                bor = prm["bor"]
                for a in range(3):
                    assert bor + 1 <= pc[a] and pc[a] + n[a] // 2 - 1 <= g.nx - bor
                    assert pr[a] == g.radx[pc[a]]
            else:
                q = int(w["pare"][p]) - 1
                assert starts[lev - 2] <= q < starts[lev - 1], (c["name"], p)
                pn = (w["pnx"][q], w["pny"][q], w["pnz"][q])
                dxp = f32(g.dx / f32(2.0) ** (lev - 1))
                ppr = (w["prx"][q], w["pry"][q], w["prz"][q])
                for a in range(3):
                    assert 1 + prm["boramr"] <= pc[a] and pc[a] + n[a] // 2 - 1 <= pn[a] - prm["boramr"]
                    assert pr[a] == f32(ppr[a] + f32(f32(pc[a]) - f32(1.5)) * dxp)
        if c.get("mpapolev"):
            assert all(npatch[l] <= c["mpapolev"][l - 1] for l in range(1, 10))


@pytest.mark.gpu
def test_gpu_create_mesh_matches_reference():
    import torch
    from asohf_amd.density import DensityBuilder
    b = DensityBuilder(0)
    for c, _, w in fixtures():
        snap = G.snapshot(c)
        prm = dict(O.MESH_DEFAULTS, **c["prm"])
        x, y, z = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (snap.x, snap.y, snap.z))
        print(f"case create_mesh {c['name']}", flush=True)
        geom = b.create_mesh(snap.geom.truncated(0), x, y, z, snap.npart_esp, nl_mesh=c["nl"],
                             mpapolev=c.get("mpapolev"), **prm)
        assert geom.nl == int(w["nl_mesh"]), c["name"]
        want_np = w["npatch"][:geom.nl + 1]
        assert np.array_equal(geom.npatch, want_np), (c["name"], geom.npatch, want_np)
        got = dict(pnx=geom.pnx, pny=geom.pny, pnz=geom.pnz, px=geom.pcx, py=geom.pcy, pz=geom.pcz,
                   pare=geom.pare, prx=geom.prx, pry=geom.pry, prz=geom.prz)
        for k in KEYS:
            assert np.array_equal(np.asarray(got[k]), w[k]), (c["name"], k)
    b.close()
