"""Golden vectors for the mesh builder: the hierarchies the reference's own
CREATE_MESH (grids.f:2-707, compiled from its sources into oracle/_ref and
run by oracle/ref_driver.f mode 6) builds for seeded synthetic snapshots.

The particles are not stored: asohf_amd.synth regenerates them from the
recorded arguments, and a SHA-256 of the arrays guards against drift.
Run from the repository root:  python tests/golden/make_mesh_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from asohf_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

CASES = [
    dict(name="default", n_side=64, nx=128, seed=9, nl=4, prm={}),
    dict(name="thr2_frac01_5lev", n_side=64, nx=128, seed=21, nl=5, prm=dict(refine_thr=2, frac_refinable=0.1)),
    dict(name="minps8_bor2_boramr1", n_side=48, nx=96, seed=4, nl=4, prm=dict(min_patchsize=8, bor=2, boramr=1)),
    dict(name="ovlp1_minps10", n_side=64, nx=128, seed=33, nl=4, prm=dict(bor_ovlp=1, min_patchsize=10)),
    dict(name="mpapolev_caps", n_side=64, nx=128, seed=9, nl=4, prm={}, mpapolev=[2, 3, 2, 5, 5, 5, 5, 5, 5]),
    dict(name="no_patches", n_side=48, nx=64, seed=5, nl=4, prm={}),
]


def snapshot(c):
    return synth.make_snapshot(synth.WORKLOADS[1], n_side=c["n_side"], nx=c["nx"], patches_per_level=(3, 2, 1),
                               seed=c["seed"])


def digest(snap):
    h = hashlib.sha256()
    for a in (snap.x, snap.y, snap.z, snap.npart_esp):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    out = {}
    for c in CASES:
        snap = snapshot(c)
        h = O.run_ref_create_mesh(snap.x, snap.y, snap.z, snap.m, snap.npart_esp, c["nx"], float(snap.geom.lado0),
                                  c["nl"], mpapolev=c.get("mpapolev"), **c["prm"])
        out[c["name"] + "__meta"] = np.frombuffer(json.dumps(dict(c, digest=digest(snap))).encode(), np.uint8)
        for k, v in h.items():
            out[c["name"] + "__" + k] = np.asarray(v)
        print(c["name"], h["nl_mesh"], h["npatch"][:6])
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "create_mesh.npz"), **out)


if __name__ == "__main__":
    main()
