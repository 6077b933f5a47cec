"""C ABI and host logic (CPU only: no compute call needs a GPU here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from asohf_amd import _lib, synth
from asohf_amd.geometry import Geometry

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "asohf_hip.h")).read()
    return sorted(set(re.findall(r"\b(asohf_hip_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    decl = declared_symbols()
    assert decl == sorted(_lib.EXPORTS)
    for name in decl:
        assert hasattr(L, name), name


def test_version_string():
    assert b"gfx950" in _lib.load().asohf_hip_version()


def test_create_without_device_reports_nodevice():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    h = C.c_void_p()
    st = _lib.load().asohf_hip_create(C.byref(h), 0)
    assert st == -5 and not h.value


def test_null_context_is_rejected():
    L = _lib.load()
    assert L.asohf_hip_density(None, None, 0, None, None, None, None, None, None, None, None, None) == -1
    assert L.asohf_hip_get_stats(None, None) == -1
    assert L.asohf_hip_release_workspaces(None) == -1
    assert L.asohf_hip_species(None, None, None, 0.0, 0.0, 0, None, None, None, None, None, None, None,
                               None) == -1
    assert L.asohf_hip_peak_candidates(None, None, 0, None, None, None, 0.0, None, None, None, None, None, 0,
                                       None, None) == -1
    assert L.asohf_hip_refinement_masks(None, None, None, None, None, None, None, None, None) == -1
    assert L.asohf_hip_overlap_mask(None, None, None, None) == -1
    assert L.asohf_hip_write_density_dump(None, b"/nonexistent", None, None, None, None, None, None, None) == -1
    assert L.asohf_hip_create_mesh(None, None, 0, None, None, None, None, None, None, None, None, None, None,
                                   None, None, None, None, None, None, None, 0, None) == -1


def test_mesh_params_layout_matches_header():
    # asohf_mesh_params: 7 scalars, namr[3], npalev, mpapolev[9] (4-byte fields, no padding)
    assert C.sizeof(_lib.MeshParams) == 4 * (7 + 3 + 1 + 9)
    txt = open(os.path.join(ROOT, "include", "asohf_hip.h")).read()
    body = txt[txt.index("typedef struct asohf_mesh_params"):txt.index("} asohf_mesh_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)          # comments out
    fields = re.findall(r"\b(\w+)(?:\[\d\])?[,;]", body)
    assert fields == [f for f, _ in _lib.MeshParams._fields_], fields


def test_synthetic_hierarchy_obeys_mesh_rules():
    snap = synth.make_snapshot(synth.WORKLOADS[2], n_side=40, nx=64, patches_per_level=(20, 15, 10, 8), seed=9)
    g = snap.geom
    assert np.all(g.pnx % 2 == 0) and np.all(g.pny % 2 == 0) and np.all(g.pnz % 2 == 0)
    lev = g.level_of_patch()
    for p in range(g.npatch_total):
        ir = lev[p]
        if ir == 1:
            assert g.pare[p] == 0
            i1 = np.nonzero(g.radx == g.prx[p])[0]
            assert len(i1) == 1 and 1 <= i1[0] <= g.nx       # PATCHRX = RADX(I1)
            continue
        q = g.pare[p] - 1
        assert lev[q] == ir - 1
        dxp = np.float32(g.dx / np.float32(2.0) ** (ir - 1))
        for a, (pr, prp, n, npar) in enumerate(((g.prx, g.prx, g.pnx, g.pnx), (g.pry, g.pry, g.pny, g.pny),
                                                  (g.prz, g.prz, g.pnz, g.pnz))):
            # child starts on a parent cell I1 and lies inside the parent
            hits = [i for i in range(1, npar[q] + 1)
                    if np.float32(prp[q] + (np.float32(i) - np.float32(1.5)) * dxp) == pr[p]]
            assert hits, (p, a)
            assert hits[0] + n[p] // 2 - 1 <= npar[q]


def test_species_partition_contiguous():
    snap = synth.make_snapshot(synth.WORKLOADS[1], n_side=32, nx=32, patches_per_level=(4, 3, 2), seed=3)
    assert snap.npart_esp.sum() == snap.n
    assert snap.npart_esp[0] > 0 and snap.npart_esp[1:].sum() > 0


def test_geometry_validation():
    g = Geometry(2, 8, 8, 8.0, 1.0, 1.0, np.zeros(1), [], [], [], [], [], [])
    with pytest.raises(ValueError):
        g.check()
    g = Geometry(8, 8, 8, 8.0, 1.0, 1.0, np.array([0, 1]), [4], [4], [0], [0.0], [0.0], [0.0])
    with pytest.raises(ValueError):
        g.check()
