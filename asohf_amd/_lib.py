"""ctypes binding of libasohf_hip.so (the C ABI declared in include/asohf_hip.h).

The library is built in-tree (asohf_amd/libasohf_hip.so).  There is no CPU
fallback: if the library is missing, or no HIP device is present, every entry
point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libasohf_hip.so")

ASOHF_OK = 0
ASOHF_ERR_CAPACITY = -2
STATUS = {0: "ASOHF_OK", -1: "ASOHF_ERR_INVALID", -2: "ASOHF_ERR_CAPACITY", -3: "ASOHF_ERR_HIP",
          -4: "ASOHF_ERR_OOM", -5: "ASOHF_ERR_NODEVICE"}
MODE_TABLE, MODE_TSC = 1, 2
REDUCE_DETERMINISTIC, REDUCE_ATOMIC = 0, 1
PATH_BATCHED, PATH_PLAIN = 0, 1

# every symbol include/asohf_hip.h declares
EXPORTS = ("asohf_hip_create", "asohf_hip_destroy", "asohf_hip_last_error", "asohf_hip_density",
           "asohf_hip_density_host", "asohf_hip_density_adjoint", "asohf_hip_membership",
           "asohf_hip_get_tune", "asohf_hip_set_tune", "asohf_hip_autotune", "asohf_hip_get_stats",
           "asohf_hip_release_workspaces", "asohf_hip_species", "asohf_hip_species_noalloc", "asohf_hip_peak_candidates",
           "asohf_hip_refinement_masks", "asohf_hip_overlap_mask", "asohf_hip_write_density_dump",
           "asohf_hip_create_mesh", "asohf_hip_version")

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


class Geom(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nl", C.c_int32),
                ("npatch", _ip), ("patchnx", _ip), ("patchny", _ip), ("patchnz", _ip),
                ("patchrx", _fp), ("patchry", _fp), ("patchrz", _fp),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("lado0", C.c_float), ("rote", C.c_float), ("rete", C.c_float), ("fdm", C.c_float),
                ("interp_degree", C.c_int32),
                ("radx", _fp), ("rady", _fp), ("radz", _fp)]


class Opts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("reduction", C.c_int32), ("path", C.c_int32),
                ("use_graph", C.c_int32), ("u1_ld", C.c_int32 * 2), ("u11_ld", C.c_int32 * 3),
                ("stream", C.c_void_p)]


class Tune(C.Structure):
    _fields_ = [("tile", C.c_int32 * 3), ("block", C.c_int32), ("chunk", C.c_int32),
                ("waves_per_cu", C.c_int32)]

    def as_dict(self):
        return {"tile": list(self.tile), "block": self.block, "chunk": self.chunk,
                "waves_per_cu": self.waves_per_cu}


class Stats(C.Structure):
    _fields_ = [("n_particles", C.c_int64), ("n_pairs", C.c_int64), ("n_entries", C.c_int64),
                ("n_tiles", C.c_int64), ("n_items", C.c_int64), ("n_cells", C.c_int64),
                ("ms_total", C.c_double), ("ms_bin", C.c_double), ("ms_scatter", C.c_double),
                ("ms_deposit", C.c_double), ("ms_sums", C.c_double), ("ms_split", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MeshParams(C.Structure):
    _fields_ = [("nl_mesh", C.c_int32), ("refine_thr", C.c_int32), ("min_patchsize", C.c_int32),
                ("frac_refinable", C.c_float), ("bor", C.c_int32), ("boramr", C.c_int32),
                ("bor_ovlp", C.c_int32), ("namr", C.c_int32 * 3), ("npalev", C.c_int32),
                ("mpapolev", C.c_int32 * 9)]


class AsohfError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


_lib = None


def load():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.asohf_hip_create.argtypes = [C.POINTER(vp), C.c_int32]
        L.asohf_hip_destroy.argtypes = [vp]
        L.asohf_hip_destroy.restype = None
        L.asohf_hip_last_error.argtypes = [vp]
        L.asohf_hip_last_error.restype = C.c_char_p
        L.asohf_hip_version.restype = C.c_char_p
        G, O = C.POINTER(Geom), C.POINTER(Opts)
        L.asohf_hip_density.argtypes = [vp, G, C.c_int64, vp, vp, vp, vp, _ip, vp, vp, vp, O]
        L.asohf_hip_density_host.argtypes = [vp, G, C.c_int64, vp, vp, vp, vp, _ip, vp, vp, vp, O]
        L.asohf_hip_density_adjoint.argtypes = [vp, G, C.c_int64, vp, vp, vp, vp, vp, vp,
                                                vp, vp, vp, vp, O]
        L.asohf_hip_membership.argtypes = [vp, G, C.c_int64, vp, vp, vp, _ip, C.c_int32, vp,
                                           C.c_int64, C.POINTER(C.c_int64)]
        L.asohf_hip_get_tune.argtypes = [vp, C.POINTER(Tune)]
        L.asohf_hip_set_tune.argtypes = [vp, C.POINTER(Tune)]
        L.asohf_hip_autotune.argtypes = [vp, G, C.c_int64, vp, vp, vp, vp, _ip, vp, vp, O,
                                         C.POINTER(Tune)]
        L.asohf_hip_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.asohf_hip_release_workspaces.argtypes = [vp]
        L.asohf_hip_species.argtypes = [vp, _ip, C.POINTER(C.c_float), C.c_float, C.c_float, C.c_int64,
                                        vp, vp, vp, vp, vp, vp, _ip, vp]
        L.asohf_hip_species_noalloc.argtypes = L.asohf_hip_species.argtypes
        L.asohf_hip_peak_candidates.argtypes = [vp, G, C.c_int32, vp, vp, vp, C.c_float, vp, vp, vp, vp, vp,
                                                C.c_int64, C.POINTER(C.c_int64), O]
        L.asohf_hip_refinement_masks.argtypes = [vp, G, _ip, _ip, _ip, _ip, vp, vp, O]
        L.asohf_hip_overlap_mask.argtypes = [vp, G, vp, O]
        L.asohf_hip_write_density_dump.argtypes = [vp, C.c_char_p, G, vp, vp, vp, vp, vp, O]
        L.asohf_hip_create_mesh.argtypes = [vp, G, C.c_int64, vp, vp, vp, _ip, C.POINTER(MeshParams), _ip, _ip,
                                            _ip, _ip, _ip, _ip, _ip, _ip, _ip, _fp, _fp, _fp, C.c_int64, O]
        for nm in EXPORTS:
            if nm not in ("asohf_hip_destroy", "asohf_hip_last_error", "asohf_hip_version"):
                getattr(L, nm).restype = C.c_int32
        _lib = L
    return _lib


def check(ctx, status):
    if status != ASOHF_OK:
        msg = load().asohf_hip_last_error(ctx) if ctx else b""
        raise AsohfError(status, (msg or b"").decode())
