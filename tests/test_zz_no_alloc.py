"""The species partition without allocation (asohf_hip_species_noalloc),
called through the C ABI, against the allocating entry point
asohf_hip_species on the same inputs: class counts, classes and the whole
permutation identical, on two calls in a row (the second smaller, on the
scratch array the first left behind), and the size limits refused before any
device work."""
import ctypes as C

import numpy as np
import pytest
import torch

from asohf_amd import _lib, synth
from asohf_amd.density import DensityBuilder, _ptr
from asohf_amd.geometry import background

pytestmark = pytest.mark.gpu

ASOHF_ERR_CAPACITY = -2


@pytest.fixture(scope="module")
def bld():
    b = DensityBuilder(0)
    yield b
    b.close()


def species_call(fn, bld, nmax, d, rodo, re0, x, y, z, m):
    n = int(x.numel())
    sp = torch.full((max(n, 1),), -7, dtype=torch.int32, device=x.device)
    order = torch.full((max(n, 1),), -7, dtype=torch.int64, device=x.device)
    esp = np.zeros(9, np.int32)
    nm = (C.c_int32 * 3)(*nmax)
    dd = (C.c_float * 3)(*d)
    st = fn(bld._h, nm, dd, C.c_float(rodo), C.c_float(re0), C.c_int64(n), _ptr(x), _ptr(y), _ptr(z), _ptr(m),
            _ptr(sp), _ptr(order), esp.ctypes.data_as(C.POINTER(C.c_int32)),
            C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return st, sp[:n].cpu().numpy(), order[:n].cpu().numpy(), esp


def lib_fns():
    L = _lib.load()
    vp = C.c_void_p
    fn = L.asohf_hip_species_noalloc
    fn.restype = C.c_int
    fn.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.c_float, C.c_float, C.c_int64,
                   vp, vp, vp, vp, vp, vp, C.POINTER(C.c_int32), vp]
    ref = L.asohf_hip_species
    ref.restype = C.c_int
    ref.argtypes = fn.argtypes
    return fn, ref


def snapshot(n_side, nx, seed):
    snap = synth.make_snapshot(synth.WORKLOADS[1], n_side=n_side, nx=nx, patches_per_level=(2, 2), seed=seed)
    d = torch.device("cuda:0")
    xyzm = [torch.from_numpy(np.ascontiguousarray(a)).to(d) for a in (snap.x, snap.y, snap.z, snap.m)]
    return snap, xyzm


def test_species_noalloc_matches_allocating_entry_point(bld):
    fn, ref = lib_fns()
    spec = synth.WORKLOADS[1]
    rodo, re0 = (float(v) for v in background(spec.h, spec.omega_m, 0.0))
    # two calls in a row: 2^20 particles on a 128^3 grid, then 32,768 on 64^3
    for n_side, nmax, seed in ((101, 128, 31), (32, 64, 32)):
        snap, (x, y, z, m) = snapshot(n_side, nmax, seed)
        d = (float(snap.geom.dx),) * 3
        print(f"case noalloc species: {snap.n} particles, NMAX {nmax}", flush=True)
        st_a, sp_a, ord_a, esp_a = species_call(fn, bld, (nmax,) * 3, d, rodo, re0, x, y, z, m)
        assert st_a == 0, _lib.load().asohf_hip_last_error(bld._h)
        st_b, sp_b, ord_b, esp_b = species_call(ref, bld, (nmax,) * 3, d, rodo, re0, x, y, z, m)
        assert st_b == 0
        assert np.array_equal(esp_a, esp_b) and int(esp_a.sum()) == snap.n
        assert np.array_equal(sp_a, sp_b)
        assert np.array_equal(ord_a, ord_b)
        assert np.array_equal(np.sort(ord_a), np.arange(snap.n))


def test_species_noalloc_refuses_sizes_beyond_its_scratch(bld):
    fn, _ = lib_fns()
    d = torch.device("cuda:0")
    n = (1 << 20) + 1
    x = torch.zeros(n, dtype=torch.float32, device=d)
    st, _, _, _ = species_call(fn, bld, (64, 64, 64), (1.0,) * 3, 1.0, 1.0, x, x, x, x)
    assert st == ASOHF_ERR_CAPACITY
    x = x[:1000]
    st, _, _, _ = species_call(fn, bld, (256, 128, 128), (1.0,) * 3, 1.0, 1.0, x, x, x, x)
    assert st == ASOHF_ERR_CAPACITY
