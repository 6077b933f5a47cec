"""Python access to the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker or the timed CPU baseline.  The product path
(asohf_amd + libasohf_hip) never imports it.

  - the C restatement in asohf_oracle.c (libasohf_oracle.so), FP32 and FP64;
  - run_ref(): the reference's own Fortran DENSITY / INTERPOLATE_DENSITY_TSC
    compiled from the reference sources into _ref/asohf_ref_density
    (built by `make ref` where the reference tree exists); run_ref_species()
    and run_ref_sort_cells() call SORT_DM_PARTICLES_LOCALDENSITY and
    SORT_CELLS_AMR through the same driver, run_ref_dump() the grid masks
    and the density dump of asohf.f:515-545, run_ref_create_mesh() the mesh
    builder CREATE_MESH.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libasohf_oracle.so")
REF_BIN = os.path.join(HERE, "_ref", "asohf_ref_density")

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)


class OrcGeom(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nl", C.c_int32),
                ("npatch", _ip), ("pnx", _ip), ("pny", _ip), ("pnz", _ip),
                ("prx", _fp), ("pry", _fp), ("prz", _fp),
                ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float),
                ("lado0", C.c_float), ("rote", C.c_float), ("rete", C.c_float),
                ("fdm", C.c_float), ("interp_degree", C.c_int32),
                ("radx", _fp), ("rady", _fp), ("radz", _fp)]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "libasohf_oracle.so"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True,
                       stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        G = C.POINTER(OrcGeom)
        L.orc_malla.restype = C.c_float
        L.orc_malla.argtypes = [C.c_int, C.c_float, _fp]
        L.orc_kernels32.argtypes = [C.c_int, C.c_int, _fp]
        L.orc_kernels64.argtypes = [C.c_int, C.c_int, _dp]
        for nm in ("orc_density_a32",):
            getattr(L, nm).argtypes = [G, C.c_int64, _fp, _fp, _fp, _fp, _ip, _fp, _fp, C.c_int]
        L.orc_density_a64.argtypes = [G, C.c_int64, _fp, _fp, _fp, _fp, _ip, _dp, _dp, _dp, C.c_int]
        L.orc_density_b32.argtypes = [G, C.c_int64, _fp, _fp, _fp, _fp, _fp, _fp, C.c_int]
        L.orc_density_b64.argtypes = [G, C.c_int64, _fp, _fp, _fp, _fp, _dp, _dp, _dp, C.c_int]
        L.orc_adjoint_b64.argtypes = [G, C.c_int64, _fp, _fp, _fp, _fp, _dp, _dp,
                                      _dp, _dp, _dp, _dp, C.c_int]
        L.orc_pairs.restype = C.c_int64
        L.orc_pairs.argtypes = [G, C.c_int, C.c_int64, _fp, _fp, _fp, _ip, _ip, C.c_int64]
        L.orc_loss_one_b64.restype = C.c_double
        L.orc_loss_one_b64.argtypes = [G, C.c_float, C.c_float, C.c_float, C.c_float, _dp, _dp]
        L.orc_stencil_work_a.argtypes = [G, C.c_int64, _fp, _fp, _fp, _ip, C.POINTER(C.c_int64)]
        for nm in ("orc_density_a32", "orc_density_a64", "orc_density_b32", "orc_density_b64",
                   "orc_adjoint_b64"):
            getattr(L, nm).restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class _G:
    """Keeps the numpy buffers behind an OrcGeom alive."""

    def __init__(self, geom, use_radx: bool = True):
        self.keep = [geom.npatch, geom.pnx, geom.pny, geom.pnz, geom.prx, geom.pry, geom.prz,
                     geom.radx, geom.rady, geom.radz]
        self.s = OrcGeom(geom.nx, geom.ny, geom.nz, geom.nl, _p(geom.npatch, _ip),
                         _p(geom.pnx, _ip), _p(geom.pny, _ip), _p(geom.pnz, _ip),
                         _p(geom.prx, _fp), _p(geom.pry, _fp), _p(geom.prz, _fp),
                         geom.dx, geom.dy, geom.dz, geom.lado0, geom.rote, geom.rete,
                         geom.fdm, geom.interp_degree,
                         _p(geom.radx, _fp) if use_radx else None,
                         _p(geom.rady, _fp) if use_radx else None,
                         _p(geom.radz, _fp) if use_radx else None)

    @property
    def ref(self):
        return C.byref(self.s)


def _arrs(snap):
    x, y, z, m = (np.ascontiguousarray(a, dtype=np.float32) for a in (snap.x, snap.y, snap.z, snap.m))
    esp = np.ascontiguousarray(snap.npart_esp, dtype=np.int32)
    return x, y, z, m, esp


def malla(n: int, lado: float):
    rad = np.zeros(n + 2, np.float32)
    d = lib().orc_malla(n, np.float32(lado), _p(rad, _fp))
    return rad, np.float32(d)


def kernels32(nl: int, degree: int) -> np.ndarray:
    kmax = 1 << nl
    K = np.zeros((nl + 1, 2 * kmax + 1), np.float32)
    lib().orc_kernels32(nl, degree, _p(K, _fp))
    return K


def kernels64(nl: int, degree: int) -> np.ndarray:
    kmax = 1 << nl
    K = np.zeros((nl + 1, 2 * kmax + 1), np.float64)
    lib().orc_kernels64(nl, degree, _p(K, _dp))
    return K


def _out(geom, dt):
    u1 = np.zeros((geom.nz, geom.ny, geom.nx), dt)
    u11 = np.zeros(max(int(geom.patch_cells().sum()), 1), dt)
    return u1, u11


def density32(snap, mode: str = "B", nthreads: int = 1):
    """Reference-faithful FP32 fields: (u1[nz,ny,nx], u11 packed)."""
    g = _G(snap.geom)
    x, y, z, m, esp = _arrs(snap)
    u1, u11 = _out(snap.geom, np.float32)
    L = lib()
    if mode == "A":
        rc = L.orc_density_a32(g.ref, len(x), _p(x, _fp), _p(y, _fp), _p(z, _fp), _p(m, _fp),
                               _p(esp, _ip), _p(u1, _fp), _p(u11, _fp), nthreads)
    else:
        rc = L.orc_density_b32(g.ref, len(x), _p(x, _fp), _p(y, _fp), _p(z, _fp), _p(m, _fp),
                               _p(u1, _fp), _p(u11, _fp), nthreads)
    if rc != 0:
        raise ValueError(f"oracle rejected the input (rc={rc})")
    return u1, u11[: int(snap.geom.patch_cells().sum())]


def density64(snap, mode: str = "B", nthreads: int = 1):
    """FP64 fields and deposited mass per grid: (u1, u11, grid_mass[1+P])."""
    g = _G(snap.geom)
    x, y, z, m, esp = _arrs(snap)
    u1, u11 = _out(snap.geom, np.float64)
    gm = np.zeros(1 + snap.geom.npatch_total, np.float64)
    L = lib()
    if mode == "A":
        rc = L.orc_density_a64(g.ref, len(x), _p(x, _fp), _p(y, _fp), _p(z, _fp), _p(m, _fp),
                               _p(esp, _ip), _p(u1, _dp), _p(u11, _dp), _p(gm, _dp), nthreads)
    else:
        rc = L.orc_density_b64(g.ref, len(x), _p(x, _fp), _p(y, _fp), _p(z, _fp), _p(m, _fp),
                               _p(u1, _dp), _p(u11, _dp), _p(gm, _dp), nthreads)
    if rc != 0:
        raise ValueError(f"oracle rejected the input (rc={rc})")
    return u1, u11[: int(snap.geom.patch_cells().sum())], gm


def adjoint64(snap, g1: np.ndarray, g11: np.ndarray, nthreads: int = 0):
    """FP64 gradients of L = sum(g1*u1) + sum(g11*u11) for mode B:
    (dm, dx, dy, dz)."""
    g = _G(snap.geom)
    x, y, z, m, _ = _arrs(snap)
    g1 = np.ascontiguousarray(g1, np.float64)
    g11 = np.ascontiguousarray(g11 if g11.size else np.zeros(1), np.float64)
    outs = [np.zeros(len(x), np.float64) for _ in range(4)]
    rc = lib().orc_adjoint_b64(g.ref, len(x), _p(x, _fp), _p(y, _fp), _p(z, _fp), _p(m, _fp),
                               _p(g1, _dp), _p(g11, _dp), *[_p(o, _dp) for o in outs], nthreads)
    if rc != 0:
        raise ValueError("oracle rejected the input")
    return tuple(outs)


def loss_one(snap, i: int, g1: np.ndarray, g11: np.ndarray, pos=None) -> float:
    """FP64 loss sum(g*rho) of particle i's own TSC deposit (optionally at a
    moved position pos = (x, y, z))."""
    g = _G(snap.geom)
    x, y, z = pos if pos is not None else (snap.x[i], snap.y[i], snap.z[i])
    g1 = np.ascontiguousarray(g1, np.float64)
    g11 = np.ascontiguousarray(g11 if g11.size else np.zeros(1), np.float64)
    return float(lib().orc_loss_one_b64(g.ref, np.float32(x), np.float32(y), np.float32(z),
                                         np.float32(snap.m[i]), _p(g1, _dp), _p(g11, _dp)))


def pairs(snap, mode: str = "B", all_members: bool = False) -> np.ndarray:
    """Membership records [n_pairs, 5] = (grid, particle, ix, jy, kz)."""
    g = _G(snap.geom)
    x, y, z, _, esp = _arrs(snap)
    md = 2 if mode == "B" else (3 if all_members else 1)
    L = lib()
    cnt = L.orc_pairs(g.ref, md, len(x), _p(x, _fp), _p(y, _fp), _p(z, _fp), _p(esp, _ip), None, 0)
    out = np.zeros((max(cnt, 1), 5), np.int32)
    L.orc_pairs(g.ref, md, len(x), _p(x, _fp), _p(y, _fp), _p(z, _fp), _p(esp, _ip),
                _p(out, _ip), cnt)
    return out[:cnt]


def npairs(snap, mode: str = "B") -> int:
    g = _G(snap.geom)
    x, y, z, _, esp = _arrs(snap)
    md = 2 if mode == "B" else 3
    return int(lib().orc_pairs(g.ref, md, len(x), _p(x, _fp), _p(y, _fp), _p(z, _fp),
                               _p(esp, _ip), None, 0))


def stencil_work_a(snap) -> np.ndarray:
    g = _G(snap.geom)
    x, y, z, _, esp = _arrs(snap)
    out = np.zeros(snap.geom.nl + 1, np.int64)
    lib().orc_stencil_work_a(g.ref, len(x), _p(x, _fp), _p(y, _fp), _p(z, _fp), _p(esp, _ip),
                             out.ctypes.data_as(C.POINTER(C.c_int64)))
    return out


# ---------------------------------------------------------------------------
# the reference's own routines (oracle/_ref)
# ---------------------------------------------------------------------------
def species32(x, y, z, m, nmax: int, dx: float, rodo: float, re0: float):
    """SORT_DM_PARTICLES_LOCALDENSITY (reader.f:140-336): (species[n],
    npart_esp[9], order[n]) on an nmax^3 grid of cell dx (the reference's
    NMAX = NMAY = NMAZ; DX = DY = DZ)."""
    x, y, z, m = (np.ascontiguousarray(a, np.float32) for a in (x, y, z, m))
    n = len(x)
    sp = np.zeros(max(n, 1), np.int32)
    esp = np.zeros(9, np.int32)
    order = np.zeros(max(n, 1), np.int64)
    L = lib()
    L.orc_species32.restype = C.c_int
    rc = L.orc_species32(C.c_int(nmax), C.c_int(nmax), C.c_int(nmax), C.c_float(dx), C.c_float(dx),
                         C.c_float(dx), C.c_float(rodo), C.c_float(re0), C.c_int64(n), _p(x, _fp),
                         _p(y, _fp), _p(z, _fp), _p(m, _fp), _p(sp, _ip), _p(esp, _ip),
                         _p(order, C.POINTER(C.c_int64)))
    if rc != 0:
        raise MemoryError("oracle species grid")
    return sp[:n], esp, order[:n]


def run_ref_species(x, y, z, m, nx: int, lado0: float, rodo: float, re0: float):
    """The reference's SORT_DM_PARTICLES_LOCALDENSITY (its compile-time
    NMAX = 128 grid; DX from MALLA for this NX and LADO0): (dx, npart_esp,
    order[n] 0-based)."""
    x, y, z, m = (np.ascontiguousarray(a, np.float32) for a in (x, y, z, m))
    n = len(x)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([3, nx, nx, nx, 0, n, 2, 0], np.int32).tofile(f)
            np.zeros(1, np.int32).tofile(f)          # NPATCH(0)
            np.zeros(9, np.int32).tofile(f)          # NPART_ESP (output)
            np.array([lado0, rodo, re0, 1.0], np.float32).tofile(f)
            for a in (x, y, z, m):
                a.tofile(f)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([REF_BIN, fin, fout], check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, env=env)
        raw = np.fromfile(fout, np.uint8)
    d = np.frombuffer(raw[:12].tobytes(), np.float32)
    esp = np.frombuffer(raw[12:48].tobytes(), np.int32)
    order = np.frombuffer(raw[48:48 + 4 * n].tobytes(), np.int32).astype(np.int64) - 1
    return float(d[0]), esp.copy(), order


def ref_available() -> bool:
    return os.path.exists(REF_BIN)


def run_ref(snap, mode: str = "A"):
    """Run the reference's DENSITY (mode A) or INTERPOLATE_DENSITY_TSC (mode B)
    on this snapshot: returns (dx, dy, dz, radx, u1[nz,ny,nx], u11 packed)."""
    geom = snap.geom
    x, y, z, m, esp = _arrs(snap)
    P = geom.npatch_total
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([1 if mode == "A" else 2, geom.nx, geom.ny, geom.nz, geom.nl, len(x),
                      geom.interp_degree, P], np.int32).tofile(f)
            geom.npatch.astype(np.int32).tofile(f)
            esp.astype(np.int32).tofile(f)
            np.array([geom.lado0, geom.rote, geom.rete, geom.fdm], np.float32).tofile(f)
            for a in (geom.pnx, geom.pny, geom.pnz):
                a.astype(np.int32).tofile(f)
            for a in (geom.prx, geom.pry, geom.prz):
                a.astype(np.float32).tofile(f)
            for a in (x, y, z, m):
                a.tofile(f)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([REF_BIN, fin, fout], check=True, stdout=subprocess.DEVNULL, env=env)
        raw = np.fromfile(fout, np.float32)
    d = raw[:3]
    radx = raw[3:3 + geom.nx + 2]
    o = 3 + geom.nx + 2
    nc = geom.ncells_base
    u1 = raw[o:o + nc].reshape(geom.nz, geom.ny, geom.nx)
    u11 = raw[o + nc:]
    return d[0], d[1], d[2], radx, u1, u11


def argsort32(a):
    """ARGSORT (num.f:150-200) restated: 1-based indices sorting a (FP32)
    increasingly, the reference's merge order on ties."""
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty(len(a), np.int32)
    L = lib()
    L.orc_argsort32.restype = None
    L.orc_argsort32(C.c_int64(len(a)), _p(a, _fp), _p(out, _ip))
    return out


def peaks32(fields, contrastec: float, masks=None):
    """Peak candidates of HALOFIND_GRID (haloes_grids.f:316-404, 700-790) on
    a list of grids (3-D FP32 arrays indexed [z, y, x]; the base grid alone or
    one level's patches in patch order), with optional CONTA1 masks (int32,
    same shapes), sorted as SORT_CELLS_AMR (haloes_grids.f:1383-1425).
    Returns (dens, ix, iy, iz, grid): grid is the position in `fields`;
    ix, iy, iz are 1-based."""
    shapes = [np.asarray(f).shape for f in fields]
    n1 = np.array([sh[2] for sh in shapes], np.int32)
    n2 = np.array([sh[1] for sh in shapes], np.int32)
    n3 = np.array([sh[0] for sh in shapes], np.int32)
    sizes = n1.astype(np.int64) * n2 * n3
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    flat = np.concatenate([np.ascontiguousarray(f, np.float32).ravel() for f in fields]) if fields else np.zeros(1, np.float32)
    mk = None
    if masks is not None:
        mk = np.concatenate([np.ascontiguousarray(q, np.int32).ravel() for q in masks])
    cap = max(int(sizes.sum()) // 8 + 16, 16)
    L = lib()
    L.orc_peaks32.restype = C.c_int64
    while True:
        outs = [np.empty(cap, np.float32)] + [np.empty(cap, np.int32) for _ in range(4)]
        k = L.orc_peaks32(C.c_int(len(fields)), _p(n1, _ip), _p(n2, _ip), _p(n3, _ip),
                          _p(off, C.POINTER(C.c_int64)), _p(flat, _fp),
                          _p(mk, _ip) if mk is not None else None, C.c_float(contrastec),
                          C.c_int64(cap), _p(outs[0], _fp), *[_p(o, _ip) for o in outs[1:]])
        if k <= cap:
            return tuple(o[:k] for o in outs)
        cap = int(k)


def run_ref_sort_cells(ddd, dddx, dddy, dddz, dddp):
    """The reference's SORT_CELLS_AMR (haloes_grids.f:1383-1425) on the given
    candidate arrays: returns the five arrays after its sort."""
    ddd = np.ascontiguousarray(ddd, np.float32)
    ints = [np.ascontiguousarray(a, np.int32) for a in (dddx, dddy, dddz, dddp)]
    k = len(ddd)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([4, 0, 0, 0, 0, k, 0, 0], np.int32).tofile(f)
            ddd.tofile(f)
            for a in ints:
                a.tofile(f)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([REF_BIN, fin, fout], check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, env=env)
        raw = np.fromfile(fout, np.uint8)
    d = np.frombuffer(raw[:4 * k].tobytes(), np.float32).copy()
    rest = [np.frombuffer(raw[4 * k * (i + 1):4 * k * (i + 2)].tobytes(), np.int32).copy() for i in range(4)]
    return (d, *rest)


def _patch_arrays(geom):
    need = (geom.pcx, geom.pcy, geom.pcz, geom.pare)
    if any(a is None for a in need):
        raise ValueError("the geometry carries no PATCHX/Y/Z or PARE")
    return [np.ascontiguousarray(a, np.int32) for a in need]


def cr0amr(geom):
    """COMPUTE_CR0AMR (grids.f:2076-2226): (cr0amr int32 [nz, ny, nx],
    cr0amr11 int32 packed per patch, x fastest)."""
    px, py, pz, pare = _patch_arrays(geom)
    c0 = np.empty(geom.ncells_base, np.int32)
    c11 = np.empty(max(int(geom.patch_cells().sum()), 1), np.int32)
    L = lib()
    L.orc_cr0amr.restype = C.c_int
    rc = L.orc_cr0amr(geom.nx, geom.ny, geom.nz, geom.nl, _p(geom.npatch, _ip), _p(geom.pnx, _ip),
                      _p(geom.pny, _ip), _p(geom.pnz, _ip), _p(px, _ip), _p(py, _ip), _p(pz, _ip),
                      _p(pare, _ip), _p(geom.prx, _fp), _p(geom.pry, _fp), _p(geom.prz, _fp),
                      C.c_float(geom.lado0), _p(c0, _ip), _p(c11, _ip))
    if rc != 0:
        raise ValueError("a mother cell falls outside its grid")
    return c0.reshape(geom.nz, geom.ny, geom.nx), c11[: int(geom.patch_cells().sum())]


def veinsgrid(geom):
    """VEINSGRID (grids.f:2229-2506) for levels 1..NL: SOLAP int32 packed."""
    s = np.empty(max(int(geom.patch_cells().sum()), 1), np.int32)
    L = lib()
    L.orc_veinsgrid.restype = C.c_int
    rc = L.orc_veinsgrid(geom.nl, _p(geom.npatch, _ip), _p(geom.pnx, _ip), _p(geom.pny, _ip),
                         _p(geom.pnz, _ip), _p(geom.prx, _fp), _p(geom.pry, _fp), _p(geom.prz, _fp),
                         C.c_float(geom.dx), C.c_float(geom.dy), C.c_float(geom.dz), C.c_float(geom.radx[1]),
                         C.c_float(geom.rady[1]), C.c_float(geom.radz[1]), _p(s, _ip))
    if rc != 0:
        raise ValueError("an overlap maps a cell outside its patch")
    return s[: int(geom.patch_cells().sum())]


def dump_bytes(geom, u1, c0, u11, c11, solap) -> bytes:
    """The FW2 density dump of asohf.f:530-545 as an unformatted sequential
    file with 4-byte record markers: U1, CR0AMR, then per patch U11,
    CR0AMR11, SOLAP (interiors, x fastest)."""
    out = []

    def rec(a):
        b = np.ascontiguousarray(a).tobytes()
        m = np.array([len(b)], np.int32).tobytes()
        out.extend((m, b, m))
    rec(np.asarray(u1, np.float32))
    rec(np.asarray(c0, np.int32))
    off = np.concatenate([[0], np.cumsum(geom.patch_cells())]).astype(np.int64)
    for p in range(geom.npatch_total):
        sl = slice(off[p], off[p + 1])
        rec(np.asarray(u11[sl], np.float32))
        rec(np.asarray(c11[sl], np.int32))
        rec(np.asarray(solap[sl], np.int32))
    return b"".join(out)


def parse_dump(raw: bytes, geom):
    """Records of a density dump: (u1, cr0amr, u11, cr0amr11, solap), the
    patch arrays packed."""
    recs, pos = [], 0
    while pos < len(raw):
        n = int(np.frombuffer(raw[pos:pos + 4], np.int32)[0])
        recs.append(raw[pos + 4:pos + 4 + n])
        assert raw[pos + 4 + n:pos + 8 + n] == raw[pos:pos + 4]
        pos += 8 + n
    u1 = np.frombuffer(recs[0], np.float32).reshape(geom.nz, geom.ny, geom.nx)
    c0 = np.frombuffer(recs[1], np.int32).reshape(geom.nz, geom.ny, geom.nx)
    P = geom.npatch_total
    u11 = np.concatenate([np.frombuffer(recs[2 + 3 * p], np.float32) for p in range(P)]) if P else np.zeros(0, np.float32)
    c11 = np.concatenate([np.frombuffer(recs[3 + 3 * p], np.int32) for p in range(P)]) if P else np.zeros(0, np.int32)
    so = np.concatenate([np.frombuffer(recs[4 + 3 * p], np.int32) for p in range(P)]) if P else np.zeros(0, np.int32)
    return u1, c0, u11, c11, so


def run_ref_dump(snap) -> bytes:
    """The reference's own sequence of asohf.f:515-545 on this snapshot
    (INTERPOLATE_DENSITY_TSC, RENORM_DENSITY, VEINSGRID for every level,
    COMPUTE_CR0AMR, the FW2 WRITE statements): the dump file's bytes."""
    import resource
    geom = snap.geom
    x, y, z, m, esp = _arrs(snap)
    px, py, pz, pare = _patch_arrays(geom)
    P = geom.npatch_total

    def big_stack():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([5, geom.nx, geom.ny, geom.nz, geom.nl, len(x), geom.interp_degree, P], np.int32).tofile(f)
            geom.npatch.astype(np.int32).tofile(f)
            esp.astype(np.int32).tofile(f)
            np.array([geom.lado0, geom.rote, geom.rete, geom.fdm], np.float32).tofile(f)
            for a in (geom.pnx, geom.pny, geom.pnz):
                a.astype(np.int32).tofile(f)
            for a in (geom.prx, geom.pry, geom.prz):
                a.astype(np.float32).tofile(f)
            for a in (x, y, z, m):
                a.tofile(f)
            for a in (px, py, pz, pare):
                a.tofile(f)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([REF_BIN, fin, fout], check=True, stdout=subprocess.DEVNULL, env=env,
                       preexec_fn=big_stack)
        with open(fout, "rb") as f:
            return f.read()


MESH_DEFAULTS = dict(refine_thr=3, min_patchsize=14, bor=4, boramr=0, bor_ovlp=0, frac_refinable=0.05)


def run_ref_create_mesh(x, y, z, m, npart_esp, nx: int, lado0: float, nl_mesh: int, mpapolev=None, **params):
    """The reference's CREATE_MESH (grids.f:2-707) on particles in reference
    order (species contiguous): a dict with nl_mesh, npatch[0:NLEVELS+1] and
    the patch arrays (pnx.., px.., pare, prx..)."""
    import resource
    prm = dict(MESH_DEFAULTS, **params)
    x, y, z, m = (np.ascontiguousarray(a, np.float32) for a in (x, y, z, m))
    n = len(x)
    mp = np.full(9, 10000, np.int32) if mpapolev is None else np.asarray(mpapolev, np.int32)

    def big_stack():
        resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([6, nx, nx, nx, nl_mesh, n, 2, 0], np.int32).tofile(f)
            np.zeros(nl_mesh + 1, np.int32).tofile(f)
            np.asarray(npart_esp, np.int32).tofile(f)
            np.array([lado0, 1.0, 1.0, 1.0], np.float32).tofile(f)
            for a in (x, y, z, m):
                a.tofile(f)
            np.array([prm["refine_thr"], prm["min_patchsize"], prm["bor"], prm["boramr"], prm["bor_ovlp"]],
                     np.int32).tofile(f)
            np.array([prm["frac_refinable"]], np.float32).tofile(f)
            mp.tofile(f)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        subprocess.run([REF_BIN, fin, fout], check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                       env=env, preexec_fn=big_stack)
        raw = np.fromfile(fout, np.uint8)
    i32 = lambda a, b: np.frombuffer(raw[a:b].tobytes(), np.int32)
    nlm = int(i32(0, 4)[0])
    npatch = i32(4, 44).copy()
    P = int(npatch.sum())
    out = {"nl_mesh": nlm, "npatch": npatch}
    pos = 44
    for k in ("pnx", "pny", "pnz", "px", "py", "pz", "pare"):
        out[k] = i32(pos, pos + 4 * P).copy()
        pos += 4 * P
    for k in ("prx", "pry", "prz"):
        out[k] = np.frombuffer(raw[pos:pos + 4 * P].tobytes(), np.float32).copy()
        pos += 4 * P
    return out
