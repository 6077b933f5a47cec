"""Python face of libasohf_hip: the density build on device tensors.

This mirrors the reference's DENSITY call (grids.f:710-713, called at
asohf.f:467-470): the same inputs (particle positions and masses in reference
order, NPART_ESP, the patch hierarchy, LADO0, the COMMON /ESPACIADO/ and
/BACK/ constants) and the same outputs (U1 on the base grid, U11 on every
patch), plus the FP64 deposited mass per grid.  Status codes of the C ABI
become AsohfError exceptions, the analogue of the reference's STOP 1 on bad
input.

PyTorch supplies device memory and streams; every computation runs in the HIP
library.  There is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
import dataclasses

import numpy as np
import torch

from . import _lib
from .geometry import N_ESP, Geometry

_MODES = {"table": _lib.MODE_TABLE, "A": _lib.MODE_TABLE, "tsc": _lib.MODE_TSC, "B": _lib.MODE_TSC}
_REDS = {"deterministic": _lib.REDUCE_DETERMINISTIC, "atomic": _lib.REDUCE_ATOMIC}
_PATHS = {"batched": _lib.PATH_BATCHED, "plain": _lib.PATH_PLAIN}


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _np_ptr(a, typ):
    return a.ctypes.data_as(C.POINTER(typ))


class _GeomArgs:
    """Keeps the host arrays behind a Geom struct alive for one call."""

    def __init__(self, g: Geometry, pass_rad: bool = True):
        self.keep = [g.npatch, g.pnx, g.pny, g.pnz, g.prx, g.pry, g.prz, g.radx, g.rady, g.radz]
        f, i = C.c_float, C.c_int32
        self.s = _lib.Geom(g.nx, g.ny, g.nz, g.nl, _np_ptr(g.npatch, i), _np_ptr(g.pnx, i),
                           _np_ptr(g.pny, i), _np_ptr(g.pnz, i), _np_ptr(g.prx, f), _np_ptr(g.pry, f),
                           _np_ptr(g.prz, f), float(g.dx), float(g.dy), float(g.dz), float(g.lado0),
                           float(g.rote), float(g.rete), float(g.fdm), int(g.interp_degree),
                           _np_ptr(g.radx, f) if pass_rad else None,
                           _np_ptr(g.rady, f) if pass_rad else None,
                           _np_ptr(g.radz, f) if pass_rad else None)

    @property
    def ref(self):
        return C.byref(self.s)


@dataclasses.dataclass
class DensityResult:
    u1: torch.Tensor            # base field, [nz, ly, lx] (lx, ly = leading dims)
    u11: torch.Tensor           # patch fields, packed (or padded) 1-D
    grid_mass: torch.Tensor | None   # FP64 [1+P], deposited mass per grid


def _check_particles(*ts):
    n = None
    for t in ts:
        if t is None:
            continue
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.dim() == 1):
            raise ValueError("particle arrays must be contiguous 1-D float32 tensors on a HIP device")
        if n is None:
            n = t.numel()
        elif t.numel() != n:
            raise ValueError("particle arrays differ in length")
    return n or 0


class DensityBuilder:
    """One library context (one device, one stream) and its workspaces."""

    def __init__(self, device: int | None = None):
        L = _lib.load()
        if device is None:
            device = torch.cuda.current_device()
        self.device = int(device)
        h = C.c_void_p()
        _lib.check(None, L.asohf_hip_create(C.byref(h), self.device))
        self._h = h

    def species(self, x, y, z, m, nmax=(128, 128, 128), d=(1.0, 1.0, 1.0), rodo=1.0, re0=1.0,
                want_species=True):
        """SORT_DM_PARTICLES_LOCALDENSITY (reader.f:140-336): each particle's
        kernel class from the local density on the reference's NMAX^3 grid of
        cell d, and the stable partition by class.  Returns (species int32
        tensor or None, order int64 tensor: original index at each position,
        npart_esp numpy int32[9]).  Apply with x[order] etc."""
        n = _check_particles(x, y, z, m)
        dev = x.device
        sp = torch.empty(max(n, 1), dtype=torch.int32, device=dev) if want_species else None
        order = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
        esp = np.zeros(N_ESP, np.int32)
        nm = np.ascontiguousarray(nmax, np.int32)
        dd = np.ascontiguousarray(d, np.float32)
        st = _lib.load().asohf_hip_species(
            self._h, _np_ptr(nm, C.c_int32), dd.ctypes.data_as(C.POINTER(C.c_float)), C.c_float(rodo),
            C.c_float(re0), n, _ptr(x), _ptr(y), _ptr(z), _ptr(m), _ptr(sp) if sp is not None else None,
            _ptr(order), _np_ptr(esp, C.c_int32), C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        _lib.check(self._h, st)
        return (sp[:n] if sp is not None else None), order[:n], esp

    def peak_candidates(self, geom: Geometry, level: int, u1=None, u11=None, conta=None, *,
                        contrastec: float, u1_ld=None, u11_ld=None, cap=None):
        """Peak candidates of the grid halo finder on one level
        (HALOFIND_GRID, haloes_grids.f:316-404 / 700-790), sorted as
        SORT_CELLS_AMR (haloes_grids.f:1383-1425): cells at or above
        CONTRASTEC/6 that exceed their six face neighbours, away from the
        grid border, with CONTA1 (conta, int32, U11's layout) equal to 1 on
        AMR levels.  Returns (dens, ix, iy, iz, patch) device tensors; patch
        is the 1-based global patch number, 0 on the base grid."""
        dev = (u1 if level == 0 else u11).device
        if cap is None:
            cap = 1 << 16
        o = self._opts("tsc", u1_ld=u1_ld, u11_ld=u11_ld)
        ga = _GeomArgs(geom)
        L = _lib.load()
        while True:
            outs = [torch.empty(max(cap, 1), dtype=torch.float32, device=dev)] + \
                   [torch.empty(max(cap, 1), dtype=torch.int32, device=dev) for _ in range(4)]
            cnt = C.c_int64()
            st = L.asohf_hip_peak_candidates(self._h, ga.ref, int(level),
                                             _ptr(u1) if u1 is not None else None,
                                             _ptr(u11) if u11 is not None else None,
                                             _ptr(conta) if conta is not None else None, C.c_float(contrastec),
                                             *[_ptr(t) for t in outs], cap, C.byref(cnt), C.byref(o))
            if st == _lib.ASOHF_ERR_CAPACITY and cnt.value > cap:
                cap = cnt.value
                continue
            _lib.check(self._h, st)
            return tuple(t[: cnt.value] for t in outs)

    def refinement_masks(self, geom: Geometry, *, u1_ld=None, u11_ld=None, device=None):
        """COMPUTE_CR0AMR (grids.f:2076-2226): (cr0amr int32 [nz, ly, lx],
        cr0amr11 int32 in u11's layout), 1 = not refined, 0 = refined."""
        if geom.pcx is None or geom.pare is None:
            raise ValueError("the geometry carries no PATCHX/Y/Z or PARE")
        dev = torch.device(device) if device is not None else self.device
        s1, n11 = self.out_shapes(geom, u1_ld, u11_ld)
        c0 = torch.empty(s1, dtype=torch.int32, device=dev)
        c11 = torch.empty(n11, dtype=torch.int32, device=dev)
        pc = [np.ascontiguousarray(a, np.int32) for a in (geom.pcx, geom.pcy, geom.pcz, geom.pare)]
        ga = _GeomArgs(geom)
        o = self._opts("tsc", u1_ld=u1_ld, u11_ld=u11_ld)
        st = _lib.load().asohf_hip_refinement_masks(self._h, ga.ref, *[_np_ptr(a, C.c_int32) for a in pc],
                                                     _ptr(c0), _ptr(c11), C.byref(o))
        _lib.check(self._h, st)
        return c0, c11

    def overlap_mask(self, geom: Geometry, *, u11_ld=None, device=None):
        """VEINSGRID (grids.f:2229-2506) for every level: SOLAP int32 in
        u11's layout."""
        dev = torch.device(device) if device is not None else self.device
        _, n11 = self.out_shapes(geom, None, u11_ld)
        so = torch.empty(n11, dtype=torch.int32, device=dev)
        ga = _GeomArgs(geom)
        o = self._opts("tsc", u11_ld=u11_ld)
        _lib.check(self._h, _lib.load().asohf_hip_overlap_mask(self._h, ga.ref, _ptr(so), C.byref(o)))
        return so

    def write_density_dump(self, path, geom: Geometry, u1, cr0amr, u11, cr0amr11, solap, *, u1_ld=None,
                           u11_ld=None):
        """The FW2 density dump of asohf.f:530-545 (unformatted sequential
        records: U1, CR0AMR, then per patch U11, CR0AMR11, SOLAP)."""
        ga = _GeomArgs(geom)
        o = self._opts("tsc", u1_ld=u1_ld, u11_ld=u11_ld)
        st = _lib.load().asohf_hip_write_density_dump(self._h, os.fsencode(str(path)), ga.ref, _ptr(u1),
                                                       _ptr(cr0amr), _ptr(u11), _ptr(cr0amr11), _ptr(solap),
                                                       C.byref(o))
        _lib.check(self._h, st)

    def create_mesh(self, base: Geometry, x, y, z, npart_esp, *, nl_mesh=4, refine_thr=3, min_patchsize=14,
                    frac_refinable=0.05, bor=4, boramr=0, bor_ovlp=0, namr=(32, 32, 32), npalev=10000,
                    mpapolev=None, cap=None):
        """CREATE_MESH (grids.f:2-707): the AMR hierarchy the reference builds
        for these particles (device FP32, reference order; npart_esp gives
        the species blocks).  Defaults are asohf.dat's mesh block and the
        reference build's NAMRX..Z = 32, NPALEV = 10000.  Returns a Geometry
        on `base`'s grid with PATCHX..Z and PARE."""
        n = _check_particles(x, y, z)
        prm = _lib.MeshParams()
        prm.nl_mesh, prm.refine_thr, prm.min_patchsize = nl_mesh, refine_thr, min_patchsize
        prm.frac_refinable, prm.bor, prm.boramr, prm.bor_ovlp = frac_refinable, bor, boramr, bor_ovlp
        prm.namr[0], prm.namr[1], prm.namr[2] = namr
        prm.npalev = npalev
        mp = [npalev] * 9 if mpapolev is None else list(mpapolev)
        for k in range(9):
            prm.mpapolev[k] = int(mp[k])
        esp = np.ascontiguousarray(npart_esp, np.int32)
        cap = npalev if cap is None else cap
        ints = [np.zeros(max(cap, 1), np.int32) for _ in range(7)]
        flts = [np.zeros(max(cap, 1), np.float32) for _ in range(3)]
        nl = C.c_int32()
        npatch = np.zeros(10, np.int32)
        ga = _GeomArgs(base)
        o = self._opts("tsc")
        st = _lib.load().asohf_hip_create_mesh(
            self._h, ga.ref, n, _ptr(x), _ptr(y), _ptr(z), _np_ptr(esp, C.c_int32), C.byref(prm), C.byref(nl),
            _np_ptr(npatch, C.c_int32), *[_np_ptr(a, C.c_int32) for a in ints],
            *[a.ctypes.data_as(C.POINTER(C.c_float)) for a in flts], cap, C.byref(o))
        _lib.check(self._h, st)
        P = int(npatch[1:nl.value + 1].sum())
        pnx, pny, pnz, px, py, pz, pare = (a[:P] for a in ints)
        prx, pry, prz = (a[:P] for a in flts)
        return Geometry(base.nx, base.ny, base.nz, base.lado0, base.rote, base.rete,
                        npatch[:nl.value + 1].copy(), pnx, pny, pnz, prx, pry, prz, pare=pare,
                        fdm=base.fdm, interp_degree=base.interp_degree, pcx=px, pcy=py, pcz=pz)

    def release_workspaces(self):
        """Free the device workspaces kept between calls (the plan stays)."""
        _lib.check(self._h, _lib.load().asohf_hip_release_workspaces(self._h))

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().asohf_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ options
    def _opts(self, mode, reduction="deterministic", path="batched", u1_ld=None, u11_ld=None,
              use_graph=False):
        o = _lib.Opts()
        o.mode = _MODES[mode]
        o.reduction = _REDS[reduction]
        o.path = _PATHS[path]
        o.use_graph = 1 if use_graph else 0
        if u1_ld:
            o.u1_ld[0], o.u1_ld[1] = u1_ld
        if u11_ld:
            o.u11_ld[0], o.u11_ld[1], o.u11_ld[2] = u11_ld
        o.stream = C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)
        return o

    @staticmethod
    def out_shapes(geom: Geometry, u1_ld=None, u11_ld=None):
        lx, ly = u1_ld if u1_ld else (geom.nx, geom.ny)
        n11 = (geom.npatch_total * u11_ld[0] * u11_ld[1] * u11_ld[2] if u11_ld
               else int(geom.patch_cells().sum()))
        return (geom.nz, ly, lx), max(n11, 1)

    # ------------------------------------------------------------------ forward
    def density(self, geom: Geometry, x, y, z, m, npart_esp=None, *, mode="table",
                reduction="deterministic", path="batched", u1=None, u11=None, want_mass=True,
                grid_mass=None, u1_ld=None, u11_ld=None, use_graph=False) -> DensityResult:
        """U1/U11 of DENSITY (mode "table") or INTERPOLATE_DENSITY_TSC (mode
        "tsc"), both divided by FDM as RENORM_DENSITY does.

        use_graph: run the batched build as one hipGraph (captured on first use,
        replayed while the geometry, sizes and every array stay the same; pass
        u1, u11 and grid_mass to keep the arrays the same)."""
        n = _check_particles(x, y, z, m)
        dev = x.device
        s1, n11 = self.out_shapes(geom, u1_ld, u11_ld)
        if u1 is None:
            u1 = torch.empty(s1, dtype=torch.float32, device=dev)
        if u11 is None:
            u11 = torch.empty(n11, dtype=torch.float32, device=dev)
        gm = grid_mass
        if gm is None and want_mass:
            gm = torch.empty(1 + geom.npatch_total, dtype=torch.float64, device=dev)
        esp = None
        if _MODES[mode] == _lib.MODE_TABLE:
            if npart_esp is None:
                raise ValueError("table mode needs NPART_ESP")
            esp = np.ascontiguousarray(npart_esp, dtype=np.int32)
            if esp.shape != (N_ESP,):
                raise ValueError("NPART_ESP must have 9 entries")
        ga = _GeomArgs(geom)
        o = self._opts(mode, reduction, path, u1_ld, u11_ld, use_graph)
        st = _lib.load().asohf_hip_density(
            self._h, ga.ref, n, _ptr(x), _ptr(y), _ptr(z), _ptr(m),
            _np_ptr(esp, C.c_int32) if esp is not None else None, _ptr(u1), _ptr(u11), _ptr(gm),
            C.byref(o))
        _lib.check(self._h, st)
        return DensityResult(u1, u11, gm)

    def density_host(self, geom: Geometry, x, y, z, m, npart_esp=None, *, mode="table",
                     reduction="deterministic", u1=None, u11=None, u1_ld=None, u11_ld=None):
        """The host-array entry point (the one the Fortran shim calls): numpy in, numpy out."""
        s1, n11 = self.out_shapes(geom, u1_ld, u11_ld)
        u1 = np.zeros(s1, np.float32) if u1 is None else u1
        u11 = np.zeros(n11, np.float32) if u11 is None else u11
        gm = np.zeros(1 + geom.npatch_total, np.float64)
        x, y, z, m = (np.ascontiguousarray(a, np.float32) for a in (x, y, z, m))
        esp = np.ascontiguousarray(npart_esp if npart_esp is not None else np.zeros(9), np.int32)
        ga = _GeomArgs(geom)
        o = self._opts(mode, reduction, "batched", u1_ld, u11_ld)
        o.stream = None
        f = C.c_float
        st = _lib.load().asohf_hip_density_host(
            self._h, ga.ref, len(x), x.ctypes.data_as(C.c_void_p), y.ctypes.data_as(C.c_void_p),
            z.ctypes.data_as(C.c_void_p), m.ctypes.data_as(C.c_void_p), _np_ptr(esp, C.c_int32),
            u1.ctypes.data_as(C.c_void_p), u11.ctypes.data_as(C.c_void_p),
            gm.ctypes.data_as(C.c_void_p), C.byref(o))
        _lib.check(self._h, st)
        del f
        return u1, u11, gm

    # ------------------------------------------------------------------ adjoint
    def adjoint(self, geom: Geometry, x, y, z, m, g1, g11, *, u1_ld=None, u11_ld=None):
        """dL/dm, dL/dx, dL/dy, dL/dz of L = sum(g1*u1) + sum(g11*u11), TSC mode.
        g11=None stands for a zero gradient on every patch cell."""
        n = _check_particles(x, y, z, m)
        g1 = g1.contiguous().float()
        if g11 is None:
            g11 = torch.zeros(self.out_shapes(geom, u1_ld, u11_ld)[1], dtype=torch.float32, device=x.device)
        g11 = g11.contiguous().float()
        outs = [torch.empty(n, dtype=torch.float32, device=x.device) for _ in range(4)]
        ga = _GeomArgs(geom)
        o = self._opts("tsc", "deterministic", "batched", u1_ld, u11_ld)
        st = _lib.load().asohf_hip_density_adjoint(
            self._h, ga.ref, n, _ptr(x), _ptr(y), _ptr(z), _ptr(m), _ptr(g1), _ptr(g11),
            *[_ptr(t) for t in outs], C.byref(o))
        _lib.check(self._h, st)
        return tuple(outs)

    # ------------------------------------------------------------------ helpers
    def membership(self, geom: Geometry, x, y, z, npart_esp=None, *, mode="tsc"):
        """(grid, particle, ix, jy, kz) of every box member, as the binning
        stage computes them (int32 [count, 5], unordered)."""
        n = _check_particles(x, y, z)
        esp = np.ascontiguousarray(npart_esp if npart_esp is not None else np.zeros(9), np.int32)
        ga = _GeomArgs(geom)
        L = _lib.load()
        cnt = C.c_int64()
        st = L.asohf_hip_membership(self._h, ga.ref, n, _ptr(x), _ptr(y), _ptr(z),
                                    _np_ptr(esp, C.c_int32), _MODES[mode], None, 0, C.byref(cnt))
        _lib.check(self._h, st)
        out = torch.empty((max(cnt.value, 1), 5), dtype=torch.int32, device=x.device)
        st = L.asohf_hip_membership(self._h, ga.ref, n, _ptr(x), _ptr(y), _ptr(z),
                                    _np_ptr(esp, C.c_int32), _MODES[mode], _ptr(out), cnt.value,
                                    C.byref(cnt))
        _lib.check(self._h, st)
        return out[: cnt.value]

    def get_tune(self) -> dict:
        t = _lib.Tune()
        _lib.check(self._h, _lib.load().asohf_hip_get_tune(self._h, C.byref(t)))
        return t.as_dict()

    def set_tune(self, tile=(16, 8, 8), block=256, chunk=4096, waves_per_cu=16):
        t = _lib.Tune()
        t.tile[0], t.tile[1], t.tile[2] = tile
        t.block, t.chunk, t.waves_per_cu = block, chunk, waves_per_cu
        _lib.check(self._h, _lib.load().asohf_hip_set_tune(self._h, C.byref(t)))

    def autotune(self, geom: Geometry, x, y, z, m, npart_esp=None, *, mode="tsc",
                 reduction="deterministic") -> dict:
        n = _check_particles(x, y, z, m)
        s1, n11 = self.out_shapes(geom)
        u1 = torch.empty(s1, dtype=torch.float32, device=x.device)
        u11 = torch.empty(n11, dtype=torch.float32, device=x.device)
        esp = np.ascontiguousarray(npart_esp if npart_esp is not None else np.zeros(9), np.int32)
        ga = _GeomArgs(geom)
        o = self._opts(mode, reduction)
        t = _lib.Tune()
        st = _lib.load().asohf_hip_autotune(self._h, ga.ref, n, _ptr(x), _ptr(y), _ptr(z), _ptr(m),
                                            _np_ptr(esp, C.c_int32), _ptr(u1), _ptr(u11), C.byref(o),
                                            C.byref(t))
        _lib.check(self._h, st)
        return t.as_dict()

    def stats(self) -> dict:
        s = _lib.Stats()
        _lib.check(self._h, _lib.load().asohf_hip_get_stats(self._h, C.byref(s)))
        return s.as_dict()


_builders: dict[int, DensityBuilder] = {}


def builder(device: int | None = None) -> DensityBuilder:
    d = torch.cuda.current_device() if device is None else int(device)
    if d not in _builders:
        _builders[d] = DensityBuilder(d)
    return _builders[d]


def density(geom: Geometry, x, y, z, m, npart_esp=None, **kw) -> DensityResult:
    """Module-level DENSITY on the tensors' device."""
    return builder(x.device.index).density(geom, x, y, z, m, npart_esp, **kw)


class TSCDeposit(torch.autograd.Function):
    """TSC deposit (INTERPOLATE_DENSITY_TSC semantics) with the hand-written
    adjoint: d(loss)/d(positions, masses).  Weights are recomputed from the
    positions in the backward pass, not stored."""

    @staticmethod
    def forward(ctx, x, y, z, m, geom, bld, reduction):
        res = bld.density(geom, x, y, z, m, None, mode="tsc", reduction=reduction, want_mass=False)
        ctx.save_for_backward(x, y, z, m)
        ctx.geom, ctx.bld = geom, bld
        return res.u1, res.u11

    @staticmethod
    def backward(ctx, g1, g11):
        x, y, z, m = ctx.saved_tensors
        if g1 is None:
            g1 = torch.zeros((ctx.geom.nz, ctx.geom.ny, ctx.geom.nx), dtype=torch.float32, device=x.device)
        dm, dx, dy, dz = ctx.bld.adjoint(ctx.geom, x, y, z, m, g1, g11)
        return dx, dy, dz, dm, None, None, None


def tsc_deposit(x, y, z, m, geom: Geometry, reduction: str = "deterministic", bld=None):
    """Differentiable TSC density: (u1 [nz,ny,nx], u11 packed)."""
    bld = bld or builder(x.device.index)
    return TSCDeposit.apply(x, y, z, m, geom, bld, reduction)
