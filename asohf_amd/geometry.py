"""AMR geometry of one density build: the base grid, the patch hierarchy and the
cosmological constants the density build reads.

The fields mirror the arguments and COMMON blocks the reference's DENSITY reads
(grids.f:710-742): NX/NY/NZ, NL, NPATCH(0:NL), PATCHNX/NY/NZ, PATCHRX/RY/RZ,
PARE, LADO0, COMMON /ESPACIADO/ DX,DY,DZ and COMMON /BACK/ RETE,ROTE.  Patch
numbering is global and level-major as in the reference: level IR holds patches
SUM(NPATCH(0:IR-1))+1 .. SUM(NPATCH(0:IR)) (grids.f:846-847), here 0-based.
"""
from __future__ import annotations

import dataclasses
import hashlib

import numpy as np

N_ESP = 9        # asohf_parameters.dat: number of particle species (kernel classes)
NLEVELS = 9      # asohf_parameters.dat: maximum number of AMR levels

f32 = np.float32


def malla(n: int, lado) -> tuple[np.ndarray, np.float32]:
    """Cell centres RADX(0:N+1) and cell size, as MALLA builds them in FP32
    (grids.f:2509-2573): A=-LADO/2, C=(B-A)/(N-1), RADX(I)=RADX(1)+(I-1)*C,
    ghost centres RADX(0) and RADX(N+1), DX=RADX(2)-RADX(1)."""
    lado = f32(lado)
    a = f32(-lado / f32(2.0))
    b = f32(lado / f32(2.0))
    c = f32((b - a) / f32(n - 1))
    rad = np.zeros(n + 2, dtype=np.float32)
    rad[1] = a
    i = np.arange(2, n + 1, dtype=np.float32)
    rad[2:n + 1] = rad[1] + (i - f32(1.0)) * c
    rad[0] = f32(rad[1] - c)
    rad[n + 1] = f32(rad[n] + c)
    return rad, f32(rad[2] - rad[1])


def background(h: float, omega_m: float, zeta: float = 0.0) -> tuple[np.float32, np.float32]:
    """ROTE and RETE for redshift zeta (asohf.f:351-358, 447-448):
    ACHE = h*3.66e-3, RODO = OMEGA0*3*ACHE**2, RE0 = 1/10.98,
    ROTE = RODO*(1+z)**3, RETE = RE0/(1+z).  ACHE, RODO and RE0 are REAL*4
    (asohf.f:28-30) computed from FP64 products; ROTE/RETE are FP32."""
    ache = f32(float(f32(h)) * 3.66e-3)
    rodo = f32(float(f32(omega_m)) * 3.0 * float(ache) ** 2)
    re0 = f32(f32(1.0) / f32(10.98))
    opz = f32(f32(1.0) + f32(zeta))
    rote = f32(rodo * f32(opz * f32(opz * opz)))
    rete = f32(re0 / opz)
    return rote, rete


@dataclasses.dataclass
class Geometry:
    nx: int
    ny: int
    nz: int
    lado0: np.float32
    rote: np.float32
    rete: np.float32
    npatch: np.ndarray            # int32 [nl+1]
    pnx: np.ndarray               # int32 [P]
    pny: np.ndarray
    pnz: np.ndarray
    prx: np.ndarray               # float32 [P]
    pry: np.ndarray
    prz: np.ndarray
    pare: np.ndarray | None = None   # int32 [P], 1-based parent, 0 = base grid
    fdm: np.float32 = f32(1.0)
    interp_degree: int = 2
    pcx: np.ndarray | None = None    # int32 [P], PATCHX/Y/Z: the parent cell (1-based)
    pcy: np.ndarray | None = None    # holding the patch's first cells
    pcz: np.ndarray | None = None

    def __post_init__(self):
        self.npatch = np.ascontiguousarray(self.npatch, dtype=np.int32)
        for k in ("pnx", "pny", "pnz"):
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=np.int32))
        for k in ("prx", "pry", "prz"):
            setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=np.float32))
        if self.pare is not None:
            self.pare = np.ascontiguousarray(self.pare, dtype=np.int32)
        for k in ("pcx", "pcy", "pcz"):
            if getattr(self, k) is not None:
                setattr(self, k, np.ascontiguousarray(getattr(self, k), dtype=np.int32))
        self.lado0 = f32(self.lado0)
        self.rote = f32(self.rote)
        self.rete = f32(self.rete)
        self.fdm = f32(self.fdm)
        lado = f32(self.lado0 - self.lado0 / f32(self.nx))   # asohf.f:323
        self.radx, self.dx = malla(self.nx, lado)
        self.rady, self.dy = malla(self.ny, lado)
        self.radz, self.dz = malla(self.nz, lado)

    @property
    def nl(self) -> int:
        return len(self.npatch) - 1

    @property
    def npatch_total(self) -> int:
        return int(self.npatch.sum())

    def level_of_patch(self) -> np.ndarray:
        """Level IR of every patch (0-based patch index)."""
        return np.repeat(np.arange(self.nl + 1, dtype=np.int32), self.npatch)

    def patch_cells(self) -> np.ndarray:
        return self.pnx.astype(np.int64) * self.pny * self.pnz

    def patch_offsets(self) -> np.ndarray:
        """Packed layout: patch p occupies cells [off[p], off[p+1])."""
        off = np.zeros(self.npatch_total + 1, dtype=np.int64)
        np.cumsum(self.patch_cells(), out=off[1:])
        return off

    @property
    def ncells_base(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def ncells_total(self) -> int:
        return self.ncells_base + int(self.patch_cells().sum())

    def truncated(self, levels: int) -> "Geometry":
        """The hierarchy cut to levels 0..levels (the patches of deeper levels
        dropped), as CREATE_MESH would leave it with NL = levels."""
        levels = max(0, min(levels, self.nl))
        k = int(self.npatch[: levels + 1].sum())
        return Geometry(self.nx, self.ny, self.nz, self.lado0, self.rote, self.rete,
                        self.npatch[: levels + 1].copy(), self.pnx[:k], self.pny[:k], self.pnz[:k],
                        self.prx[:k], self.pry[:k], self.prz[:k],
                        pare=None if self.pare is None else self.pare[:k],
                        fdm=self.fdm, interp_degree=self.interp_degree,
                        pcx=None if self.pcx is None else self.pcx[:k],
                        pcy=None if self.pcy is None else self.pcy[:k],
                        pcz=None if self.pcz is None else self.pcz[:k])

    def denbas(self, ir: int) -> np.float32:
        """DENBAS of level ir in FP32: DXPA*DYPA*DZPA*ROTE*RETE**3 (grids.f:824, 844)."""
        p2 = f32(2.0) ** ir
        dxpa, dypa, dzpa = f32(self.dx / p2), f32(self.dy / p2), f32(self.dz / p2)
        r3 = f32(self.rete * f32(self.rete * self.rete))
        return f32(f32(f32(f32(dxpa * dypa) * dzpa) * self.rote) * r3)

    def digest(self) -> str:
        h = hashlib.sha256()
        for a in (self.npatch, self.pnx, self.pny, self.pnz, self.prx, self.pry, self.prz):
            h.update(np.ascontiguousarray(a).tobytes())
        h.update(np.array([self.nx, self.ny, self.nz, self.interp_degree], np.int64).tobytes())
        h.update(np.array([self.lado0, self.rote, self.rete, self.fdm], np.float32).tobytes())
        return h.hexdigest()[:16]

    def check(self) -> None:
        """Host-side validation: the same capacity/shape guards the library
        applies before touching a device."""
        if min(self.nx, self.ny, self.nz) < 3:
            raise ValueError("base grid needs at least 3 cells per axis")
        if self.nl < 0 or self.nl > NLEVELS:
            raise ValueError(f"nl={self.nl} outside 0..{NLEVELS}")
        p = self.npatch_total
        for k in ("pnx", "pny", "pnz", "prx", "pry", "prz"):
            if len(getattr(self, k)) != p:
                raise ValueError(f"{k} has {len(getattr(self, k))} entries, expected {p}")
        if p and (min(self.pnx.min(), self.pny.min(), self.pnz.min()) < 1):
            raise ValueError("patch with a non-positive side")
        if self.interp_degree not in (1, 2):
            raise ValueError("interp_degree must be 1 or 2")
