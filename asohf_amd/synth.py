"""Seeded synthetic snapshots and AMR hierarchies for the density build.

Particles: NFW haloes with masses from dN/dM ~ M^-1.9 plus a uniform
background, wrapped periodically into the box-centred frame [-L/2, L/2) that
the reference uses internally (readers/particle_general.f:146-149).  Masses are
equal and set so that the box-mean density contrast is 1 (total mass =
ROTE*RETE**3*L**3).  With equal-mass particles the drawn halo masses cannot be
literal, so they act as weights: halo h receives n_h ~ M_h particles, and its
virial radius is the Delta=200 radius of n_h*m_p.  Velocities are not generated:
the density build never reads them.

Species (kernel classes) follow the local-density rule of
SORT_DM_PARTICLES_LOCALDENSITY (reader.f:140-336): NGP density contrast on the
base grid, s = clamp(int(log(rho)/log(8)), 0, 8), then a stable partition so the
species are contiguous, and NPART_ESP(0:8).  (The reference bins on an
NMAX-sized grid; here the grid is NX-sized.)

AMR hierarchy: patches that obey CREATE_MESH's geometric rules
(grids.f:186-313, 457-624): even sides, a level-1 patch starts on a base cell
I1 with PATCHRX = RADX(I1) (grids.f:300-302); a level-IR patch lies inside its
parent and starts on the parent's cell I1 with
PATCHRX = PATCHRX(parent) + (I1-1.5)*DXPA(parent) (grids.f:586-588); same-level
overlaps are allowed.  Placement follows the particle density: every patch is
centred near a halo drawn with probability ~ n_h, at a radius that shrinks with
the level.  Sides are log-uniform in [8, 64] fine cells per axis, clipped to
the parent.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .geometry import N_ESP, Geometry, background, f32


@dataclasses.dataclass(frozen=True)
class SnapshotSpec:
    name: str
    n_side: int                 # N = n_side**3 particles
    box: float                  # box side, Mpc/h (internal length unit)
    nx: int                     # base grid cells per axis
    patches_per_level: tuple    # target patch count per AMR level 1..NL
    n_halos: int
    f_halo: float
    m_lo: float
    m_hi: float
    seed: int
    h: float = 0.7
    omega_m: float = 0.3
    side_lo: int = 8
    side_hi: int = 64

    @property
    def n(self) -> int:
        return self.n_side ** 3

    @property
    def nl(self) -> int:
        return len(self.patches_per_level)


# The workloads of the density build (see DESIGN.md for the mapping).
CONFIGS = {
    1: SnapshotSpec("c1_128cubed_3lev", 128, 40.0, 128, (24, 18, 12), 200, 0.70, 1e11, 1e14, 1),
    2: SnapshotSpec("c2_256cubed_4lev", 256, 40.0, 256, (320, 260, 200, 140), 200, 0.70,
                    1e11, 1e14, 2),
    3: SnapshotSpec("c3_512cubed_6lev", 512, 200.0, 256, (5000, 4200, 3500, 3000, 2500, 1800),
                    5000, 0.80, 1e11, 1e14, 3),
    4: SnapshotSpec("c4_256cubed_single", 256, 40.0, 256, (), 200, 0.70, 1e11, 1e14, 2),
    5: SnapshotSpec("c5_1024cubed_7lev", 1024, 500.0, 1024,
                    (36000, 30000, 25000, 21000, 17000, 12000, 9000), 50000, 0.85,
                    1e10, 1e15, 5),
}


@dataclasses.dataclass
class Snapshot:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    m: np.ndarray
    npart_esp: np.ndarray       # int32 [9]
    geom: Geometry
    spec: SnapshotSpec | None = None

    @property
    def n(self) -> int:
        return len(self.x)


def _halo_catalogue(spec: SnapshotSpec, rng: np.random.Generator, n: int):
    a, b = spec.m_lo ** -0.9, spec.m_hi ** -0.9
    mass = (a + rng.random(spec.n_halos) * (b - a)) ** (-1.0 / 0.9)
    n_in = int(round(spec.f_halo * n))
    w = mass / mass.sum()
    nh = np.floor(w * n_in).astype(np.int64)
    rest = n_in - int(nh.sum())
    if rest > 0:
        nh[np.argsort(-(w * n_in - nh), kind="stable")[:rest]] += 1
    conc = rng.uniform(5.0, 10.0, spec.n_halos)
    centre = rng.uniform(-spec.box / 2, spec.box / 2, size=(spec.n_halos, 3))
    # R_vir from Delta = 200 rho_crit on n_h * m_p with m_p = Om rho_crit L^3 / N
    rvir = (3.0 * nh * spec.omega_m * spec.box ** 3 / (n * 4.0 * np.pi * 200.0)) ** (1.0 / 3.0)
    return nh, conc, centre, rvir


def _nfw_radii(u: np.ndarray, c: np.ndarray) -> np.ndarray:
    """x = r/r_s solving mu(x) = u*mu(c), mu(x) = ln(1+x) - x/(1+x), by bisection."""
    target = u * (np.log1p(c) - c / (1.0 + c))
    lo = np.zeros_like(c)
    hi = c.copy()
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        f = np.log1p(mid) - mid / (1.0 + mid)
        below = f < target
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    return 0.5 * (lo + hi)


def _wrap(v: np.ndarray, box: float) -> np.ndarray:
    return (v + box / 2) % box - box / 2


def particles(spec: SnapshotSpec, n_side: int | None = None, rng=None):
    """Positions (FP64, box-centred), halo id per particle (-1 = background)
    and the halo catalogue."""
    n = (n_side or spec.n_side) ** 3
    rng = rng if rng is not None else np.random.default_rng(spec.seed)
    nh, conc, centre, rvir = _halo_catalogue(spec, rng, n)
    n_in = int(nh.sum())
    hid = np.repeat(np.arange(spec.n_halos), nh)
    xs = _nfw_radii(rng.random(n_in), conc[hid]) * (rvir / conc)[hid]
    cost = rng.uniform(-1.0, 1.0, n_in)
    phi = rng.uniform(0.0, 2 * np.pi, n_in)
    sint = np.sqrt(1.0 - cost ** 2)
    pos = np.empty((n, 3))
    pos[:n_in, 0] = centre[hid, 0] + xs * sint * np.cos(phi)
    pos[:n_in, 1] = centre[hid, 1] + xs * sint * np.sin(phi)
    pos[:n_in, 2] = centre[hid, 2] + xs * cost
    pos[n_in:] = rng.uniform(-spec.box / 2, spec.box / 2, size=(n - n_in, 3))
    pos = _wrap(pos, spec.box)
    halo_of = np.full(n, -1, dtype=np.int64)
    halo_of[:n_in] = hid
    return pos, halo_of, (nh, conc, centre, rvir)


def species_local_density(x, y, z, m, geom: Geometry, rodo, re0) -> np.ndarray:
    """Kernel class per particle by the local-density rule (reader.f:176-233),
    with an NX-sized NGP grid."""
    nx = geom.nx
    dx = geom.dx
    xmin = f32(-f32(nx) * dx / f32(2.0))
    def idx(v):
        i = ((v - xmin) / dx).astype(np.int64) + 1
        return np.clip(i, 1, nx) - 1
    ix, iy, iz = idx(x), idx(y), idx(z)
    flat = ix + nx * (iy + nx * iz)
    dens = np.bincount(flat, weights=m.astype(np.float64), minlength=nx ** 3).astype(np.float32)
    bas = f32(f32(f32(dx * dx) * dx) * f32(rodo) * f32(f32(re0) * f32(f32(re0) * f32(re0))))
    dens = dens / bas
    d = dens[flat]
    with np.errstate(divide="ignore"):
        s = np.where(d > 0, (np.log(np.maximum(d, f32(1e-30))) / np.log(f32(8.0))).astype(np.float32), 0)
    s = np.clip(np.trunc(s).astype(np.int64), 0, N_ESP - 1)
    return s


def hierarchy(spec: SnapshotSpec, geom_base: dict, halos, rng: np.random.Generator):
    """Patch arrays obeying CREATE_MESH's nesting and lattice rules."""
    nh, conc, centre, rvir = halos
    nx = geom_base["nx"]
    radx, dx = geom_base["radx"], geom_base["dx"]
    nl = spec.nl
    p_halo = nh / max(nh.sum(), 1)
    # per level: arrays of (pnx, pny, pnz, i1, j1, k1, prx, pry, prz, parent, halo)
    levels = []
    logs = np.log(spec.side_lo), np.log(spec.side_hi)
    for ir in range(1, nl + 1):
        want = spec.patches_per_level[ir - 1]
        if ir == 1:
            par_n = None
            halo = rng.choice(len(nh), size=want, p=p_halo)
            parent = np.zeros(want, dtype=np.int64)
        else:
            prev = levels[-1]
            if len(prev["halo"]) == 0:
                levels.append({k: v[:0] for k, v in prev.items()})
                continue
            w = nh[prev["halo"]].astype(np.float64)
            w /= w.sum()
            pick = rng.choice(len(prev["halo"]), size=want, p=w)
            halo = prev["halo"][pick]
            parent = pick
        dxp = f32(dx / f32(2.0) ** (ir - 1))      # parent cell size
        sides = np.exp(rng.uniform(logs[0], logs[1], size=(want, 3)))
        sides = np.clip((np.round(sides / 2) * 2).astype(np.int64), spec.side_lo, spec.side_hi)
        # target point: halo centre + offset shrinking with the level
        sig = rvir[halo] * 0.6 / (1.5 ** (ir - 1))
        pt = centre[halo] + rng.normal(size=(want, 3)) * sig[:, None]
        out = {k: np.zeros(want, dtype=np.int64) for k in ("nx", "ny", "nz", "i1", "j1", "k1")}
        rr = np.zeros((want, 3), dtype=np.float32)
        for ax, (nk, ik) in enumerate((("nx", "i1"), ("ny", "j1"), ("nz", "k1"))):
            if ir == 1:
                ext = nx                                   # parent cells available
                lo_edge = f32(radx[1] - dx / f32(2.0))
                cell = np.floor((pt[:, ax] - lo_edge) / dx).astype(np.int64) + 1
                border = 1
            else:
                prev = levels[-1]
                ext = prev["n" + "xyz"[ax]][parent]
                porg = prev["r"][parent, ax] - dxp        # parent left edge
                cell = np.floor((pt[:, ax] - porg) / dxp).astype(np.int64) + 1
                border = 0
            s = np.minimum(sides[:, ax], 2 * (ext - 2 * border))
            s = np.maximum(s - (s % 2), 2)
            npc = s // 2                                   # parent cells covered
            i1 = cell - npc // 2
            i1 = np.clip(i1, 1 + border, ext - border - npc + 1)
            out[nk] = s
            out[ik] = i1
            if ir == 1:
                rr[:, ax] = radx[i1]                       # grids.f:300-302
            else:
                prr = prev["r"][parent, ax].astype(np.float32)
                rr[:, ax] = prr + (i1.astype(np.float32) - f32(1.5)) * dxp  # grids.f:586-588
        levels.append({"nx": out["nx"], "ny": out["ny"], "nz": out["nz"],
                       "r": rr, "parent": parent, "halo": halo})
    npatch = [0] + [len(l["halo"]) for l in levels]
    base = np.cumsum([0] + npatch)              # global 0-based start of level ir at base[ir]
    pnx = np.concatenate([l["nx"] for l in levels]) if levels else np.zeros(0, np.int64)
    pny = np.concatenate([l["ny"] for l in levels]) if levels else np.zeros(0, np.int64)
    pnz = np.concatenate([l["nz"] for l in levels]) if levels else np.zeros(0, np.int64)
    r = np.concatenate([l["r"] for l in levels]) if levels else np.zeros((0, 3), np.float32)
    pare = []
    for ir, l in enumerate(levels, start=1):
        if ir == 1:
            pare.append(np.zeros(len(l["halo"]), np.int64))
        else:
            pare.append(l["parent"] + base[ir - 1] + 1)   # 1-based global parent index
    pare = np.concatenate(pare) if pare else np.zeros(0, np.int64)
    return np.array(npatch, np.int32), pnx, pny, pnz, r[:, 0], r[:, 1], r[:, 2], pare


def make_snapshot(spec: SnapshotSpec, n_side: int | None = None, nx: int | None = None,
                  patches_per_level: tuple | None = None, seed: int | None = None,
                  interp_degree: int = 2, side_hi: int | None = None) -> Snapshot:
    """Build one seeded snapshot.  Optional overrides shrink a workload for tests
    (side_hi = 32 keeps every patch within the reference's NAMRX = 32)."""
    spec = dataclasses.replace(
        spec,
        n_side=n_side or spec.n_side,
        nx=nx or spec.nx,
        patches_per_level=spec.patches_per_level if patches_per_level is None else tuple(patches_per_level),
        seed=spec.seed if seed is None else seed,
        side_hi=side_hi or spec.side_hi)
    rng = np.random.default_rng(spec.seed)
    pos, _, halos = particles(spec, rng=rng)
    n = len(pos)
    rote, rete = background(spec.h, spec.omega_m, 0.0)
    lado0 = f32(spec.box)
    g0 = Geometry(spec.nx, spec.nx, spec.nx, lado0, rote, rete, np.zeros(1, np.int32),
                  np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0))
    npatch, pnx, pny, pnz, prx, pry, prz, pare = hierarchy(
        spec, {"nx": spec.nx, "radx": g0.radx, "dx": g0.dx}, halos, rng)
    geom = Geometry(spec.nx, spec.nx, spec.nx, lado0, rote, rete, npatch, pnx, pny, pnz,
                    prx, pry, prz, pare=pare, interp_degree=interp_degree)
    x = pos[:, 0].astype(np.float32)
    y = pos[:, 1].astype(np.float32)
    z = pos[:, 2].astype(np.float32)
    # keep every position strictly inside [-L/2, L/2) after the FP32 rounding
    half = f32(spec.box / 2)
    for v in (x, y, z):
        v[v >= half] = -half
    mp = f32(f32(rote * f32(rete * f32(rete * rete))) * f32(f32(spec.box) ** 3) / f32(n))
    m = np.full(n, mp, dtype=np.float32)
    s = species_local_density(x, y, z, m, geom, rote, rete)
    order = np.argsort(s, kind="stable")
    npart_esp = np.bincount(s, minlength=N_ESP).astype(np.int32)
    return Snapshot(np.ascontiguousarray(x[order]), np.ascontiguousarray(y[order]),
                    np.ascontiguousarray(z[order]), np.ascontiguousarray(m[order]),
                    npart_esp, geom, spec)
