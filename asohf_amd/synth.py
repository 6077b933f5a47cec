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
I1 with PATCHRX = RADX(I1) (grids.f:300-302); a level-IR patch lies inside one
parent and starts on the parent's cell I1 with
PATCHRX = PATCHRX(parent) + (I1-1.5)*DXPA(parent) (grids.f:586-588).
Placement follows CREATE_MESH's refinement criterion (REFINE_THR = 8 particles
per parent cell, grids.f:107-129, 357-405, i.e. density above 8**IR times the
mean at level IR): each halo's refinable region is the cube out to its NFW
radius of that density, tiled inside each parent by non-overlapping patches
with sides log-uniform in [8, 64] fine cells.  Overlaps only arise where the
regions of neighbouring haloes meet.  The threshold is scaled by one factor
per workload so the total patch count lands near the target.
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
    fdm: float = 0.845          # FDM of RENORM_DENSITY: the reference's default input (asohf.dat:11)

    @property
    def n(self) -> int:
        return self.n_side ** 3

    @property
    def nl(self) -> int:
        return len(self.patches_per_level)


# The workloads of the density build: 1 is the small CPU-runnable case, 2 the
# 256^3-particle single-GPU benchmark, 3 and 5 the large multi-level runs, 4 the
# base grid alone (no refinement).
WORKLOADS = {
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


def _nfw_overdensity_radius(rhos_over_thr: np.ndarray) -> np.ndarray:
    """x = r/r_s where the NFW density rho_s/(x(1+x)^2) falls to the threshold:
    solves x(1+x)^2 = rho_s/threshold by bisection (0 where never reached)."""
    lo = np.zeros_like(rhos_over_thr)
    hi = np.maximum(np.cbrt(rhos_over_thr), rhos_over_thr) + 1.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        f = mid * (1.0 + mid) ** 2
        above = f < rhos_over_thr
        lo = np.where(above, mid, lo)
        hi = np.where(above, hi, mid)
    return 0.5 * (lo + hi)


def _segments(total: int, rng, lo_fine: int, hi_fine: int) -> list:
    """Split `total` parent cells into runs whose fine-cell sides (2 per parent
    cell) are log-uniform in [lo_fine, hi_fine]; a short remainder is merged
    into the previous run (or the two are split evenly) so no run is below
    lo_fine/2 parent cells when total allows it."""
    lo_r, hi_r = max(lo_fine // 2, 1), max(hi_fine // 2, 1)
    out, left = [], total
    while left > 0:
        side = int(np.exp(rng.uniform(np.log(lo_fine), np.log(hi_fine))))
        run = min(max(side // 2, lo_r), left)
        out.append(run)
        left -= run
    if len(out) > 1 and out[-1] < lo_r:
        rest = out.pop()
        if out[-1] + rest <= hi_r:
            out[-1] += rest
        else:
            both = out.pop() + rest
            out += [both // 2, both - both // 2]
    return out


def hierarchy(spec: SnapshotSpec, geom_base: dict, halos, rng: np.random.Generator, n: int,
              alpha: float = 1.0):
    """Patch arrays built the way CREATE_MESH refines: a level-IR patch covers
    parent cells that hold at least REFINE_THR = 8 particles, i.e. where the
    density exceeds 8**IR times the mean (alpha scales that threshold).  For
    every halo the refinable region is the cube around its centre out to the
    NFW radius of that density; inside each parent patch (the base grid at
    level 1) the region is tiled by non-overlapping patches with sides
    log-uniform in [side_lo, side_hi] fine cells."""
    nh, conc, centre, rvir = halos
    nx = geom_base["nx"]
    radx, dx = geom_base["radx"], geom_base["dx"]
    L = spec.box
    rs = rvir / conc
    mu_c = np.log1p(conc) - conc / (1.0 + conc)
    with np.errstate(divide="ignore", invalid="ignore"):
        rhos = np.where(nh > 0, nh * L ** 3 / (n * 4.0 * np.pi * rs ** 3 * mu_c), 0.0)
    left_edge = float(radx[1]) - float(dx) / 2.0         # -L/2 up to FP32 rounding
    # per level: list of dicts (n[3], i1[3] in parent cells, r[3] FP32, parent idx, halo)
    levels = []
    for ir in range(1, spec.nl + 1):
        thr = alpha * 8.0 ** ir
        rad = np.minimum(_nfw_overdensity_radius(rhos / thr) * rs, rvir)
        dxp = f32(dx / f32(2.0) ** (ir - 1))          # parent cell size
        cur = []
        if ir == 1:
            parents = [(-1, h) for h in np.nonzero(2 * rad / (dx / 2.0) >= 4)[0]]
        else:
            parents = [(k, p["halo"]) for k, p in enumerate(levels[-1])]
        for pk, h in parents:
            c, r = centre[h], rad[h]
            if 2 * r / (float(dxp) / 2.0) < 4:
                continue
            box = []
            for a in range(3):
                if pk < 0:
                    org, ext = left_edge, nx                  # base cells 1..nx
                    lo_cell = max(int(np.floor((c[a] - r - org) / float(dxp))) + 1, 2)
                    hi_cell = min(int(np.floor((c[a] + r - org) / float(dxp))) + 1, ext - 1)
                else:
                    par = levels[-1][pk]
                    org = float(par["r"][a]) - float(dxp)     # parent's left edge (its XLFIX)
                    ext = par["n"][a]
                    lo_cell = max(int(np.floor((c[a] - r - org) / float(dxp))) + 1, 1)
                    hi_cell = min(int(np.floor((c[a] + r - org) / float(dxp))) + 1, ext)
                if hi_cell < lo_cell:
                    box = None
                    break
                # CREATE_MESH's minimum patch size: at least side_lo fine cells
                minrun = spec.side_lo // 2
                first = 2 if pk < 0 else 1
                if hi_cell - lo_cell + 1 < minrun and ext - first + 1 >= minrun:
                    need = minrun - (hi_cell - lo_cell + 1)
                    lo_cell -= need // 2
                    hi_cell += need - need // 2
                    if lo_cell < first:
                        hi_cell += first - lo_cell
                        lo_cell = first
                    last = ext - 1 if pk < 0 else ext
                    if hi_cell > last:
                        lo_cell -= hi_cell - last
                        hi_cell = last
                box.append((lo_cell, hi_cell))
            if box is None:
                continue
            segs = [_segments(b1 - b0 + 1, rng, spec.side_lo, spec.side_hi) for b0, b1 in box]
            starts = [np.concatenate([[b0], b0 + np.cumsum(sg)[:-1]]) for (b0, _), sg in zip(box, segs)]
            for kz, sz in enumerate(segs[2]):
                for ky, sy in enumerate(segs[1]):
                    for kx, sx in enumerate(segs[0]):
                        i1 = (int(starts[0][kx]), int(starts[1][ky]), int(starts[2][kz]))
                        nn = (2 * sx, 2 * sy, 2 * sz)
                        rr = np.zeros(3, np.float32)
                        for a in range(3):
                            if pk < 0:
                                rr[a] = radx[i1[a]]                                     # grids.f:300-302
                            else:
                                rr[a] = f32(levels[-1][pk]["r"][a] +
                                            (f32(i1[a]) - f32(1.5)) * dxp)               # grids.f:586-588
                        cur.append({"n": nn, "r": rr, "parent": pk, "halo": h, "i1": i1})
        levels.append(cur)
    npatch = [0] + [len(l) for l in levels]
    base = np.cumsum([0] + npatch)
    flat = [p for l in levels for p in l]
    pn = np.array([p["n"] for p in flat], np.int64).reshape(-1, 3)
    pr = np.array([p["r"] for p in flat], np.float32).reshape(-1, 3)
    pc = np.array([p["i1"] for p in flat], np.int64).reshape(-1, 3)   # PATCHX/Y/Z: parent cell
    pare = []
    for ir, l in enumerate(levels, start=1):
        for p in l:
            pare.append(0 if ir == 1 else p["parent"] + base[ir - 1] + 1)   # 1-based global parent
    return (np.array(npatch, np.int32), pn[:, 0], pn[:, 1], pn[:, 2], pr[:, 0], pr[:, 1], pr[:, 2],
            np.array(pare, np.int64), pc)


def _tuned_hierarchy(spec, geom_base, halos, seed, n):
    """Scales the refinement threshold so the total patch count lands near the
    workload's target (sum of patches_per_level)."""
    target = sum(spec.patches_per_level)
    lo, hi = -6.0, 6.0          # log2 of the threshold factor
    best = None
    for _ in range(12):
        mid = 0.5 * (lo + hi)
        h = hierarchy(spec, geom_base, halos, np.random.default_rng(seed + 1000), n, alpha=2.0 ** mid)
        cnt = int(h[0].sum())
        if best is None or abs(cnt - target) < abs(int(best[0].sum()) - target):
            best = h
        if abs(cnt - target) <= 0.05 * target:
            break
        if cnt > target:
            lo = mid
        else:
            hi = mid
    return best


def make_snapshot(spec: SnapshotSpec, n_side: int | None = None, nx: int | None = None,
                  patches_per_level: tuple | None = None, seed: int | None = None,
                  interp_degree: int = 2, side_hi: int | None = None, fdm: float | None = None) -> Snapshot:
    """Build one seeded snapshot.  Optional overrides shrink a workload for tests
    (side_hi = 32 keeps every patch within the reference's NAMRX = 32)."""
    spec = dataclasses.replace(
        spec,
        n_side=n_side or spec.n_side,
        nx=nx or spec.nx,
        patches_per_level=spec.patches_per_level if patches_per_level is None else tuple(patches_per_level),
        seed=spec.seed if seed is None else seed,
        side_hi=side_hi or spec.side_hi,
        fdm=spec.fdm if fdm is None else fdm)
    rng = np.random.default_rng(spec.seed)
    pos, _, halos = particles(spec, rng=rng)
    n = len(pos)
    rote, rete = background(spec.h, spec.omega_m, 0.0)
    lado0 = f32(spec.box)
    g0 = Geometry(spec.nx, spec.nx, spec.nx, lado0, rote, rete, np.zeros(1, np.int32),
                  np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0))
    npatch, pnx, pny, pnz, prx, pry, prz, pare, pc = _tuned_hierarchy(
        spec, {"nx": spec.nx, "radx": g0.radx, "dx": g0.dx}, halos, spec.seed, n)
    geom = Geometry(spec.nx, spec.nx, spec.nx, lado0, rote, rete, npatch, pnx, pny, pnz,
                    prx, pry, prz, pare=pare, fdm=f32(spec.fdm), interp_degree=interp_degree,
                    pcx=pc[:, 0], pcy=pc[:, 1], pcz=pc[:, 2])
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


def make_snapshot_device(spec: SnapshotSpec, device="cuda:0", chunk: int = 1 << 26,
                         interp_degree: int = 2) -> Snapshot:
    """The same workload generated on the GPU (for the 134M- and 1.07G-particle
    configurations, which numpy cannot produce in reasonable time or memory).
    The halo catalogue and the AMR hierarchy come from the same host RNG stream
    as make_snapshot; the particles from torch's device RNG seeded with the
    workload seed, so they differ from make_snapshot's for the same spec.
    Returns a Snapshot whose x, y, z, m are device tensors in reference order
    (species contiguous) and whose npart_esp is a host array."""
    import torch

    rng = np.random.default_rng(spec.seed)
    n = spec.n
    nh, conc, centre, rvir = _halo_catalogue(spec, rng, n)
    rote, rete = background(spec.h, spec.omega_m, 0.0)
    lado0 = f32(spec.box)
    g0 = Geometry(spec.nx, spec.nx, spec.nx, lado0, rote, rete, np.zeros(1, np.int32),
                  np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0))
    npatch, pnx, pny, pnz, prx, pry, prz, pare, pc = _tuned_hierarchy(
        spec, {"nx": spec.nx, "radx": g0.radx, "dx": g0.dx}, (nh, conc, centre, rvir), spec.seed, n)
    geom = Geometry(spec.nx, spec.nx, spec.nx, lado0, rote, rete, npatch, pnx, pny, pnz,
                    prx, pry, prz, pare=pare, fdm=f32(spec.fdm), interp_degree=interp_degree,
                    pcx=pc[:, 0], pcy=pc[:, 1], pcz=pc[:, 2])

    dev = torch.device(device)
    gen = torch.Generator(device=dev)
    gen.manual_seed(spec.seed)
    L = float(spec.box)
    pos = [torch.empty(n, dtype=torch.float32, device=dev) for _ in range(3)]
    n_in = int(nh.sum())
    conc_t = torch.from_numpy(conc).to(dev, torch.float32)
    rs_t = torch.from_numpy(rvir / conc).to(dev, torch.float32)
    cen_t = torch.from_numpy(centre).to(dev, torch.float32)
    starts = np.concatenate([[0], np.cumsum(nh)])
    for s0 in range(0, n_in, chunk):
        s1 = min(s0 + chunk, n_in)
        idx = torch.arange(s0, s1, device=dev)
        hid = torch.searchsorted(torch.from_numpy(starts[1:]).to(dev), idx, right=True)
        c = conc_t[hid]
        target = torch.rand(s1 - s0, generator=gen, device=dev) * (torch.log1p(c) - c / (1 + c))
        lo, hi = torch.zeros_like(c), c.clone()
        for _ in range(32):                       # NFW radii by bisection
            mid = 0.5 * (lo + hi)
            below = (torch.log1p(mid) - mid / (1 + mid)) < target
            lo = torch.where(below, mid, lo)
            hi = torch.where(below, hi, mid)
        r = 0.5 * (lo + hi) * rs_t[hid]
        cost = torch.rand(s1 - s0, generator=gen, device=dev) * 2 - 1
        phi = torch.rand(s1 - s0, generator=gen, device=dev) * (2 * np.pi)
        sint = torch.sqrt(torch.clamp(1 - cost * cost, min=0))
        comp = (sint * torch.cos(phi), sint * torch.sin(phi), cost)
        for a in range(3):
            v = cen_t[hid, a] + r * comp[a]
            pos[a][s0:s1] = torch.remainder(v + L / 2, L) - L / 2
        del idx, hid, c, target, lo, hi, r, cost, phi, sint, comp
    for s0 in range(n_in, n, chunk):
        s1 = min(s0 + chunk, n)
        for a in range(3):
            pos[a][s0:s1] = torch.rand(s1 - s0, generator=gen, device=dev) * L - L / 2
    half = float(f32(spec.box / 2))
    for a in range(3):
        pos[a][pos[a] >= half] = -half
    mp = f32(f32(rote * f32(rete * f32(rete * rete))) * f32(f32(spec.box) ** 3) / f32(n))
    # species by the local-density rule on an NX-sized NGP grid (equal masses:
    # the density is count * m / (DX**3 * RODO * RE0**3))
    nx = spec.nx
    dx = float(geom.dx)
    xmin = float(f32(-f32(nx) * geom.dx / f32(2.0)))
    flat = torch.zeros(n, dtype=torch.int64, device=dev)
    for a in (2, 1, 0):
        i = torch.clamp(((pos[a] - xmin) / dx).to(torch.int64), 0, nx - 1)
        flat = flat * nx + i
        del i
    cnt = torch.bincount(flat, minlength=nx ** 3)
    bas = float(f32(f32(f32(geom.dx * geom.dx) * geom.dx) * rote * f32(rete * f32(rete * rete))))
    dens = cnt[flat].to(torch.float32) * float(mp) / bas
    del cnt, flat
    s = torch.where(dens > 0, torch.log(torch.clamp(dens, min=1e-30)) / np.log(8.0),
                    torch.zeros_like(dens))
    s = torch.clamp(torch.trunc(s), 0, N_ESP - 1).to(torch.int32)
    del dens
    npart_esp = torch.bincount(s, minlength=N_ESP).cpu().numpy().astype(np.int32)
    order = torch.sort(s, stable=True).indices
    del s
    out = [p[order].contiguous() for p in pos]
    del pos, order
    m = torch.full((n,), float(mp), dtype=torch.float32, device=dev)
    if dev.type == "cuda":
        torch.cuda.synchronize(dev)
        torch.cuda.empty_cache()
    return Snapshot(out[0], out[1], out[2], m, npart_esp, geom, spec)
