"""Synthetic ice scenes in the hssoa layout (seeded, reproducible).

Point initialisation follows SnapshotManager::PopulatePoints
(simulation/data_manager/snapshotmanager.cpp:150-270): points in the two-node
border band are dropped (:191-200), ParticleVolume = h^2 gx gy / N_generated
(:180), Fe = I, Jp_inv = 1, v = 0, P = Q = 0 (:241-255), and positions are
converted to the (nearest node, cell-local offset) format of
ProxyPoint::ConvertToIntegerCellFormat (proxypoint2d.cpp:125-139).  Points are
generated sorted by cell index, as generate_and_save sorts them
(snapshotmanager.cpp:515-529).  Instead of Poisson-disk sampling, 4 points
per cell are placed stratified (one per quarter cell) with seeded uniform
jitter.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import abi
from .params import SimParams, sea_ice_params


@dataclass
class Scene:
    params: SimParams
    soa: np.ndarray          # float32 [nlines, n] (integer lines as bit patterns)
    status: np.ndarray       # uint8 [gx*gy], node index x*gy + y
    current: np.ndarray | None   # float32 [2, gx*gy] or None
    wind: np.ndarray | None      # float32 [2, gx*gy] (Pa) or None
    name: str = ""

    @property
    def n(self) -> int:
        return self.soa.shape[1]

    @property
    def nlines(self) -> int:
        return self.soa.shape[0]

    def copy(self) -> "Scene":
        return Scene(self.params.replace(), self.soa.copy(), self.status.copy(),
                     None if self.current is None else self.current.copy(),
                     None if self.wind is None else self.wind.copy(), self.name)


def u32_line(soa: np.ndarray, line: int) -> np.ndarray:
    return soa[line].view(np.uint32)


def border_status(gx: int, gy: int, ring: int = 2) -> np.ndarray:
    """Status 100 (water) inside, wall region ids on a `ring`-node border:
    1 left, 2 right, 3 bottom, 4 top (corners take the x side)."""
    st = np.full((gx, gy), abi.GRID_STATUS_WATER, np.uint8)
    st[:, :ring] = 3
    st[:, gy - ring:] = 4
    st[:ring, :] = 1
    st[gx - ring:, :] = 2
    return st.reshape(-1)


def stratified_positions(gx: int, gy: int, rng: np.random.Generator, ppc_side: int = 2,
                         cell_mask: np.ndarray | None = None):
    """Positions (in cell units) of ppc_side^2 jittered points per square cell
    [i, i+1] x [j, j+1], i < gx-1, j < gy-1.  Returns (x, y, n_generated)
    where n_generated counts every cell before masking."""
    ncx, ncy = gx - 1, gy - 1
    n_gen = ncx * ncy * ppc_side * ppc_side
    if cell_mask is None:
        ii, jj = np.meshgrid(np.arange(ncx, dtype=np.int32), np.arange(ncy, dtype=np.int32), indexing="ij")
        ii, jj = ii.reshape(-1), jj.reshape(-1)
    else:
        ii, jj = np.nonzero(cell_mask.reshape(ncx, ncy))
        ii, jj = ii.astype(np.int32), jj.astype(np.int32)
    k = ppc_side * ppc_side
    sub = np.arange(k, dtype=np.int32)
    sa, sb = sub // ppc_side, sub % ppc_side
    x = (ii[:, None] + (sa[None, :] + rng.random((ii.size, k))) / ppc_side).reshape(-1)
    y = (jj[:, None] + (sb[None, :] + rng.random((ii.size, k))) / ppc_side).reshape(-1)
    return x, y, n_gen


def make_points(x: np.ndarray, y: np.ndarray, gx: int, gy: int, *, with_cohesion: bool = False) -> np.ndarray:
    """Initial state of points at cell-unit positions (x, y), border band
    dropped, sorted by cell index (stable)."""
    xi = np.floor(x + 0.5).astype(np.int64)
    yi = np.floor(y + 0.5).astype(np.int64)
    keep = (xi > 1) & (yi > 1) & (xi < gx - 2) & (yi < gy - 2)
    x, y, xi, yi = x[keep], y[keep], xi[keep], yi[keep]
    key = xi * gy + yi
    order = np.argsort(key, kind="stable")
    x, y, xi, yi = x[order], y[order], xi[order], yi[order]
    n = x.size
    nl = abi.N_LINES_MAX if with_cohesion else abi.N_LINES
    soa = np.zeros((nl, n), np.float32)
    u32_line(soa, abi.IDX_UTILITY)[:] = 0
    u32_line(soa, abi.IDX_CELL)[:] = ((yi.astype(np.uint32) << 16) | xi.astype(np.uint32))
    u32_line(soa, abi.IDX_POINT)[:] = np.arange(n, dtype=np.uint32)
    soa[abi.IDX_JP_INV] = 1.0
    soa[abi.IDX_POSX] = (x - xi).astype(np.float32)
    soa[abi.IDX_POSX + 1] = (y - yi).astype(np.float32)
    soa[abi.IDX_FE00] = 1.0
    soa[abi.IDX_FE00 + 3] = 1.0
    soa[abi.IDX_THICKNESS] = 1.0
    return soa


def gaussian_random_field(gx: int, gy: int, corr_len: float, rng: np.random.Generator) -> np.ndarray:
    """Zero-mean, unit-std stationary field with Gaussian correlation of length
    corr_len (cells), shape [gx, gy], by spectral filtering of white noise."""
    noise = rng.standard_normal((gx, gy))
    kx = np.fft.fftfreq(gx)[:, None]
    ky = np.fft.rfftfreq(gy)[None, :]
    filt = np.exp(-2.0 * (np.pi * corr_len) ** 2 * (kx ** 2 + ky ** 2) / 2.0)
    f = np.fft.irfft2(np.fft.rfft2(noise) * filt, s=(gx, gy))
    f -= f.mean()
    s = f.std()
    return f / s if s > 0 else f


def floe_256x128(seed: int = 12345, gx: int = 256, gy: int = 128) -> Scene:
    """Plain MPM, 256x128 nodes h = 1 m, ice in the left 60 % of the domain,
    4 points per cell, uniform wind stress 0.1 Pa along +x, still ocean,
    dt = 1e-4 s, walls on all sides."""
    rng = np.random.default_rng(seed)
    x, y, n_gen = stratified_positions(gx, gy, rng)
    ice = x < 0.6 * (gx - 1)
    soa = make_points(x[ice], y[ice], gx, gy)
    prm = sea_ice_params(gx, gy)
    prm.ParticleVolume = prm.cellsize ** 2 * gx * gy / n_gen
    wind = np.zeros((2, gx * gy), np.float32)
    wind[0] = 0.1
    return Scene(prm, soa, border_status(gx, gy), None, wind, "floe_256x128")


def voronoi_ice_mask(ncx: int, ncy: int, n_floes: int, coverage: float, rng: np.random.Generator) -> np.ndarray:
    """Cell mask of a random polygonal floe field: Voronoi cells of n_floes
    seeds, with open-water leads along the cell edges chosen so that
    `coverage` of the cells are ice."""
    from scipy.spatial import cKDTree
    seeds = np.column_stack([rng.random(n_floes) * ncx, rng.random(n_floes) * ncy])
    tree = cKDTree(seeds)
    ii, jj = np.meshgrid(np.arange(ncx) + 0.5, np.arange(ncy) + 0.5, indexing="ij")
    pts = np.column_stack([ii.reshape(-1), jj.reshape(-1)])
    d, _ = tree.query(pts, k=2, workers=-1)
    gap = (d[:, 1] - d[:, 0]).astype(np.float32)
    thr = np.quantile(gap, 1.0 - coverage)
    return (gap > thr).reshape(ncx, ncy)


def icefield_4096x2048(seed: int = 2024, gx: int = 4096, gy: int = 2048, cohesion_cap: float = 1e6) -> Scene:
    """Plain MPM at scale: 4096x2048 nodes, 4 points per cell in a seeded
    random polygonal ice field covering ~95 % of the domain, wind stress from
    a Gaussian random field (correlation length 200 cells, mean 0.2 Pa, std
    0.05 Pa along x; mean 0, std 0.05 Pa along y), per-point cohesion
    lognormal (median 1e5 Pa, sigma 0.5, capped at cohesion_cap so the DP apex
    stays above the tensile strength)."""
    rng = np.random.default_rng(seed)
    ncx, ncy = gx - 1, gy - 1
    n_floes = max(8, (ncx * ncy) // (120 * 120))
    mask = voronoi_ice_mask(ncx, ncy, n_floes, 0.95, rng)
    x, y, n_gen = stratified_positions(gx, gy, rng, cell_mask=mask)
    soa = make_points(x, y, gx, gy, with_cohesion=True)
    del x, y
    coh = np.exp(np.log(1e5) + 0.5 * rng.standard_normal(soa.shape[1]))
    soa[abi.IDX_COHESION] = np.minimum(coh, cohesion_cap).astype(np.float32)
    prm = sea_ice_params(gx, gy, flags=abi.FLAG_COHESION_LINE)
    prm.ParticleVolume = prm.cellsize ** 2 * gx * gy / n_gen
    wind = np.empty((2, gx * gy), np.float32)
    wind[0] = (0.2 + 0.05 * gaussian_random_field(gx, gy, 200.0, rng)).reshape(-1)
    wind[1] = (0.05 * gaussian_random_field(gx, gy, 200.0, rng)).reshape(-1)
    return Scene(prm, soa, border_status(gx, gy), None, wind, "icefield_4096x2048")


def dynamic_scene(seed: int = 7, gx: int = 48, gy: int = 40, *, speed: float = 300.0, strain: float = 4e-3,
                  crushed_frac: float = 0.5, with_cohesion: bool = False, with_current: bool = True,
                  fill: float = 0.8, cellsize: float = 1.0, bp_scale: float = 5.0) -> Scene:
    """A small, strongly perturbed scene for parity tests: a swirling velocity
    field of up to `speed` cells/s (speed * cellsize m/s; cells cross every
    ~10-100 substeps at dt = 1e-4 with the default speed), random elastic
    strain up to `strain`, APIC matrices Bp of `bp_scale` 1/s, a fraction of
    points already crushed, random Jp_inv and thickness, a random current and
    wind, on a grid of cell size `cellsize` m.  Exercises every branch of the
    update and return map, cell crossings, the disable band and the wall force
    accumulator."""
    rng = np.random.default_rng(seed)
    x, y, n_gen = stratified_positions(gx, gy, rng)
    cx, cy = (gx - 1) / 2, (gy - 1) / 2
    r = np.maximum(np.abs(x - cx) / cx, np.abs(y - cy) / cy)
    keep = r < fill
    soa = make_points(x[keep], y[keep], gx, gy, with_cohesion=with_cohesion)
    n = soa.shape[1]
    xi = (u32_line(soa, abi.IDX_CELL) & 0xFFFF).astype(np.float64) + soa[abi.IDX_POSX]
    yi = (u32_line(soa, abi.IDX_CELL) >> 16).astype(np.float64) + soa[abi.IDX_POSX + 1]
    ang = np.arctan2(yi - cy, xi - cx)
    rad = np.hypot((xi - cx) / cx, (yi - cy) / cy)
    vs = speed * cellsize
    soa[abi.IDX_VELX] = (-vs * rad * np.sin(ang) + 0.2 * vs * rng.standard_normal(n)).astype(np.float32)
    soa[abi.IDX_VELX + 1] = (vs * rad * np.cos(ang) + 0.2 * vs * rng.standard_normal(n)).astype(np.float32)
    F = np.eye(2)[None] + strain * rng.standard_normal((n, 2, 2))
    soa[abi.IDX_FE00:abi.IDX_FE00 + 4] = F.reshape(n, 4).T.astype(np.float32)
    soa[abi.IDX_BP00:abi.IDX_BP00 + 4] = (bp_scale * rng.standard_normal((4, n))).astype(np.float32)
    soa[abi.IDX_JP_INV] = rng.uniform(0.05, 1.3, n).astype(np.float32)
    soa[abi.IDX_THICKNESS] = rng.uniform(0.5, 2.0, n).astype(np.float32)
    util = u32_line(soa, abi.IDX_UTILITY)
    util[:] = np.where(rng.random(n) < crushed_frac, abi.STATUS_CRUSHED, 0).astype(np.uint32)
    util[:] |= rng.integers(0, 0x10000, n).astype(np.uint32)   # grain bits are carried through
    flags = abi.FLAG_COHESION_LINE if with_cohesion else 0
    if with_cohesion:
        soa[abi.IDX_COHESION] = np.exp(np.log(1e5) + 0.5 * rng.standard_normal(n)).astype(np.float32)
    prm = sea_ice_params(gx, gy, frame_steps=10, flags=flags, cellsize=cellsize)
    prm.ParticleVolume = prm.cellsize ** 2 * gx * gy / n_gen
    N = gx * gy
    current = None
    if with_current:
        current = (0.5 * rng.standard_normal((2, N))).astype(np.float32)
    wind = (0.2 + 0.3 * rng.standard_normal((2, N))).astype(np.float32)
    status = border_status(gx, gy)
    return Scene(prm, soa, status, current, wind, "dynamic")


def batch_domain(seed: int, side_cells: int, coverage: float, wind_pa: float, wind_dir: float,
                 base_params=None) -> Scene:
    """One domain of the batched workload: side_cells x side_cells cells,
    ice in a seeded Voronoi floe field covering `coverage`, uniform wind stress
    of `wind_pa` Pa toward `wind_dir` radians, 4 points per cell.  All domains
    share ParticleVolume = h^2/4 so they can share one material block."""
    rng = np.random.default_rng(seed)
    gx = gy = side_cells + 1
    ncx, ncy = gx - 1, gy - 1
    n_floes = max(4, (ncx * ncy) // (60 * 60))
    mask = voronoi_ice_mask(ncx, ncy, n_floes, coverage, rng)
    x, y, _ = stratified_positions(gx, gy, rng, cell_mask=mask)
    soa = make_points(x, y, gx, gy)
    prm = sea_ice_params(gx, gy) if base_params is None else base_params.replace(GridXTotal=gx, GridYTotal=gy)
    prm.ParticleVolume = prm.cellsize ** 2 / 4.0
    N = gx * gy
    wind = np.empty((2, N), np.float32)
    wind[0] = wind_pa * np.cos(wind_dir)
    wind[1] = wind_pa * np.sin(wind_dir)
    return Scene(prm, soa, border_status(gx, gy), None, wind, f"batch{seed}")


def batch64_uneven(seed: int = 3003, n_domains: int = 64, min_side: int = 128, max_side: int = 1024) -> list:
    """Batched plain MPM: n_domains independent domains, sides log-uniform in
    [min_side, max_side] cells, ice coverage 25-95 % (so 16k-4M points at 4 per
    cell), per-domain seeded wind 0.05-0.3 Pa in a random direction."""
    rng = np.random.default_rng(seed)
    sides = np.exp(rng.uniform(np.log(min_side), np.log(max_side), n_domains)).astype(int)
    out = []
    for d in range(n_domains):
        out.append(batch_domain(seed * 1000 + d, int(sides[d]), float(rng.uniform(0.25, 0.95)),
                                float(rng.uniform(0.05, 0.3)), float(rng.uniform(0, 2 * np.pi))))
    return out


def batch64_compressed(seed: int = 5005, n_domains: int = 64, min_side: int = 128, max_side: int = 1024,
                       strain_rate: float = 1.0, wind_pa: float = 1.0) -> list:
    """Surrogate + batched workload: the batch64_uneven domain set under
    high-strain-rate forcing.  Each domain's ice starts in uniaxial
    compression against the fixed wall at x_max (the first wall node column,
    x_w = gx - 2): v_x = strain_rate * (x_w - x) (m/s, with Bp = grad v),
    under a 1.0 Pa wind stress along +x, so that most points project
    plastically onto the ridge-capped yield surface every substep (the trial
    strain increment dt * strain_rate exceeds the yield strain)."""
    doms = batch64_uneven(seed=seed, n_domains=n_domains, min_side=min_side, max_side=max_side)
    return [compressed(sc, strain_rate, wind_pa) for sc in doms]


def compressed(sc: Scene, strain_rate: float = 1.0, wind_pa: float = 1.0, band: float | None = None) -> Scene:
    """The scene's ice in uniaxial compression against the fixed wall at x_max
    (the first wall node column, x_w = gx - 2): v_x = strain_rate (x_w - x)
    (m/s, Bp = grad v), v_y = 0, under a uniform wind stress of wind_pa along
    +x (replacing the scene's forcing), so that points project plastically
    onto the ridge-capped yield surface every substep.  With `band` (cells),
    the compression acts within that distance of the wall and the ice beyond
    moves at the band's edge speed, strain_rate * band * h (Bp = 0 there), so
    that no velocity reaches the node clamp (sqrt(0.1) h / (4 dt))."""
    soa = sc.soa
    gx, gy = sc.params.GridXTotal, sc.params.GridYTotal
    xi = (u32_line(soa, abi.IDX_CELL) & 0xFFFF).astype(np.float64) + soa[abi.IDX_POSX]
    x_wall = gx - 2.0
    dist = x_wall - xi
    inside = np.ones(dist.shape, bool) if band is None else dist < band
    if band is not None:
        dist = np.minimum(dist, band)
    soa[abi.IDX_VELX] = (strain_rate * dist * sc.params.cellsize).astype(np.float32)
    soa[abi.IDX_VELX + 1] = 0.0
    soa[abi.IDX_BP00] = np.where(inside, -strain_rate, 0.0).astype(np.float32)
    soa[abi.IDX_BP00 + 1:abi.IDX_BP00 + 4] = 0.0
    wind = np.zeros((2, gx * gy), np.float32)
    wind[0] = wind_pa
    return Scene(sc.params, soa, sc.status, None, wind, sc.name + "c")
