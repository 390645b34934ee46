"""Seeded SPH scenes: the dam break (one world) and batched independent worlds.

Dam break: a fluid lattice of (nx, ny, nz) particles at spacing s placed at (i + 1/2) s
from the corner of an open box [0, Lx] x [0, Ly] x [0, Lz] (z up), jitter 0.1 s
(uniform +-0.05 s from numpy.random.default_rng(42 + 0), builder.py:8095-8096), mass
rho0 s^3, zero velocity.  Walls and floor are two layers of static boundary particles
(ACTIVE, mass 0) continuing the same lattice outside the box.  Support radius H = 2 s.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .sim import ModelBuilder, ParticleFlags

REST_DENSITY = 1000.0
DT = 1.0e-4


@dataclass
class DamBreak:
    name: str
    spacing: float
    fluid_dims: tuple[int, int, int]  # particles along x, y, z
    box: tuple[float, float, float]
    layers: int = 2
    steps: int = 200

    @property
    def support(self) -> float:
        return 2.0 * self.spacing

    @property
    def n_fluid(self) -> int:
        return int(np.prod(self.fluid_dims))


# Configurations (particle count authoritative; lattice shapes as frozen for this build).
DAM_BREAK_32K = DamBreak("dam_break_32k", 0.0125, (32, 32, 32), (1.0, 1.0, 1.0))
DAM_BREAK_64K = DamBreak("dam_break_64k", 0.0125, (32, 32, 64), (1.0, 1.0, 1.0))
DAM_BREAK_1M = DamBreak("dam_break_1m", 0.004, (64, 128, 128), (2.0, 2.0, 1.0))


def _register_sph(b: ModelBuilder) -> None:
    from .solvers import SolverSPH

    SolverSPH.register_custom_attributes(b)


def small_dam_break(nx: int = 12, ny: int = 10, nz: int = 14, spacing: float = 0.0125,
                    box: tuple[float, float, float] = (0.3, 0.3, 0.3)) -> DamBreak:
    return DamBreak(f"dam_break_{nx}x{ny}x{nz}", spacing, (nx, ny, nz), box)


def box_boundary_points(spacing: float, box: tuple[float, float, float], layers: int = 2,
                        origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Lattice points (i + 1/2) s outside [0, L] on the floor and four walls (no lid),
    `layers` deep, walls up to the box height. float64, (M, 3)."""
    s = spacing
    n = [int(round(L / s)) for L in box]
    ii = np.arange(-layers, n[0] + layers)
    jj = np.arange(-layers, n[1] + layers)
    pts = []
    # floor
    kk = np.arange(-layers, 0)
    g = np.stack(np.meshgrid(ii, jj, kk, indexing="ij"), axis=-1).reshape(-1, 3)
    pts.append(g)
    # walls: ring around [0, n) x [0, n) for 0 <= k < nz
    I, J = np.meshgrid(ii, jj, indexing="ij")
    ring = (I < 0) | (I >= n[0]) | (J < 0) | (J >= n[1])
    ri, rj = I[ring], J[ring]
    kk = np.arange(0, n[2])
    K = np.repeat(kk, ri.size)
    pts.append(np.stack([np.tile(ri, kk.size), np.tile(rj, kk.size), K], axis=-1))
    idx = np.concatenate(pts).astype(np.float64)
    return (idx + 0.5) * s + np.asarray(origin, dtype=np.float64)


def build_dam_break(cfg: DamBreak, builder: ModelBuilder | None = None,
                    rest_density: float = REST_DENSITY) -> ModelBuilder:
    """Add the fluid lattice first (generator seed 42 + 0), then the boundary.  The SPH
    solver's State attribute (sph:pressure) is registered."""
    b = builder if builder is not None else ModelBuilder()
    _register_sph(b)
    s = cfg.spacing
    nx, ny, nz = cfg.fluid_dims
    b.add_particle_grid(pos=(0.5 * s, 0.5 * s, 0.5 * s), rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0),
                        dim_x=nx, dim_y=ny, dim_z=nz, cell_x=s, cell_y=s, cell_z=s,
                        mass=rest_density * s ** 3, jitter=0.1 * s, radius_mean=0.5 * s)
    bp = box_boundary_points(s, cfg.box, cfg.layers)
    b.add_particles(bp, np.zeros_like(bp), np.zeros(bp.shape[0]), np.full(bp.shape[0], 0.5 * s),
                    np.full(bp.shape[0], int(ParticleFlags.ACTIVE)))
    return b


# ---- batched worlds -------------------------------------------------------------------------


@dataclass
class BatchedWorlds:
    name: str
    n_worlds: int
    count_range: tuple[int, int]
    log_uniform: bool
    seed: int = 7
    max_obstacles: int = 3
    layers: int = 2
    spacings: list = field(default_factory=list)


BATCH_4096 = BatchedWorlds("batched_4096", 4096, (256, 8192), True)
BATCH_1024 = BatchedWorlds("batched_1024", 1024, (4096, 65536), False)


def small_batch(n_worlds: int = 12, seed: int = 7) -> BatchedWorlds:
    return BatchedWorlds(f"batched_{n_worlds}", n_worlds, (256, 2048), True, seed=seed)


def build_batched(cfg: BatchedWorlds, rest_density: float = REST_DENSITY):
    """Independent worlds: per world a random water block (extents U(0.2, 0.6) m) with
    about N_w particles (N_w log-uniform or uniform in count_range), in an open box twice
    the block's x extent, with 0..max_obstacles random static boxes (boundary particles,
    two-layer shells) on the floor. Returns (builder, support per world)."""
    rng = np.random.default_rng(cfg.seed)
    lo, hi = cfg.count_range
    b = ModelBuilder()
    _register_sph(b)
    supports = []
    for w in range(cfg.n_worlds):
        if cfg.log_uniform:
            n_target = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        else:
            n_target = float(rng.uniform(lo, hi))
        ext = rng.uniform(0.2, 0.6, size=3)
        s = float((np.prod(ext) / n_target) ** (1.0 / 3.0))
        dims = [max(2, int(round(e / s))) for e in ext]
        box = (2.0 * dims[0] * s, dims[1] * s, 1.5 * dims[2] * s)
        n_cells = [int(round(L / s)) for L in box]
        box = tuple(n_cells[a] * s for a in range(3))
        b.begin_world()
        b.add_particle_grid(pos=(0.5 * s, 0.5 * s, 0.5 * s), rot=(0.0, 0.0, 0.0, 1.0),
                            vel=(0.0, 0.0, 0.0), dim_x=dims[0], dim_y=dims[1], dim_z=dims[2],
                            cell_x=s, cell_y=s, cell_z=s, mass=rest_density * s ** 3,
                            jitter=0.1 * s, radius_mean=0.5 * s)
        pts = [box_boundary_points(s, box, cfg.layers)]
        n_obs = int(rng.integers(0, cfg.max_obstacles + 1))
        for _ in range(n_obs):
            # obstacle: lattice box in the empty half of the container, resting on the floor
            ox = int(rng.integers(dims[0] + 1, max(dims[0] + 2, n_cells[0] - 3)))
            oy = int(rng.integers(0, max(1, n_cells[1] - 3)))
            sx, sy, sz = (int(rng.integers(2, 6)) for _ in range(3))
            ii = np.arange(ox, min(ox + sx, n_cells[0]))
            jj = np.arange(oy, min(oy + sy, n_cells[1]))
            kk = np.arange(0, min(sz, n_cells[2]))
            g = np.stack(np.meshgrid(ii, jj, kk, indexing="ij"), axis=-1).reshape(-1, 3)
            pts.append((g.astype(np.float64) + 0.5) * s)
        bp = np.concatenate(pts)
        bp = np.unique(np.round(bp / s - 0.5).astype(np.int64), axis=0).astype(np.float64)
        bp = (bp + 0.5) * s
        b.add_particles(bp, np.zeros_like(bp), np.zeros(bp.shape[0]), np.full(bp.shape[0], 0.5 * s),
                        np.full(bp.shape[0], int(ParticleFlags.ACTIVE)))
        b.end_world()
        supports.append(2.0 * s)
    return b, np.asarray(supports)


def extract_world(model, w: int) -> ModelBuilder:
    """A one-world builder holding world w's particles of a finalized model (positions as
    stored, float32 -> float64 exactly), with that world's gravity."""
    q = model.particle_q.detach().cpu().numpy()
    qd = model.particle_qd.detach().cpu().numpy()
    mass = model.particle_mass.detach().cpu().numpy()
    flags = model.particle_flags.detach().cpu().numpy()
    radius = model.particle_radius.detach().cpu().numpy()
    world = model.particle_world.detach().cpu().numpy()
    g = model.gravity.detach().cpu().numpy().reshape(-1, 3)[w]
    sel = world == w
    b = ModelBuilder()
    _register_sph(b)
    b.begin_world(gravity=tuple(float(x) for x in g))
    b.add_particles(q[sel].astype(np.float64), qd[sel].astype(np.float64), mass[sel].astype(np.float64),
                    radius[sel].astype(np.float64), flags[sel])
    b.end_world()
    return b
