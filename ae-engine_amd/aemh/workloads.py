"""Synthetic workloads of the benchmark configurations.

Every scene is generated from a fixed seed (xoshiro256** seeded through
splitmix64(42 + index), 42 being the engine's world seed) with v = 0, C = 0,
F = I, J = 1, mass = rho * V_region / N.  Values the configuration text does
not fix (snow Young's modulus and density, jelly density, box placement) are
chosen so every scene satisfies CFL <= 0.5 at its dt; they are listed in the
docstrings.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import (ACCUM_F32, ACCUM_Q32_32, CLAMP_AUTO, DRUCKER_PRAGER, FX_ELASTOPLASTIC,
               NEO_HOOKEAN, NEWTONIAN_FLUID, SNOW, STORE_F32, Material, Scene, default_config,
               fixed_config_default, uniform_box)

SEED_BASE = 42


@dataclass
class Workload:
    name: str
    pos: np.ndarray
    mass: np.ndarray
    mat: np.ndarray
    materials: list
    grid: tuple            # (nx, ny, nz, dx, origin)
    dt: float
    steps: int
    gravity: tuple = (0.0, -9.8, 0.0)
    bc_margin: int = 3
    accum: int = ACCUM_F32
    storage: int = STORE_F32
    scenes: list = field(default_factory=list)   # batched scenes (aemh Scene)

    @property
    def n(self):
        return self.pos.shape[0]

    def config(self):
        cfg = default_config()
        cfg.substeps = 1
        cfg.gravity[:] = self.gravity
        cfg.damping = 0.995
        cfg.bc_margin = self.bc_margin
        cfg.bc_restitution = 0.0
        cfg.accum = self.accum
        cfg.clamp_mode = CLAMP_AUTO
        cfg.storage = self.storage
        return cfg


def snow_material():
    """Snow of config 2: theta_c 2.5e-2, theta_s 7.5e-3, xi 10; E 1.4e5,
    nu 0.2, rho 400 (Stomakhin et al. 2013 values)."""
    return Material(young_modulus=1.4e5, poisson_ratio=0.2, density=400.0, model=SNOW,
                    theta_c=2.5e-2, theta_s=7.5e-3, xi=10.0)


def c1_elastic_cube(n=8192, steps=100):
    """8,192 particles in [0.3,0.6]^3, 32^3 grid, dx 1/32, dt 1e-4,
    Neo-Hookean E 1e4, nu 0.2, rho 1, gravity -9.8, 3-cell walls."""
    lo, hi = (0.3, 0.3, 0.3), (0.6, 0.6, 0.6)
    pos = uniform_box(SEED_BASE + 0, n, lo, hi)
    vol = 0.3 ** 3
    mass = np.full(n, 1.0 * vol / n, np.float32)
    mats = [Material(young_modulus=1e4, poisson_ratio=0.2, density=1.0, model=NEO_HOOKEAN)]
    return Workload("c1_elastic_8k_32", pos, mass, np.zeros(n, np.uint16), mats,
                    (32, 32, 32, 1.0 / 32.0, (0.0, 0.0, 0.0)), 1e-4, steps)


def c2_snow_dam(n=1 << 20, steps=200, accum=ACCUM_F32, res=128):
    """1,048,576 snow particles in [0,0.4]x[0,0.6]x[0,1], 128^3 grid,
    dx 1/128, dt 5e-5 (dam break)."""
    lo, hi = (0.0, 0.0, 0.0), (0.4, 0.6, 1.0)
    pos = uniform_box(SEED_BASE + 1, n, lo, hi)
    rho = 400.0
    mass = np.full(n, rho * 0.4 * 0.6 * 1.0 / n, np.float32)
    return Workload("c2_snow_1m_128", pos, mass, np.zeros(n, np.uint16), [snow_material()],
                    (res, res, res, 1.0 / res, (0.0, 0.0, 0.0)), 5e-5, steps, accum=accum)


def c3_three_boxes(n=1 << 24, steps=100, res=256):
    """16,777,216 particles in three seeded boxes falling onto the floor of a
    256^3 grid of 4^3 blocks.  Box k holds material k: weakly compressible
    fluid (K 400, rho 1), elastic jelly (E 1e4, nu 0.2, rho 10) and snow.
    Box sides U(0.30, 0.45), centres x,z U(0.3, 0.7), y U(0.45, 0.7)."""
    rng = np.random.default_rng(SEED_BASE + 2)
    counts = [n // 3, n // 3, n - 2 * (n // 3)]
    pos_l, mass_l, mat_l = [], [], []
    rhos = [1.0, 10.0, 400.0]
    for k in range(3):
        side = rng.uniform(0.30, 0.45)
        cx, cz = rng.uniform(0.3, 0.7, 2)
        cy = rng.uniform(0.45, 0.7)
        lo = (cx - side / 2, cy - side / 2, cz - side / 2)
        hi = (cx + side / 2, cy + side / 2, cz + side / 2)
        p = uniform_box(SEED_BASE + 2 + 1000 * (k + 1), counts[k], lo, hi)
        pos_l.append(p)
        mass_l.append(np.full(counts[k], rhos[k] * side ** 3 / counts[k], np.float32))
        mat_l.append(np.full(counts[k], k, np.uint16))
    # weakly compressible fluid: K = E / (3 (1 - 2 nu)) = 400 with nu = 0.2
    fluid = Material(young_modulus=400.0 * 3.0 * (1.0 - 0.4), poisson_ratio=0.2, density=1.0,
                     model=NEWTONIAN_FLUID)
    jelly = Material(young_modulus=1e4, poisson_ratio=0.2, density=10.0, model=NEO_HOOKEAN)
    mats = [fluid, jelly, snow_material()]
    return Workload("c3_three_boxes_16m_256", np.concatenate(pos_l), np.concatenate(mass_l),
                    np.concatenate(mat_l), mats, (res, res, res, 1.0 / res, (0.0, 0.0, 0.0)),
                    5e-5, steps)


def c4_batched_scenes(n_scenes=256, steps=100, n_min=4096, n_max=262144):
    """256 independent scenes, N_i log-uniform in [4K, 256K], grid
    n_i = 4 ceil(cbrt(8 N_i) / 4) clamped to [32, 128], dx = 1/n_i, particles
    in a cube of side 0.25 centred (0.5, 0.6, 0.5), scene material alternating
    Neo-Hookean (E 1e4, nu 0.2, rho 10) and snow; dt 5e-5."""
    rng = np.random.default_rng(SEED_BASE + 3)
    sizes = np.round(np.exp(rng.uniform(math.log(n_min), math.log(n_max), n_scenes))).astype(np.int64)
    pos_l, mass_l, mat_l, scenes = [], [], [], []
    first = 0
    for s, ns in enumerate(sizes):
        ns = int(ns)
        res = int(4 * math.ceil((8 * ns) ** (1.0 / 3.0) / 4.0 - 1e-9))
        res = max(32, min(128, res))
        lo = (0.375, 0.475, 0.375)
        hi = (0.625, 0.725, 0.625)
        pos_l.append(uniform_box(SEED_BASE + 3 + 1000 * (s + 1), ns, lo, hi))
        mat = s % 2
        rho = 10.0 if mat == 0 else 400.0
        mass_l.append(np.full(ns, rho * 0.25 ** 3 / ns, np.float32))
        mat_l.append(np.full(ns, mat, np.uint16))
        sc = Scene()
        sc.first = first
        sc.count = ns
        sc.nx = sc.ny = sc.nz = res
        sc.dx = 1.0 / res
        sc.origin[:] = (0.0, 0.0, 0.0)
        sc.dt = 5e-5
        sc.velocity_only = 0
        sc.clamp_mode = CLAMP_AUTO
        scenes.append(sc)
        first += ns
    mats = [Material(young_modulus=1e4, poisson_ratio=0.2, density=10.0, model=NEO_HOOKEAN),
            snow_material()]
    wl = Workload("c4_batched_256_scenes", np.concatenate(pos_l), np.concatenate(mass_l),
                  np.concatenate(mat_l), mats, (128, 128, 128, 1.0 / 128, (0.0, 0.0, 0.0)),
                  5e-5, steps, scenes=scenes)
    return wl


def c5_mixed_precision(n=1 << 22, steps=500, res=192, storage=STORE_F32, nh_only=False):
    """4,194,304 particles on a 192^3 grid: a snow box and a Neo-Hookean box
    (E 1e4, nu 0.2, rho 10) side by side; dt 5e-5.  nh_only makes both boxes
    Neo-Hookean (no snow SVD in G2P, which then is bound by HBM traffic)."""
    half = n // 2
    p0 = uniform_box(SEED_BASE + 4, half, (0.15, 0.2, 0.2), (0.45, 0.5, 0.8))
    p1 = uniform_box(SEED_BASE + 4 + 1000, n - half, (0.55, 0.2, 0.2), (0.85, 0.5, 0.8))
    vol = 0.3 * 0.3 * 0.6
    rho0 = 10.0 if nh_only else 400.0
    mass = np.concatenate([np.full(half, rho0 * vol / half, np.float32),
                           np.full(n - half, 10.0 * vol / (n - half), np.float32)])
    mat = np.concatenate([np.full(half, 1 if nh_only else 0, np.uint16), np.ones(n - half, np.uint16)])
    mats = [snow_material(), Material(young_modulus=1e4, poisson_ratio=0.2, density=10.0,
                                      model=NEO_HOOKEAN)]
    return Workload("c5_mixed_4m_192", np.concatenate([p0, p1]), mass, mat, mats,
                    (res, res, res, 1.0 / res, (0.0, 0.0, 0.0)), 5e-5, steps, storage=storage)


def load_engine(eng, wl: Workload):
    """Configure an aemh Engine with a workload (uploads particles)."""
    nx, ny, nz, dx, origin = wl.grid
    eng.set_config(wl.config())
    eng.set_grid(nx, ny, nz, dx, origin)
    eng.set_materials(wl.materials)
    eng.upload(wl.pos, mass=wl.mass, mat=wl.mat)
    if wl.scenes:
        eng.set_scenes(wl.scenes)


def tier_materials():
    """The engine's material kinds (simulation.rs:243-294: elastic, fluid,
    sand) with E scaled so each tier is CFL-stable at dt 1/60 s, dx 1.6 m
    (wave speed <= 26 m/s); the steel entry is left out."""
    return [
        Material(young_modulus=1e5, poisson_ratio=0.3, yield_stress=1e4, hardening=0.01,
                 density=1000.0, friction_angle=35.0, cohesion=0.0, model=NEO_HOOKEAN),
        Material(young_modulus=2e5, poisson_ratio=0.45, yield_stress=0.0, hardening=0.0,
                 density=1000.0, friction_angle=0.0, cohesion=0.0, model=NEWTONIAN_FLUID),
        Material(young_modulus=1e6, poisson_ratio=0.3, yield_stress=0.0, hardening=0.0,
                 density=1600.0, friction_angle=35.0, cohesion=0.0, model=DRUCKER_PRAGER),
    ]


@dataclass
class TierWorld:
    name: str
    capacity: int
    pos: np.ndarray
    mat: np.ndarray
    materials: list

    @property
    def n(self):
        return self.pos.shape[0]


def _shell(rng, n, r0, r1):
    """n points uniform in volume in the upper half shell r0 <= r < r1."""
    u = rng.random(n)
    r = np.cbrt(u * (r1 ** 3 - r0 ** 3) + r0 ** 3)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:, 1] = np.abs(d[:, 1])
    return (d * r[:, None]).astype(np.float32)


def engine_tiers(n_live=2_000_000, capacity=3_000_000, n_near=20_000, n_mid=490_000, seed=6):
    """The engine's resident buffer at scale: MpssBuffer::new(3_000_000)
    (simulation.rs:301) holding n_live particles around a player at the
    origin -- n_near within the 50 m near radius (twice max_near_particles,
    so the cull runs), n_mid in the 50-200 m mid shell (the 490k-particle
    velocity-only tier of ARCHITECTURE_V7_CRITIQUE.md:549), the rest far
    (200-400 m); materials uniform over tier_materials()."""
    rng = np.random.default_rng(SEED_BASE + seed)
    n_far = n_live - n_near - n_mid
    pos = np.concatenate([_shell(rng, n_near, 0.0, 50.0), _shell(rng, n_mid, 50.0, 200.0),
                          _shell(rng, n_far, 200.0, 400.0)])
    perm = rng.permutation(n_live)
    mats = tier_materials()
    return TierWorld(name=f"engine_tiers_{n_live}", capacity=capacity, pos=pos[perm],
                     mat=rng.integers(0, len(mats), n_live).astype(np.uint16), materials=mats)


def to_q32(v):
    """FixedPoint::from_f32 (fixed_point.rs:22-25) of f32 values: f64 of the
    f32 times 2^32, truncated toward zero (the values here are in range)."""
    return (np.asarray(v, np.float32).astype(np.float64) * 4294967296.0).astype(np.int64)


@dataclass
class FixedScene:
    name: str
    config: object        # FixedConfig
    pos: np.ndarray       # [n][3] raw Q32.32
    mass: np.ndarray      # [n] raw Q32.32

    @property
    def n(self):
        return self.pos.shape[0]


def fixed_cloud(n=1 << 20, res=64, seed=7):
    """The all-integer MpmSimulation at scale: n particles uniform
    (xoshiro256**) in the middle half of a res^3 grid, with the material and
    step of the reference's own MpmSimulation tests (small_config,
    ae_physics/src/mpm.rs:957-968: cell 0.5, E 50, density 1, dt 1/60, one
    substep) and its default elastoplastic model; mass = density x cube
    volume / n.  The reference force term (stress x offset / cell^2, no
    volume or dt) makes stiffer settings leave the grid within a few steps;
    this one stays inside it for the first ten."""
    cfg = fixed_config_default()
    cfg.grid_resolution[:] = (res, res, res)
    cfg.cell_size = int(to_q32(0.5))
    cfg.youngs_modulus = int(to_q32(50.0))
    cfg.density = int(to_q32(1.0))
    cfg.dt = int(to_q32(1.0 / 60.0))
    cfg.substeps = 1
    cfg.particle_count = 0
    cfg.material_model = FX_ELASTOPLASTIC
    side = res * 0.25  # metres: half the grid's extent of res * 0.5
    lo = np.full(3, res * 0.125, np.float32)
    pos = uniform_box(SEED_BASE + seed, n, lo, lo + np.float32(side))
    mass = np.full(n, to_q32(np.float32(side ** 3 / n)), np.int64)
    return FixedScene(name=f"fixed_cloud_{n}_{res}", config=cfg, pos=to_q32(pos), mass=mass)
