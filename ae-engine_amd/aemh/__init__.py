"""aemh -- Python binding of the MI355X MLS-MPM C ABI (include/aemh.h).

A thin ctypes layer over ``libaemh.so`` for tests and the benchmark.  It
mirrors the reference step API (ae_particle/src/mpm_solver.rs:1063-1084):
an ``Engine`` owns a device-resident MpssBuffer-shaped particle set, a grid,
a material table and a config, and ``mpm_step_parallel_with_grid`` advances
the particles listed in ``particle_indices`` by ``total_dt``.

There is no CPU fallback: importing works anywhere, but creating an Engine
without the built library or without a gfx950 GPU raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaemh.so")

# aemh_status
OK, ERR_ARG, ERR_HIP, ERR_OOM, ERR_STATE, ERR_OVERFLOW, ERR_NO_DEVICE = range(7)
# ConstitutiveModel order (ae_particle/src/constitutive.rs:11-27) + SNOW
NEO_HOOKEAN, DRUCKER_PRAGER, VON_MISES, NEWTONIAN_FLUID, COROTATED, FIXED_COROTATED, RIGID, SNOW = range(8)
ACCUM_F32, ACCUM_Q32_32, ACCUM_F32_SERIAL = 0, 1, 2
CLAMP_AUTO, CLAMP_SERIAL, CLAMP_WINDOW = 0, 1, 2
STORE_F32, STORE_BF16, STORE_F16 = 0, 1, 2


class AemhError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"aemh status {status}: {msg}")
        self.status = status


class LodConfig(C.Structure):
    """aemh_lod_config: PhysicsLod (simulation.rs:27-48) + tier grids."""
    _fields_ = [
        ("near_field_distance", C.c_float),
        ("mid_field_distance", C.c_float),
        ("far_field_update_interval", C.c_uint32),
        ("max_near_particles", C.c_uint32),
        ("mid_interval", C.c_uint32),
        ("near_n", C.c_uint32),
        ("near_dx", C.c_float),
        ("mid_n", C.c_uint32),
        ("mid_dx", C.c_float),
        ("reserved", C.c_uint32 * 3),
    ]


class TierStats(C.Structure):
    _fields_ = [
        ("near", C.c_uint32), ("mid", C.c_uint32), ("far", C.c_uint32),
        ("reclassified", C.c_uint32), ("near_ran", C.c_uint32), ("mid_ran", C.c_uint32),
        ("far_ran", C.c_uint32), ("expired", C.c_uint32),
        ("near_origin", C.c_float * 3), ("mid_origin", C.c_float * 3),
    ]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["near_origin"] = tuple(self.near_origin)
        d["mid_origin"] = tuple(self.mid_origin)
        return d


class Config(C.Structure):
    _fields_ = [
        ("substeps", C.c_uint32),
        ("gravity", C.c_float * 3),
        ("damping", C.c_float),
        ("bc_margin", C.c_uint32),
        ("bc_restitution", C.c_float),
        ("accum", C.c_uint32),
        ("clamp_mode", C.c_uint32),
        ("storage", C.c_uint32),
        ("resort_interval", C.c_uint32),
        ("reserved", C.c_uint32 * 3),
    ]


class Scene(C.Structure):
    _fields_ = [
        ("first", C.c_uint32),
        ("count", C.c_uint32),
        ("nx", C.c_uint32),
        ("ny", C.c_uint32),
        ("nz", C.c_uint32),
        ("dx", C.c_float),
        ("origin", C.c_float * 3),
        ("dt", C.c_float),
        ("velocity_only", C.c_uint32),
        ("clamp_mode", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


# --- all-integer MpmSimulation (ae_physics/src/mpm.rs) ----------------------
FX_ELASTIC, FX_ELASTOPLASTIC, FX_GRANULAR, FX_BRITTLE, FX_FLUID, FX_SNOW = range(6)
Q32_ONE = 1 << 32


class FixedConfig(C.Structure):
    """MpmConfig (ae_physics/src/mpm.rs:15-31), FixedPoint fields as raw int64."""
    _fields_ = [("grid_resolution", C.c_uint32 * 3), ("cell_size", C.c_int64),
                ("youngs_modulus", C.c_int64), ("poissons_ratio", C.c_int64),
                ("yield_stress", C.c_int64), ("hardening", C.c_int64), ("density", C.c_int64),
                ("gravity", C.c_int64 * 3), ("dt", C.c_int64), ("substeps", C.c_uint32),
                ("material_model", C.c_uint32), ("model_param", C.c_int64),
                ("enable_fracture", C.c_uint32), ("fracture_strain", C.c_int64),
                ("particle_count", C.c_uint64)]


class FixedState(C.Structure):
    _fields_ = [("particles", C.c_uint64), ("step_count", C.c_uint64), ("time", C.c_double),
                ("next_particle_id", C.c_uint64), ("fracture_events", C.c_uint64),
                ("hardening", C.c_int64), ("fracture_strain", C.c_int64),
                ("enable_fracture", C.c_uint32), ("material_model", C.c_uint32),
                ("model_param", C.c_int64), ("checked_substeps", C.c_uint64),
                ("serial_substeps", C.c_uint64), ("max_scatter_bound", C.c_uint64)]


class FractureEvent(C.Structure):
    _fields_ = [("time", C.c_double), ("position", C.c_float * 3),
                ("strain_magnitude", C.c_float), ("num_fragments", C.c_uint32),
                ("pad", C.c_uint32)]


class FixedStats(C.Structure):
    _fields_ = [("active_particles", C.c_uint64), ("fractured_count", C.c_uint64),
                ("total_mass", C.c_int64), ("average_temperature", C.c_int64),
                ("kinetic_energy", C.c_int64), ("total_strain_energy", C.c_int64)]


class FixedSnapshot(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("velocity", C.c_float * 3),
                ("stress_magnitude", C.c_float), ("plastic_strain", C.c_float),
                ("temperature", C.c_float), ("color", C.c_float * 4), ("fractured", C.c_uint32)]


# Every symbol include/aemh.h declares; tests check the library exports them.
EXPORTS = [
    "aemh_config_default", "aemh_version", "aemh_create", "aemh_destroy", "aemh_last_error",
    "aemh_set_config", "aemh_set_grid", "aemh_set_materials", "aemh_upload_particles",
    "aemh_upload_plastic_j", "aemh_download_particles", "aemh_download_plastic_j",
    "aemh_set_indices", "aemh_step", "aemh_mpm_step_parallel_with_grid",
    "aemh_step_velocity_only", "aemh_step_async", "aemh_sync", "aemh_set_scenes",
    "aemh_batch_step", "aemh_autotune",
    "aemh_get_launch_shape", "aemh_set_launch_shape",
    "aemh_download_grid", "aemh_download_grid_q32", "aemh_conservation_f64",
    "aemh_conservation_f64_noalloc", "aemh_svd_fp64_count",
    "aemh_last_substep_info", "aemh_profile_enable", "aemh_profile_get", "aemh_profile_reset",
    "aemh_scene_uniform_box",
    "aemh_mpss_reserve", "aemh_mpss_spawn", "aemh_mpss_kill", "aemh_mpss_compact",
    "aemh_mpss_update_lifetimes", "aemh_mpss_write", "aemh_mpss_read", "aemh_mpss_state",
    "aemh_lod_config_default", "aemh_tier_set_config", "aemh_tier_tick", "aemh_tier_lists",
    "aemh_tier_render",
    "aemh_fixed_config_default", "aemh_fixed_create", "aemh_fixed_destroy",
    "aemh_fixed_last_error", "aemh_fixed_add_particle", "aemh_fixed_add_particles",
    "aemh_fixed_remove_particle", "aemh_fixed_step", "aemh_fixed_set_material_model",
    "aemh_fixed_enable_fracture", "aemh_fixed_disable_fracture", "aemh_fixed_get_state",
    "aemh_fixed_set_sort_interval",
    "aemh_fixed_read_particles", "aemh_fixed_write_particles", "aemh_fixed_fracture_events",
    "aemh_fixed_get_stats", "aemh_fixed_snapshots",
]

NO_SLOT = 0xFFFFFFFF  # spawn() on a full buffer (None in the reference)

_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libaemh.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libaemh.so not built at {path}; run __graft_entry__.build()")
    lib = C.CDLL(path)
    P = C.c_void_p
    fp = C.POINTER(C.c_float)
    u32 = C.c_uint32
    lib.aemh_version.restype = C.c_char_p
    lib.aemh_last_error.restype = C.c_char_p
    lib.aemh_last_error.argtypes = [P]
    lib.aemh_config_default.argtypes = [C.POINTER(Config)]
    lib.aemh_config_default.restype = None
    lib.aemh_create.argtypes = [C.POINTER(P), C.c_int]
    lib.aemh_destroy.argtypes = [P]
    lib.aemh_set_config.argtypes = [P, C.POINTER(Config)]
    lib.aemh_set_grid.argtypes = [P, u32, u32, u32, C.c_float, fp]
    lib.aemh_set_materials.argtypes = [P, u32, fp, C.POINTER(C.c_uint8), fp]
    lib.aemh_upload_particles.argtypes = [P, u32, fp, fp, fp, fp, fp, fp, C.POINTER(C.c_uint16)]
    lib.aemh_upload_plastic_j.argtypes = [P, fp]
    lib.aemh_download_particles.argtypes = [P, fp, fp, fp, fp, fp]
    lib.aemh_download_plastic_j.argtypes = [P, fp]
    lib.aemh_set_indices.argtypes = [P, C.POINTER(u32), u32]
    lib.aemh_step.argtypes = [P, C.c_float]
    lib.aemh_mpm_step_parallel_with_grid.argtypes = [P, C.POINTER(u32), u32, C.c_float]
    lib.aemh_step_velocity_only.argtypes = [P, C.c_float]
    lib.aemh_step_async.argtypes = [P, C.c_float]
    lib.aemh_sync.argtypes = [P]
    lib.aemh_set_scenes.argtypes = [P, C.POINTER(Scene), u32]
    lib.aemh_batch_step.argtypes = [P, u32]
    lib.aemh_autotune.argtypes = [P]
    lib.aemh_get_launch_shape.argtypes = [P, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32),
                                          C.POINTER(u32), fp]
    lib.aemh_set_launch_shape.argtypes = [P, u32, u32, u32, u32]
    lib.aemh_download_grid.argtypes = [P, u32, fp, fp, fp, fp]
    lib.aemh_download_grid_q32.argtypes = [P, u32, C.POINTER(C.c_int64)]
    lib.aemh_conservation_f64.argtypes = [P, C.POINTER(C.c_double)]
    lib.aemh_conservation_f64_noalloc.argtypes = [P, C.POINTER(C.c_double)]
    lib.aemh_svd_fp64_count.argtypes = [P, C.POINTER(C.c_uint64)]
    lib.aemh_last_substep_info.argtypes = [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.aemh_profile_enable.argtypes = [P, C.c_int]
    lib.aemh_profile_get.argtypes = [P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    lib.aemh_profile_reset.argtypes = [P]
    lib.aemh_scene_uniform_box.argtypes = [C.c_uint64, u32, fp, fp, fp]
    lib.aemh_scene_uniform_box.restype = None
    pu32 = C.POINTER(u32)
    pu16 = C.POINTER(C.c_uint16)
    lib.aemh_mpss_reserve.argtypes = [P, u32]
    lib.aemh_mpss_spawn.argtypes = [P, u32, pu32, pu32]
    lib.aemh_mpss_kill.argtypes = [P, pu32, u32]
    lib.aemh_mpss_compact.argtypes = [P]
    lib.aemh_mpss_update_lifetimes.argtypes = [P, C.c_float, pu32, u32, pu32]
    lib.aemh_mpss_write.argtypes = [P, pu32, u32, fp, fp, fp, fp, fp, fp, fp, pu16, fp, fp]
    lib.aemh_mpss_read.argtypes = [P, pu32, u32, fp, fp, fp, fp, fp, fp, fp, pu16, fp, fp]
    lib.aemh_mpss_state.argtypes = [P, pu32, pu32, C.POINTER(C.c_uint8)]
    lib.aemh_lod_config_default.argtypes = [C.POINTER(LodConfig)]
    lib.aemh_lod_config_default.restype = None
    lib.aemh_tier_set_config.argtypes = [P, C.POINTER(LodConfig)]
    lib.aemh_tier_tick.argtypes = [P, fp, C.c_float, C.POINTER(TierStats)]
    lib.aemh_tier_lists.argtypes = [P, pu32, pu32, pu32, pu32]
    lib.aemh_tier_render.argtypes = [P, u32, u32, fp, pu32, pu32]
    i64p = C.POINTER(C.c_int64)
    u64p = C.POINTER(C.c_uint64)
    lib.aemh_fixed_config_default.argtypes = [C.POINTER(FixedConfig)]
    lib.aemh_fixed_create.argtypes = [C.POINTER(P), C.c_int, C.POINTER(FixedConfig)]
    lib.aemh_fixed_destroy.argtypes = [P]
    lib.aemh_fixed_last_error.restype = C.c_char_p
    lib.aemh_fixed_last_error.argtypes = [P]
    lib.aemh_fixed_add_particle.argtypes = [P, i64p, i64p, C.c_int64, fp, u64p]
    lib.aemh_fixed_add_particles.argtypes = [P, C.c_uint64, i64p, i64p, i64p, fp, u64p]
    lib.aemh_fixed_remove_particle.argtypes = [P, C.c_uint64, C.POINTER(C.c_int)]
    lib.aemh_fixed_step.argtypes = [P, u32]
    lib.aemh_fixed_set_material_model.argtypes = [P, u32, C.c_int64]
    lib.aemh_fixed_enable_fracture.argtypes = [P, C.c_int64]
    lib.aemh_fixed_set_sort_interval.argtypes = [P, u32]
    lib.aemh_fixed_disable_fracture.argtypes = [P]
    lib.aemh_fixed_get_state.argtypes = [P, C.POINTER(FixedState)]
    lib.aemh_fixed_read_particles.argtypes = [P, C.c_uint64, C.c_uint64, i64p,
                                              C.POINTER(C.c_uint8), u64p, fp]
    lib.aemh_fixed_write_particles.argtypes = [P, C.c_uint64, C.c_uint64, i64p,
                                               C.POINTER(C.c_uint8), u64p, fp]
    lib.aemh_fixed_fracture_events.argtypes = [P, C.c_uint64, C.c_uint64,
                                               C.POINTER(FractureEvent)]
    lib.aemh_fixed_get_stats.argtypes = [P, C.POINTER(FixedStats)]
    lib.aemh_fixed_snapshots.argtypes = [P, C.POINTER(FixedSnapshot), C.c_uint64, u64p]
    for name in EXPORTS:
        f = getattr(lib, name)
        if name not in ("aemh_version", "aemh_last_error", "aemh_config_default",
                        "aemh_scene_uniform_box", "aemh_lod_config_default",
                        "aemh_fixed_config_default", "aemh_fixed_destroy",
                        "aemh_fixed_last_error"):
            f.restype = C.c_int
    _lib = lib
    return lib


def _f(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _c(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def _u32(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint32))


def _u16(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_uint16))


# Per-slot fields of the resident buffer: name -> (components, dtype)
MPSS_FIELDS = {"pos": (3, np.float32), "vel": (3, np.float32), "C": (9, np.float32),
               "F": (9, np.float32), "J": (1, np.float32), "Jp": (1, np.float32),
               "mass": (1, np.float32), "mat": (1, np.uint16), "lifetime": (1, np.float32),
               "age": (1, np.float32)}


def default_lod_config() -> LodConfig:
    cfg = LodConfig()
    load_library().aemh_lod_config_default(C.byref(cfg))
    return cfg


def default_config() -> Config:
    cfg = Config()
    load_library().aemh_config_default(C.byref(cfg))
    return cfg


def uniform_box(seed: int, n: int, lo, hi) -> np.ndarray:
    """Seeded xoshiro256** positions uniform in the box [lo, hi)."""
    lib = load_library()
    out = np.empty((n, 3), np.float32)
    lo_a = np.asarray(lo, np.float32)
    hi_a = np.asarray(hi, np.float32)
    lib.aemh_scene_uniform_box(C.c_uint64(seed), n, _f(lo_a), _f(hi_a), _f(out))
    return out


@dataclass
class Material:
    """MaterialParams (ae_particle/src/constitutive.rs:31-39) + model + snow params."""
    young_modulus: float = 1e6
    poisson_ratio: float = 0.3
    yield_stress: float = 1e4
    hardening: float = 0.01
    density: float = 1000.0
    friction_angle: float = 35.0
    cohesion: float = 0.0
    model: int = NEO_HOOKEAN
    theta_c: float = 2.5e-2
    theta_s: float = 7.5e-3
    xi: float = 10.0

    def params7(self):
        return [self.young_modulus, self.poisson_ratio, self.yield_stress, self.hardening,
                self.density, self.friction_angle, self.cohesion]


def material_arrays(mats):
    params = np.array([m.params7() for m in mats], np.float32)
    models = np.array([m.model for m in mats], np.uint8)
    ext = np.array([[m.theta_c, m.theta_s, m.xi] for m in mats], np.float32)
    return params, models, ext


class Engine:
    """One device context: resident particles, grid, materials, config."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        h = C.c_void_p()
        st = self._lib.aemh_create(C.byref(h), device)
        if st != OK:
            raise AemhError(st, "aemh_create failed (no gfx950 device or HIP unusable)")
        self._h = h
        self.n = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.aemh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != OK:
            raise AemhError(st, self._lib.aemh_last_error(self._h).decode())

    # -- configuration --------------------------------------------------------
    def set_config(self, cfg: Config):
        self._check(self._lib.aemh_set_config(self._h, C.byref(cfg)))

    def set_grid(self, nx, ny, nz, dx, origin=(0.0, 0.0, 0.0)):
        o = np.asarray(origin, np.float32)
        self._check(self._lib.aemh_set_grid(self._h, nx, ny, nz, C.c_float(dx), _f(o)))

    def set_materials(self, mats):
        params, models, ext = material_arrays(mats)
        self._check(self._lib.aemh_set_materials(
            self._h, len(mats), _f(params), models.ctypes.data_as(C.POINTER(C.c_uint8)), _f(ext)))

    # -- particles ------------------------------------------------------------
    def upload(self, pos, vel=None, C9=None, F9=None, J=None, mass=None, mat=None, Jp=None):
        pos = _c(pos, np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel, C9, F9, J, mass = (_c(a, np.float32) for a in (vel, C9, F9, J, mass))
        mat = _c(mat, np.uint16)
        self._check(self._lib.aemh_upload_particles(
            self._h, n, _f(pos), _f(vel), _f(C9), _f(F9), _f(J), _f(mass),
            None if mat is None else mat.ctypes.data_as(C.POINTER(C.c_uint16))))
        if Jp is not None:
            Jp = _c(Jp, np.float32)
            self._check(self._lib.aemh_upload_plastic_j(self._h, _f(Jp)))
        self.n = n

    def download(self):
        n = self.n
        pos = np.empty((n, 3), np.float32)
        vel = np.empty((n, 3), np.float32)
        C9 = np.empty((n, 9), np.float32)
        F9 = np.empty((n, 9), np.float32)
        J = np.empty(n, np.float32)
        self._check(self._lib.aemh_download_particles(self._h, _f(pos), _f(vel), _f(C9), _f(F9), _f(J)))
        Jp = np.empty(n, np.float32)
        self._check(self._lib.aemh_download_plastic_j(self._h, _f(Jp)))
        return dict(pos=pos, vel=vel, C=C9, F=F9, J=J, Jp=Jp)

    def set_indices(self, idx):
        idx = _c(idx, np.uint32)
        self._check(self._lib.aemh_set_indices(self._h, idx.ctypes.data_as(C.POINTER(C.c_uint32)), idx.size))

    # -- steps ----------------------------------------------------------------
    def step(self, total_dt: float):
        self._check(self._lib.aemh_step(self._h, C.c_float(total_dt)))

    def step_async(self, total_dt: float):
        """aemh_step without waiting for the GPU (sync() waits)."""
        self._check(self._lib.aemh_step_async(self._h, C.c_float(total_dt)))

    def sync(self):
        self._check(self._lib.aemh_sync(self._h))

    def mpm_step_parallel_with_grid(self, particle_indices, total_dt: float):
        idx = _c(particle_indices, np.uint32)
        self._check(self._lib.aemh_mpm_step_parallel_with_grid(
            self._h, idx.ctypes.data_as(C.POINTER(C.c_uint32)), idx.size, C.c_float(total_dt)))

    def step_velocity_only(self, total_dt: float):
        self._check(self._lib.aemh_step_velocity_only(self._h, C.c_float(total_dt)))

    def set_scenes(self, scenes):
        arr = (Scene * len(scenes))(*scenes)
        self._check(self._lib.aemh_set_scenes(self._h, arr, len(scenes)))

    def batch_step(self, n_substeps: int):
        self._check(self._lib.aemh_batch_step(self._h, n_substeps))

    def autotune(self):
        """Pick P2G tiles, grid-update workgroups and G2P workgroup size by
        timing them on the current state (which is not advanced)."""
        self._check(self._lib.aemh_autotune(self._h))
        return self.launch_shape()

    def launch_shape(self):
        """dict(p2g_tiles, p2g_partials, p2g_pairs, p2g_layers, g2p_wg,
        upd_wgs, tuned_ms=(p2g + update, update, g2p)); the ABI carries
        p2g_pairs and p2g_layers as bits 1 and 2 of p2g_partials."""
        a, p, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        ms = np.zeros(3, np.float32)
        self._check(self._lib.aemh_get_launch_shape(self._h, C.byref(a), C.byref(p), C.byref(b),
                                                    C.byref(c), _f(ms)))
        return dict(p2g_tiles=a.value, p2g_partials=p.value & 1, p2g_pairs=(p.value >> 1) & 1,
                    p2g_layers=(p.value >> 2) & 1,
                    g2p_wg=b.value, upd_wgs=c.value, tuned_ms=tuple(float(x) for x in ms))

    def set_launch_shape(self, p2g_tiles=3, p2g_partials=1, g2p_wg=256, upd_wgs=2048,
                         p2g_pairs=0, p2g_layers=0):
        self._check(self._lib.aemh_set_launch_shape(
            self._h, p2g_tiles, p2g_partials | (p2g_pairs << 1) | (p2g_layers << 2), g2p_wg,
            upd_wgs))

    # -- readback ---------------------------------------------------------------
    def download_grid(self, scene=0, nodes=None):
        m = np.zeros(nodes, np.float32)
        p = np.zeros((nodes, 3), np.float32)
        f = np.zeros((nodes, 3), np.float32)
        v = np.zeros((nodes, 3), np.float32)
        self._check(self._lib.aemh_download_grid(self._h, scene, _f(m), _f(p), _f(f), _f(v)))
        return dict(m=m, p=p, f=f, v=v)

    def download_grid_q32(self, scene=0, nodes=None):
        q = np.zeros((nodes, 7), np.int64)
        self._check(self._lib.aemh_download_grid_q32(self._h, scene, q.ctypes.data_as(C.POINTER(C.c_int64))))
        return q

    def conservation_f64(self):
        out = (C.c_double * 5)()
        self._check(self._lib.aemh_conservation_f64(self._h, out))
        return np.array(out[:], np.float64)

    def conservation_f64_noalloc(self):
        """conservation_f64 through the entry point that allocates nothing
        (its partial sums use a static device buffer)."""
        out = (C.c_double * 5)()
        self._check(self._lib.aemh_conservation_f64_noalloc(self._h, out))
        return np.array(out[:], np.float64)

    def svd_fp64_count(self) -> int:
        """Snow SVDs that took the FP64 path since the particles were uploaded."""
        v = C.c_uint64()
        self._check(self._lib.aemh_svd_fp64_count(self._h, C.byref(v)))
        return int(v.value)

    def last_substep_info(self):
        """(active 4^3 node blocks, particles) of the last substep."""
        nb = C.c_uint32()
        npart = C.c_uint32()
        self._check(self._lib.aemh_last_substep_info(self._h, C.byref(nb), C.byref(npart)))
        return nb.value, npart.value

    # -- profiling ----------------------------------------------------------------
    def profile(self, on=True):
        self._check(self._lib.aemh_profile_enable(self._h, 1 if on else 0))

    def profile_reset(self):
        self._check(self._lib.aemh_profile_reset(self._h))

    def profile_get(self, kernel: str):
        ms = C.c_double()
        cnt = C.c_uint64()
        self._check(self._lib.aemh_profile_get(self._h, kernel.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    # -- resident MpssBuffer (mpss.rs:161-297, 431-445) ---------------------------
    def mpss_reserve(self, capacity: int):
        self._check(self._lib.aemh_mpss_reserve(self._h, capacity))
        self.n = capacity

    def mpss_spawn(self, n: int) -> np.ndarray:
        """spawn() n times; slots, NO_SLOT where the buffer was full."""
        out = np.empty(max(n, 1), np.uint32)
        got = C.c_uint32(0)
        self._check(self._lib.aemh_mpss_spawn(self._h, n, _u32(out), C.byref(got)))
        return out[:n]

    def mpss_kill(self, slots):
        slots = _c(np.atleast_1d(slots), np.uint32)
        self._check(self._lib.aemh_mpss_kill(self._h, _u32(slots), slots.size))

    def mpss_compact(self):
        self._check(self._lib.aemh_mpss_compact(self._h))

    def mpss_update_lifetimes(self, dt: float) -> np.ndarray:
        out = np.empty(max(self.n, 1), np.uint32)
        ne = C.c_uint32(0)
        self._check(self._lib.aemh_mpss_update_lifetimes(self._h, C.c_float(dt), _u32(out),
                                                         out.size, C.byref(ne)))
        return out[:ne.value].copy()

    def mpss_write(self, slots, **fields):
        slots = _c(np.atleast_1d(slots), np.uint32)
        n = slots.size
        args = []
        for name, (comps, dt) in MPSS_FIELDS.items():
            a = fields.pop(name, None)
            if a is not None:
                a = _c(a, dt).reshape(-1)
                if a.size != n * comps:
                    raise ValueError(f"{name}: expected {n * comps} values, got {a.size}")
            args.append(a)
        if fields:
            raise ValueError(f"unknown fields {sorted(fields)}")
        ptrs = [_u16(a) if name == "mat" else _f(a) for name, a in zip(MPSS_FIELDS, args)]
        self._check(self._lib.aemh_mpss_write(self._h, _u32(slots), n, *ptrs))

    def mpss_read(self, slots, fields=tuple(MPSS_FIELDS)):
        slots = _c(np.atleast_1d(slots), np.uint32)
        n = slots.size
        out = {}
        ptrs = []
        for name, (comps, dt) in MPSS_FIELDS.items():
            if name in fields:
                a = np.empty((n, comps) if comps > 1 else n, dt)
                out[name] = a
                ptrs.append(_u16(a) if name == "mat" else _f(a))
            else:
                ptrs.append(None)
        self._check(self._lib.aemh_mpss_read(self._h, _u32(slots), n, *ptrs))
        return out

    def mpss_state(self):
        """(count, capacity, active flags as bool[capacity])."""
        cnt = C.c_uint32(0)
        cap = C.c_uint32(0)
        self._check(self._lib.aemh_mpss_state(self._h, C.byref(cnt), C.byref(cap), None))
        act = np.empty(cap.value, np.uint8)
        self._check(self._lib.aemh_mpss_state(self._h, None, None,
                                              act.ctypes.data_as(C.POINTER(C.c_uint8))))
        return cnt.value, cap.value, act.astype(bool)

    # -- LOD tier orchestrator (simulation.rs:480-700) ----------------------------
    def tier_set_config(self, lod: LodConfig):
        self._check(self._lib.aemh_tier_set_config(self._h, C.byref(lod)))

    def tier_tick(self, player, dt: float) -> dict:
        p = np.ascontiguousarray(player, np.float32).reshape(3)
        st = TierStats()
        self._check(self._lib.aemh_tier_tick(self._h, _f(p), C.c_float(dt), C.byref(st)))
        return st.as_dict()

    def tier_lists(self):
        """(near, mid, far) cached slot lists."""
        n3 = np.zeros(3, np.uint32)
        self._check(self._lib.aemh_tier_lists(self._h, None, None, None, _u32(n3)))
        ls = [np.empty(max(int(k), 1), np.uint32) for k in n3]
        self._check(self._lib.aemh_tier_lists(self._h, _u32(ls[0]), _u32(ls[1]), _u32(ls[2]),
                                              None))
        return tuple(a[:int(k)] for a, k in zip(ls, n3))

    def tier_render(self, tier: int, take: int):
        """Live entries among the first `take` of a cached list: (pos[n,3], slots[n])."""
        m = max(take, 1)
        pos = np.empty((m, 3), np.float32)
        slots = np.empty(m, np.uint32)
        n = C.c_uint32(0)
        self._check(self._lib.aemh_tier_render(self._h, tier, take, _f(pos), _u32(slots),
                                               C.byref(n)))
        return pos[:n.value].copy(), slots[:n.value].copy()


class MpssBuffer:
    """The reference's MpssBuffer interface (ae_particle/src/mpss.rs) over an
    Engine whose particle arrays stay in HBM.

    spawn() returns a slot or None, kill(idx) ignores dead or out-of-range
    slots, compact() and update_lifetimes(dt) follow mpss.rs:283-297 and
    :431-445; field access goes through read()/write() by slot list.
    """

    def __init__(self, engine: Engine, capacity: int):
        self.engine = engine
        self.capacity = capacity
        engine.mpss_reserve(capacity)

    def spawn(self):
        s = int(self.engine.mpss_spawn(1)[0])
        return None if s == NO_SLOT else s

    def spawn_many(self, n: int) -> np.ndarray:
        return self.engine.mpss_spawn(n)

    def kill(self, idx):
        self.engine.mpss_kill(idx)

    def compact(self):
        self.engine.mpss_compact()

    def update_lifetimes(self, dt: float):
        return [int(i) for i in self.engine.mpss_update_lifetimes(dt)]

    def len(self) -> int:
        return self.engine.mpss_state()[0]

    def __len__(self) -> int:
        return self.len()

    def is_empty(self) -> bool:
        return self.len() == 0

    @property
    def active(self) -> np.ndarray:
        return self.engine.mpss_state()[2]

    def active_indices(self) -> np.ndarray:
        return np.flatnonzero(self.active).astype(np.uint32)

    def write(self, slots, **fields):
        self.engine.mpss_write(slots, **fields)

    def read(self, slots, fields=tuple(MPSS_FIELDS)):
        return self.engine.mpss_read(slots, fields)


def fixed_config_default() -> FixedConfig:
    """MpmConfig::default (ae_physics/src/mpm.rs:33-52)."""
    cfg = FixedConfig()
    load_library().aemh_fixed_config_default(C.byref(cfg))
    return cfg


class FixedSim:
    """MpmSimulation (ae_physics/src/mpm.rs:116-792) resident on one gfx950
    device; FixedPoint values are raw int64 (value = raw / 2^32).

    `cfg` is a FixedConfig or any ctypes structure with its layout."""

    def __init__(self, cfg, device: int = 0):
        self._lib = load_library()
        if not isinstance(cfg, FixedConfig):
            cfg = FixedConfig.from_buffer_copy(bytes(cfg))
        h = C.c_void_p()
        st = self._lib.aemh_fixed_create(C.byref(h), device, C.byref(cfg))
        if st != OK:
            raise AemhError(st, "aemh_fixed_create failed (bad config, no gfx950 device or HIP "
                                "unusable)")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.aemh_fixed_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != OK:
            raise AemhError(st, self._lib.aemh_fixed_last_error(self._h).decode())

    def step(self, n: int = 1):
        self._check(self._lib.aemh_fixed_step(self._h, n))

    def state(self) -> dict:
        s = FixedState()
        self._check(self._lib.aemh_fixed_get_state(self._h, C.byref(s)))
        return dict(count=s.particles, step_count=s.step_count, time=s.time,
                    next_id=s.next_particle_id, n_events=s.fracture_events,
                    hardening=s.hardening, fracture_strain=s.fracture_strain,
                    fracture_on=s.enable_fracture, model=s.material_model,
                    model_param=s.model_param, checked_substeps=s.checked_substeps,
                    serial_substeps=s.serial_substeps, max_scatter_bound=s.max_scatter_bound)

    def count(self) -> int:
        return self.state()["count"]

    def add_particle(self, pos, vel, mass, color=(1.0, 0.0, 0.0, 1.0)) -> int:
        p = (C.c_int64 * 3)(*[int(v) for v in pos])
        v = (C.c_int64 * 3)(*[int(x) for x in vel])
        col = (C.c_float * 4)(*color)
        pid = C.c_uint64()
        self._check(self._lib.aemh_fixed_add_particle(self._h, p, v, int(mass), col,
                                                       C.byref(pid)))
        return pid.value

    def add_particles(self, pos, vel, mass, color=None) -> int:
        pos = np.ascontiguousarray(pos, np.int64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.int64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, np.int64).reshape(-1)
        col = None if color is None else np.ascontiguousarray(color, np.float32).reshape(-1, 4)
        first = C.c_uint64()
        i64p = C.POINTER(C.c_int64)
        self._check(self._lib.aemh_fixed_add_particles(
            self._h, len(mass), pos.ctypes.data_as(i64p), vel.ctypes.data_as(i64p),
            mass.ctypes.data_as(i64p), _f(col), C.byref(first)))
        return first.value

    def remove_particle(self, pid: int) -> bool:
        r = C.c_int()
        self._check(self._lib.aemh_fixed_remove_particle(self._h, C.c_uint64(pid), C.byref(r)))
        return bool(r.value)

    def set_model(self, model: int, param: int = 0):
        self._check(self._lib.aemh_fixed_set_material_model(self._h, model, int(param)))

    def set_sort_interval(self, substeps: int):
        """Re-sort interval of the cell-ordered P2G (default 1)."""
        self._check(self._lib.aemh_fixed_set_sort_interval(self._h, int(substeps)))

    def set_fracture(self, on: bool, threshold: int = 0):
        if on:
            self._check(self._lib.aemh_fixed_enable_fracture(self._h, int(threshold)))
        else:
            self._check(self._lib.aemh_fixed_disable_fracture(self._h))

    def read(self, first: int = 0, count=None) -> dict:
        n = self.count()
        count = n - first if count is None else count
        q = np.zeros((max(count, 1), 29), np.int64)
        fl = np.zeros((max(count, 1), 3), np.uint8)
        ids = np.zeros((max(count, 1), 9), np.uint64)
        col = np.zeros((max(count, 1), 4), np.float32)
        self._check(self._lib.aemh_fixed_read_particles(
            self._h, first, count, q.ctypes.data_as(C.POINTER(C.c_int64)),
            fl.ctypes.data_as(C.POINTER(C.c_uint8)), ids.ctypes.data_as(C.POINTER(C.c_uint64)),
            _f(col)))
        return dict(q=q[:count], flags=fl[:count], ids=ids[:count], color=col[:count])

    def write(self, first: int, q, flags, ids, color):
        q = np.ascontiguousarray(q, np.int64).reshape(-1, 29)
        fl = np.ascontiguousarray(flags, np.uint8).reshape(-1, 3)
        ids = np.ascontiguousarray(ids, np.uint64).reshape(-1, 9)
        col = np.ascontiguousarray(color, np.float32).reshape(-1, 4)
        self._check(self._lib.aemh_fixed_write_particles(
            self._h, first, len(q), q.ctypes.data_as(C.POINTER(C.c_int64)),
            fl.ctypes.data_as(C.POINTER(C.c_uint8)), ids.ctypes.data_as(C.POINTER(C.c_uint64)),
            _f(col)))

    def events(self) -> dict:
        ne = self.state()["n_events"]
        ev = (FractureEvent * max(ne, 1))()
        self._check(self._lib.aemh_fixed_fracture_events(self._h, 0, ne, ev))
        return dict(time=np.array([e.time for e in ev[:ne]], np.float64),
                    pos=np.array([list(e.position) for e in ev[:ne]], np.float32).reshape(-1, 3),
                    strain=np.array([e.strain_magnitude for e in ev[:ne]], np.float32),
                    frags=np.array([e.num_fragments for e in ev[:ne]], np.uint32))

    def stats(self) -> dict:
        s = FixedStats()
        self._check(self._lib.aemh_fixed_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in FixedStats._fields_}

    def snapshots(self) -> list:
        n = C.c_uint64()
        self._check(self._lib.aemh_fixed_snapshots(self._h, None, 0, C.byref(n)))
        out = (FixedSnapshot * max(n.value, 1))()
        self._check(self._lib.aemh_fixed_snapshots(self._h, out, n.value, C.byref(n)))
        return list(out[:n.value])

    def active_count(self) -> int:
        return self.stats()["active_particles"]
