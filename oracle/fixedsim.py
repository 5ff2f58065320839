"""ctypes wrapper of the FixedPoint MpmSimulation oracle (libfixedsim.so,
built from oracle/fixedsim.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the device
solver; never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# AEMH_ORACLE_FIXEDSIM names a throw-away build with a planted mistake (only
# tests/test_oracle_pins.py sets it, to show the pins fail on it).
LIB_PATH = os.environ.get("AEMH_ORACLE_FIXEDSIM") or os.path.join(_HERE, "libfixedsim.so")

ELASTIC, ELASTOPLASTIC, GRANULAR, BRITTLE, FLUID, SNOW = range(6)
ONE = 1 << 32


class FxConfig(C.Structure):
    _fields_ = [("res", C.c_uint32 * 3), ("cell", C.c_int64), ("E", C.c_int64), ("nu", C.c_int64),
                ("yield_", C.c_int64), ("hardening", C.c_int64), ("density", C.c_int64),
                ("gravity", C.c_int64 * 3), ("dt", C.c_int64), ("substeps", C.c_uint32),
                ("model", C.c_uint32), ("model_param", C.c_int64),
                ("enable_fracture", C.c_uint32), ("fracture_strain", C.c_int64),
                ("particle_count", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", _HERE, "libfixedsim.so"], check=True)
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.fx_sim_new.restype = P
        L.fx_sim_new.argtypes = [C.POINTER(FxConfig)]
        L.fx_sim_free.argtypes = [P]
        L.fx_sim_step.argtypes = [P, C.c_uint32]
        L.fx_sim_count.restype = C.c_uint64
        L.fx_sim_count.argtypes = [P]
        L.fx_sim_add_particle.restype = C.c_uint64
        L.fx_sim_add_particle.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int64,
                                          C.POINTER(C.c_float)]
        L.fx_sim_add_particles.restype = C.c_uint64
        L.fx_sim_add_particles.argtypes = [P, C.c_uint64, C.POINTER(C.c_int64),
                                           C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.fx_sim_remove_particle.restype = C.c_int
        L.fx_sim_remove_particle.argtypes = [P, C.c_uint64]
        L.fx_sim_set_model.argtypes = [P, C.c_uint32, C.c_int64]
        L.fx_sim_set_fracture.argtypes = [P, C.c_uint32, C.c_int64]
        L.fx_sim_read.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_uint8),
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_float)]
        L.fx_sim_state.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                   C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_uint32)]
        L.fx_sim_events.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_float),
                                    C.POINTER(C.c_float), C.POINTER(C.c_uint32)]
        L.fx_sim_diag.argtypes = [P, C.POINTER(C.c_int64)]
        L.fx_sim_write.argtypes = [P, C.c_uint64, C.c_uint64, C.POINTER(C.c_int64),
                                   C.POINTER(C.c_uint8), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_float)]
        L.fx_sim_scatter_bound.restype = C.c_uint64
        L.fx_sim_scatter_bound.argtypes = [P]
        L.fx_sim_fracture_only.argtypes = [P]
        L.fx_constitutive.argtypes = [C.POINTER(FxConfig), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                      C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_int64)]
        L.fx_sim_stats.argtypes = [P, C.POINTER(C.c_int64)]
        L.fx_sim_snapshots.restype = C.c_uint64
        L.fx_sim_snapshots.argtypes = [P, C.POINTER(C.c_float), C.POINTER(C.c_uint8)]
        L.fx_op.restype = C.c_int64
        L.fx_op.argtypes = [C.c_int, C.c_int64, C.c_int64]
        L.fx_from_f32.restype = C.c_int64
        L.fx_from_f32.argtypes = [C.c_float]
        L.fx_to_f32.restype = C.c_float
        L.fx_to_f32.argtypes = [C.c_int64]
        L.fx_sin_table.restype = C.c_int64
        L.fx_sin_table.argtypes = [C.c_int]
        L.fx_mat_op.argtypes = [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def from_f32(v):
    return int(lib().fx_from_f32(float(v)))


def to_f32(r):
    return float(lib().fx_to_f32(int(r)))


OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "sqrt": 4, "ln": 5, "exp": 6, "sin": 7, "cos": 8,
       "atan2": 9, "neg": 10, "abs": 11, "bspline": 12}


def op(name, a, b=0):
    return int(lib().fx_op(OPS[name], int(a), int(b)))


def mat_op(which, m9):
    m = np.ascontiguousarray(m9, np.int64).reshape(9)
    out = np.zeros(27, np.int64)
    lib().fx_mat_op(which, m.ctypes.data_as(C.POINTER(C.c_int64)),
                    out.ctypes.data_as(C.POINTER(C.c_int64)))
    return out


def constitutive(cfg, F9, S9, vel3=(0, 0, 0), plastic=0):
    """handle_constitutive_model on one particle (F, stress column-major);
    returns (F9, S9, vel3, plastic, hardening, fracture_strain, fracture_on)."""
    P64 = C.POINTER(C.c_int64)
    f = np.ascontiguousarray(F9, np.int64)
    st = np.ascontiguousarray(S9, np.int64)
    v = np.ascontiguousarray(vel3, np.int64)
    out = np.zeros(25, np.int64)
    lib().fx_constitutive(C.byref(cfg), f.ctypes.data_as(P64), st.ctypes.data_as(P64),
                          v.ctypes.data_as(P64), int(plastic), out.ctypes.data_as(P64))
    return out[:9], out[9:18], out[18:21], int(out[21]), int(out[22]), int(out[23]), int(out[24])


def default_config(**kw):
    """MpmConfig::default (ae_physics/src/mpm.rs:33-52), overridable."""
    c = FxConfig()
    c.res[:] = (64, 64, 64)
    c.cell = from_f32(0.1)
    c.E = from_f32(1e6)
    c.nu = from_f32(0.3)
    c.yield_ = from_f32(1e4)
    c.hardening = from_f32(0.1)
    c.density = from_f32(1000.0)
    c.gravity[:] = (0, from_f32(-9.81), 0)
    c.dt = from_f32(1.0 / 240.0)
    c.substeps = 4
    c.model = ELASTOPLASTIC
    c.model_param = 0
    c.enable_fracture = 0
    c.fracture_strain = from_f32(0.15)
    c.particle_count = 10000
    for k, v in kw.items():
        if k == "res":
            c.res[:] = v
        elif k == "gravity":
            c.gravity[:] = v
        else:
            setattr(c, k, v)
    return c


def small_config(**kw):
    """small_config of the reference tests (mpm.rs:957-968)."""
    base = dict(res=(4, 4, 4), cell=from_f32(0.5), particle_count=8, E=from_f32(50.0),
                density=from_f32(1.0), dt=from_f32(1.0 / 60.0), substeps=1)
    base.update(kw)
    return default_config(**base)


class Sim:
    def __init__(self, cfg):
        self.h = None
        if not isinstance(cfg, FxConfig):  # any structure with the MpmConfig layout
            cfg = FxConfig.from_buffer_copy(bytes(cfg))
        self.h = lib().fx_sim_new(C.byref(cfg))

    def close(self):
        if self.h:
            lib().fx_sim_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n=1):
        lib().fx_sim_step(self.h, n)

    def count(self):
        return int(lib().fx_sim_count(self.h))

    def add_particle(self, pos, vel, mass, color=(1.0, 0.0, 0.0, 1.0)):
        p = (C.c_int64 * 3)(*pos)
        v = (C.c_int64 * 3)(*vel)
        col = (C.c_float * 4)(*color)
        return int(lib().fx_sim_add_particle(self.h, p, v, int(mass), col))

    def add_particles(self, pos, vel, mass):
        pos = np.ascontiguousarray(pos, np.int64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, np.int64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, np.int64).reshape(-1)
        P64 = C.POINTER(C.c_int64)
        return int(lib().fx_sim_add_particles(self.h, len(mass), pos.ctypes.data_as(P64),
                                              vel.ctypes.data_as(P64), mass.ctypes.data_as(P64)))

    def remove_particle(self, pid):
        return bool(lib().fx_sim_remove_particle(self.h, pid))

    def set_model(self, model, param=0):
        lib().fx_sim_set_model(self.h, model, param)

    def set_fracture(self, on, threshold=0):
        lib().fx_sim_set_fracture(self.h, 1 if on else 0, threshold)

    def read(self):
        n = self.count()
        q = np.zeros((max(n, 1), 29), np.int64)
        fl = np.zeros((max(n, 1), 3), np.uint8)
        ids = np.zeros((max(n, 1), 9), np.uint64)
        col = np.zeros((max(n, 1), 4), np.float32)
        lib().fx_sim_read(self.h, q.ctypes.data_as(C.POINTER(C.c_int64)),
                          fl.ctypes.data_as(C.POINTER(C.c_uint8)),
                          ids.ctypes.data_as(C.POINTER(C.c_uint64)),
                          col.ctypes.data_as(C.POINTER(C.c_float)))
        return dict(q=q[:n], flags=fl[:n], ids=ids[:n], color=col[:n])

    def write(self, first, q, flags, ids, color):
        q = np.ascontiguousarray(q, np.int64).reshape(-1, 29)
        fl = np.ascontiguousarray(flags, np.uint8).reshape(-1, 3)
        ids = np.ascontiguousarray(ids, np.uint64).reshape(-1, 9)
        col = np.ascontiguousarray(color, np.float32).reshape(-1, 4)
        lib().fx_sim_write(self.h, first, len(q), q.ctypes.data_as(C.POINTER(C.c_int64)),
                           fl.ctypes.data_as(C.POINTER(C.c_uint8)),
                           ids.ctypes.data_as(C.POINTER(C.c_uint64)),
                           col.ctypes.data_as(C.POINTER(C.c_float)))

    def state(self):
        sc, nid, ne, fo = C.c_uint64(), C.c_uint64(), C.c_uint32(), C.c_uint32()
        t = C.c_double()
        hd, fs = C.c_int64(), C.c_int64()
        lib().fx_sim_state(self.h, C.byref(sc), C.byref(t), C.byref(nid), C.byref(ne),
                           C.byref(hd), C.byref(fs), C.byref(fo))
        return dict(step_count=sc.value, time=t.value, next_id=nid.value, n_events=ne.value,
                    hardening=hd.value, fracture_strain=fs.value, fracture_on=fo.value)

    def events(self):
        ne = self.state()["n_events"]
        t = np.zeros(max(ne, 1), np.float64)
        pos = np.zeros((max(ne, 1), 3), np.float32)
        st = np.zeros(max(ne, 1), np.float32)
        fr = np.zeros(max(ne, 1), np.uint32)
        lib().fx_sim_events(self.h, t.ctypes.data_as(C.POINTER(C.c_double)),
                            pos.ctypes.data_as(C.POINTER(C.c_float)),
                            st.ctypes.data_as(C.POINTER(C.c_float)),
                            fr.ctypes.data_as(C.POINTER(C.c_uint32)))
        return dict(time=t[:ne], pos=pos[:ne], strain=st[:ne], frags=fr[:ne])

    def diag(self):
        out = np.zeros(3, np.int64)
        lib().fx_sim_diag(self.h, out.ctypes.data_as(C.POINTER(C.c_int64)))
        return dict(total_mass=int(out[0]), kinetic_energy=int(out[1]), strain_energy=int(out[2]))

    STATS = ("active_particles", "fractured_count", "total_mass", "average_temperature",
             "kinetic_energy", "total_strain_energy")

    def stats(self):
        out = np.zeros(6, np.int64)
        lib().fx_sim_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_int64)))
        return {k: int(v) for k, v in zip(self.STATS, out)}

    def snapshots(self):
        n = self.count()
        f = np.zeros((max(n, 1), 13), np.float32)
        fr = np.zeros(max(n, 1), np.uint8)
        k = lib().fx_sim_snapshots(self.h, f.ctypes.data_as(C.POINTER(C.c_float)),
                                   fr.ctypes.data_as(C.POINTER(C.c_uint8)))
        return f[:k], fr[:k]

    def fracture_only(self):
        lib().fx_sim_fracture_only(self.h)

    def scatter_bound(self):
        """Largest per-substep sum of (|c| >> 24) + 1 over the nonzero grid
        scatter contributions c."""
        return int(lib().fx_sim_scatter_bound(self.h))

    def active_count(self):
        return int(self.read()["flags"][:, 0].sum())
