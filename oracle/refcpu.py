"""ctypes wrapper of the oracle (librefcpu.so built from oracle/refcpu.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker and the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# AEMH_ORACLE_REFCPU names a throw-away build with a planted mistake (only
# tests/test_oracle_pins.py sets it, to show the pins fail on it).
LIB_PATH = os.environ.get("AEMH_ORACLE_REFCPU") or os.path.join(_HERE, "librefcpu.so")

NEO_HOOKEAN, DRUCKER_PRAGER, VON_MISES, NEWTONIAN_FLUID, COROTATED, FIXED_COROTATED, RIGID, SNOW = range(8)
ACCUM_F32, ACCUM_Q32 = 0, 1
CLAMP_AUTO, CLAMP_SERIAL, CLAMP_WINDOW = 0, 1, 2


class RefConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_uint32), ("ny", C.c_uint32), ("nz", C.c_uint32),
        ("dx", C.c_float), ("origin", C.c_float * 3),
        ("substeps", C.c_uint32), ("gravity", C.c_float * 3), ("damping", C.c_float),
        ("bc_margin", C.c_uint32), ("bc_restitution", C.c_float),
        ("accum", C.c_uint32), ("clamp_mode", C.c_uint32), ("det_math", C.c_uint32),
        ("threads", C.c_uint32),
    ]


class RefMaterials(C.Structure):
    _fields_ = [("n", C.c_uint32), ("params", C.POINTER(C.c_float)),
                ("models", C.POINTER(C.c_uint8)), ("ext", C.POINTER(C.c_float))]


_lib = None


def use_library(path):
    """Load the oracle from another build of oracle/refcpu.cpp (bench.py's
    CPU baseline builds one with -march=native -flto on the host that runs it)."""
    global _lib, LIB_PATH
    LIB_PATH = path
    _lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        L.ref_b_spline.restype = C.c_float
        L.ref_b_spline.argtypes = [C.c_float]
        L.ref_b_spline_deriv.restype = C.c_float
        L.ref_b_spline_deriv.argtypes = [C.c_float]
        L.ref_fp_from_f32.restype = C.c_int64
        L.ref_fp_from_f32.argtypes = [C.c_float]
        L.ref_fp_to_f32.restype = C.c_float
        L.ref_fp_to_f32.argtypes = [C.c_int64]
        for nm in ("ref_fp_add", "ref_fp_mul", "ref_fp_div"):
            getattr(L, nm).restype = C.c_int64
            getattr(L, nm).argtypes = [C.c_int64, C.c_int64]
        L.ref_fp_sqrt.restype = C.c_int64
        L.ref_fp_sqrt.argtypes = [C.c_int64]
        L.ref_det_logf.restype = C.c_float
        L.ref_det_logf.argtypes = [C.c_float]
        L.ref_det_expf.restype = C.c_float
        L.ref_det_expf.argtypes = [C.c_float]
        L.ref_compute_stress.restype = C.c_int
        L.ref_compute_stress.argtypes = [C.c_int, fp, fp, fp, C.c_float, C.c_float, C.c_int, fp, fp]
        L.ref_svd3.restype = C.c_int
        L.ref_svd3.argtypes = [fp, fp, fp, fp]
        L.ref_svd3_d.restype = C.c_int
        L.ref_svd3_d.argtypes = [dp, dp, dp, dp]
        L.ref_snow_clamp.restype = None
        L.ref_snow_clamp.argtypes = [fp, fp, fp]
        L.ref_step_f32.restype = C.c_int
        L.ref_step_f32.argtypes = [C.POINTER(RefConfig), C.POINTER(RefMaterials), C.c_uint32,
                                   fp, fp, fp, fp, fp, fp, C.POINTER(C.c_uint16), fp,
                                   C.POINTER(C.c_uint32), C.c_uint32, C.c_float,
                                   fp, fp, fp, fp, C.POINTER(C.c_int64)]
        L.ref_step_f64.restype = C.c_int
        L.ref_step_f64.argtypes = [C.POINTER(RefConfig), C.POINTER(RefMaterials), C.c_uint32,
                                   dp, dp, dp, dp, dp, dp, C.POINTER(C.c_uint16), dp,
                                   C.POINTER(C.c_uint32), C.c_uint32, C.c_double]
        L.ref_step_velocity_only.restype = C.c_int
        L.ref_step_velocity_only.argtypes = [C.POINTER(RefConfig), C.c_uint32, fp, fp, fp,
                                             C.POINTER(C.c_uint32), C.c_uint32, C.c_float]
        _lib = L
    return _lib


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def make_config(nx, ny, nz, dx, origin=(0, 0, 0), substeps=1, gravity=(0, -9.81, 0),
                damping=0.995, bc_margin=0, bc_restitution=0.0, accum=ACCUM_F32,
                clamp_mode=CLAMP_AUTO, det_math=None, threads=1):
    cfg = RefConfig()
    cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
    cfg.dx = dx
    cfg.origin[:] = origin
    cfg.substeps = substeps
    cfg.gravity[:] = gravity
    cfg.damping = damping
    cfg.bc_margin = bc_margin
    cfg.bc_restitution = bc_restitution
    cfg.accum = accum
    cfg.clamp_mode = clamp_mode
    cfg.det_math = (1 if accum == ACCUM_Q32 else 0) if det_math is None else det_math
    cfg.threads = threads
    return cfg


class Materials:
    """Keeps the arrays alive while the C struct points at them."""

    def __init__(self, params7, models, ext=None):
        self.params = np.ascontiguousarray(params7, np.float32).reshape(-1, 7)
        self.models = np.ascontiguousarray(models, np.uint8)
        n = self.params.shape[0]
        if ext is None:
            ext = np.tile(np.array([2.5e-2, 7.5e-3, 10.0], np.float32), (n, 1))
        self.ext = np.ascontiguousarray(ext, np.float32).reshape(-1, 3)
        self.s = RefMaterials(n, _fp(self.params), self.models.ctypes.data_as(C.POINTER(C.c_uint8)),
                              _fp(self.ext))


class State:
    """MpssBuffer arrays of n particles (FP32 or FP64)."""

    def __init__(self, pos, vel=None, C9=None, F9=None, J=None, mass=None, mat=None, Jp=None,
                 dtype=np.float32):
        pos = np.array(pos, dtype).reshape(-1, 3)
        n = pos.shape[0]
        self.dtype = dtype
        self.pos = pos
        self.vel = np.zeros((n, 3), dtype) if vel is None else np.array(vel, dtype).reshape(n, 3)
        self.C = np.zeros((n, 9), dtype) if C9 is None else np.array(C9, dtype).reshape(n, 9)
        if F9 is None:
            F9 = np.tile(np.eye(3).reshape(9), (n, 1))
        self.F = np.array(F9, dtype).reshape(n, 9)
        self.J = np.ones(n, dtype) if J is None else np.array(J, dtype).reshape(n)
        self.mass = np.ones(n, dtype) if mass is None else np.array(mass, dtype).reshape(n)
        self.mat = np.zeros(n, np.uint16) if mat is None else np.array(mat, np.uint16).reshape(n)
        self.Jp = np.ones(n, dtype) if Jp is None else np.array(Jp, dtype).reshape(n)

    @property
    def n(self):
        return self.pos.shape[0]

    def copy(self, dtype=None):
        dt = dtype or self.dtype
        return State(self.pos, self.vel, self.C, self.F, self.J, self.mass, self.mat, self.Jp, dt)


def step(cfg, mats: Materials, st: State, total_dt, idx=None, want_grid=False):
    """One mpm_step_parallel_with_grid call in place on `st`; optional grid of
    the last substep (dense x-fastest)."""
    L = lib()
    n = st.n
    idx = np.arange(n, dtype=np.uint32) if idx is None else np.ascontiguousarray(idx, np.uint32)
    idxp = idx.ctypes.data_as(C.POINTER(C.c_uint32))
    if st.dtype == np.float64:
        rc = L.ref_step_f64(C.byref(cfg), C.byref(mats.s), n, _dp(st.pos), _dp(st.vel), _dp(st.F),
                            _dp(st.J), _dp(st.C), _dp(st.mass),
                            st.mat.ctypes.data_as(C.POINTER(C.c_uint16)), _dp(st.Jp), idxp,
                            idx.size, total_dt)
        if rc != 0:
            raise RuntimeError(f"ref_step_f64 failed: {rc}")
        return None
    nodes = cfg.nx * cfg.ny * cfg.nz
    g = None
    if want_grid:
        g = dict(m=np.zeros(nodes, np.float32), p=np.zeros((nodes, 3), np.float32),
                 f=np.zeros((nodes, 3), np.float32), v=np.zeros((nodes, 3), np.float32))
        if cfg.accum == ACCUM_Q32:
            g["q"] = np.zeros((nodes, 7), np.int64)
    rc = L.ref_step_f32(C.byref(cfg), C.byref(mats.s), n, _fp(st.pos), _fp(st.vel), _fp(st.F),
                        _fp(st.J), _fp(st.C), _fp(st.mass),
                        st.mat.ctypes.data_as(C.POINTER(C.c_uint16)), _fp(st.Jp), idxp, idx.size,
                        total_dt,
                        _fp(g["m"]) if g else None, _fp(g["p"]) if g else None,
                        _fp(g["f"]) if g else None, _fp(g["v"]) if g else None,
                        g["q"].ctypes.data_as(C.POINTER(C.c_int64)) if g and "q" in g else None)
    if rc != 0:
        raise RuntimeError(f"ref_step_f32 failed: {rc}")
    return g


def step_velocity_only(cfg, st: State, total_dt, idx=None):
    L = lib()
    idx = np.arange(st.n, dtype=np.uint32) if idx is None else np.ascontiguousarray(idx, np.uint32)
    rc = L.ref_step_velocity_only(C.byref(cfg), st.n, _fp(st.pos), _fp(st.vel), _fp(st.mass),
                                  idx.ctypes.data_as(C.POINTER(C.c_uint32)), idx.size, total_dt)
    if rc != 0:
        raise RuntimeError(f"ref_step_velocity_only failed: {rc}")


def compute_stress(model, params7, F, J, Jp=1.0, det_math=0, ext=(2.5e-2, 7.5e-3, 10.0)):
    p = np.ascontiguousarray(params7, np.float32)
    e = np.ascontiguousarray(ext, np.float32)
    Fa = np.ascontiguousarray(np.asarray(F, np.float32).reshape(9))
    st = np.zeros(9, np.float32)
    nf = np.zeros(9, np.float32)
    plastic = lib().ref_compute_stress(model, _fp(p), _fp(e), _fp(Fa), J, Jp, det_math, _fp(st), _fp(nf))
    return st.reshape(3, 3), nf.reshape(3, 3), bool(plastic)


def svd3(F):
    Fa = np.ascontiguousarray(np.asarray(F, np.float32).reshape(9))
    U = np.zeros(9, np.float32)
    s = np.zeros(3, np.float32)
    V = np.zeros(9, np.float32)
    used64 = lib().ref_svd3(_fp(Fa), _fp(U), _fp(s), _fp(V))
    return U.reshape(3, 3), s, V.reshape(3, 3), bool(used64)


def snow_clamp(F, Jp, theta_c=2.5e-2, theta_s=7.5e-3, xi=10.0):
    e = np.array([theta_c, theta_s, xi], np.float32)
    Fa = np.ascontiguousarray(np.asarray(F, np.float32).reshape(9)).copy()
    j = np.array([Jp], np.float32)
    lib().ref_snow_clamp(_fp(e), _fp(Fa), _fp(j))
    return Fa.reshape(3, 3), float(j[0])


def svd3_d(F):
    Fa = np.ascontiguousarray(np.asarray(F, np.float64).reshape(9))
    U = np.zeros(9)
    s = np.zeros(3)
    V = np.zeros(9)
    lib().ref_svd3_d(_dp(Fa), _dp(U), _dp(s), _dp(V))
    return U.reshape(3, 3), s, V.reshape(3, 3)
