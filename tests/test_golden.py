"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces every fixture bit-exactly (single-threaded, the
reference's serial summation order).  GPU: the HIP path reproduces the
Q32.32 fixtures bit-exactly, and the FP32 fixture bit-exactly in the
reference-order FP32 mode and within FP32 summation-order noise in the
default FP32 mode, reading only the committed vectors.
"""
import glob
import os

import numpy as np
import pytest

import aemh
from oracle import refcpu as R

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# the MPM-step fixtures of make_golden.py (fixed_*.npz belong to the
# FixedPoint solver and are checked in test_fixed_sim.py)
FIXTURES = sorted(p for p in glob.glob(os.path.join(HERE, "*.npz"))
                  if not os.path.basename(p).startswith("fixed_"))
OUT = ("pos", "vel", "C", "F", "J", "Jp")


def load(path):
    with np.load(path, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def test_fixtures_present():
    names = {os.path.basename(p)[:-4] for p in FIXTURES}
    assert {"cube512_nh_q32_10", "cube512_nh_f32_10", "snow512_q32_20", "velonly512_q32_10",
            "snow_ill512_q32_5"} <= names


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_fixture(path):
    d = load(path)
    nx, ny, nz = (int(v) for v in d["grid_n"])
    cfg = R.make_config(nx, ny, nz, float(d["dx"]), tuple(d["origin"]), 1, tuple(d["gravity"]),
                        0.995, int(d["bc_margin"]), 0.0, int(d["accum"]), R.CLAMP_AUTO, threads=1)
    M = R.Materials(d["params"], d["models"], d["ext"])
    st = R.State(d["pos0"], vel=d["vel0"], F9=d["F0"], J=d["J0"], mass=d["mass"], mat=d["mat"])
    for _ in range(int(d["steps"])):
        if int(d["velocity_only"]):
            R.step_velocity_only(cfg, st, float(d["dt"]))
        else:
            R.step(cfg, M, st, float(d["dt"]))
    for f in OUT:
        got = getattr(st, f)
        assert np.array_equal(got, d[f].reshape(got.shape)), f


def _gpu_run(d, accum=None):
    params, models, ext = d["params"], d["models"], d["ext"]
    mats = [aemh.Material(*[float(x) for x in params[i]], model=int(models[i]),
                          theta_c=float(ext[i][0]), theta_s=float(ext[i][1]), xi=float(ext[i][2]))
            for i in range(params.shape[0])]
    cfg = aemh.default_config()
    cfg.substeps = 1
    cfg.gravity[:] = tuple(float(g) for g in d["gravity"])
    cfg.damping = 0.995
    cfg.bc_margin = int(d["bc_margin"])
    cfg.accum = aemh.ACCUM_Q32_32 if int(d["accum"]) == R.ACCUM_Q32 else aemh.ACCUM_F32
    if accum is not None:
        cfg.accum = accum
    eng = aemh.Engine()
    eng.set_config(cfg)
    nx, ny, nz = (int(v) for v in d["grid_n"])
    eng.set_grid(nx, ny, nz, float(d["dx"]), tuple(float(o) for o in d["origin"]))
    eng.set_materials(mats)
    eng.upload(d["pos0"], vel=d["vel0"], F9=d["F0"], J=d["J0"], mass=d["mass"], mat=d["mat"])
    for _ in range(int(d["steps"])):
        if int(d["velocity_only"]):
            eng.step_velocity_only(float(d["dt"]))
        else:
            eng.step(float(d["dt"]))
    out = eng.download()
    eng.close()
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_gpu_reproduces_fixture(path):
    d = load(path)
    got = _gpu_run(d)
    if int(d["accum"]) == R.ACCUM_Q32:
        for f in OUT:
            want = d[f]
            g = got[f].reshape(want.shape)
            assert np.array_equal(g, want), f"{f}: {int((g != want).sum())} values differ"
    else:
        # the reference's serial order: bit for bit
        ser = _gpu_run(d, aemh.ACCUM_F32_SERIAL)
        for f in OUT:
            want = d[f]
            g = ser[f].reshape(want.shape)
            assert np.array_equal(g, want), f"serial {f}: {int((g != want).sum())} values differ"
        # default FP32: only the P2G summation order differs; 10 steps of the cube
        for f in ("pos", "vel", "F"):
            want = d[f].astype(np.float64)
            err = np.linalg.norm(got[f].reshape(want.shape) - want) / np.linalg.norm(want)
            assert err < 1e-5, (f, err)
