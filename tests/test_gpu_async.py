"""aemh_step_async / aemh_sync: steps enqueued without waiting for the GPU.

Bars: the deterministic mode bit-exact with the same steps taken
synchronously (integer sums do not depend on the re-sort decisions the
asynchronous steps defer); the FP32 mode within float regrouping; the
deterministic mode's range check reported by aemh_sync; downloads and
synchronous calls ordered after pending steps.
"""
import numpy as np
import pytest

import aemh
from aemh import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "C", "F", "J")


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _run(wl, steps, mode):
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    for s in range(steps):
        if mode == "sync":
            eng.step(wl.dt)
        else:
            eng.step_async(wl.dt)
            if mode == "mixed" and s % 7 == 6:
                eng.sync()
    if mode != "sync":
        eng.sync()
    d = eng.download()
    eng.close()
    return d


@pytest.mark.parametrize("mode", ["async", "mixed"])
def test_async_steps_q32_bitexact(mode):
    wl = W.c2_snow_dam(n=1 << 15, steps=30, res=48)
    wl.accum = aemh.ACCUM_Q32_32
    want = _run(wl, 30, "sync")
    got = _run(wl, 30, mode)
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f


def test_async_steps_fp32_match_sync():
    wl = W.c2_snow_dam(n=1 << 15, steps=20, res=48)
    want = _run(wl, 20, "sync")
    got = _run(wl, 20, "async")
    # the runs differ only in the order of float sums (block ranks and
    # re-sort timing), 20 steps of which stay well below 1e-5
    for f in ("pos", "vel", "F"):
        assert rel(got[f], want[f]) < 1e-5, f


def test_download_after_pending_async_steps():
    """A download right after asynchronous steps returns their results."""
    wl = W.c1_elastic_cube(n=4000)
    wl.accum = aemh.ACCUM_Q32_32
    want = _run(wl, 12, "sync")
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    for _ in range(12):
        eng.step_async(wl.dt)
    got = eng.download()  # no sync() first
    eng.sync()
    eng.close()
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f


def test_sync_reports_q32_overflow_of_async_steps():
    wl = W.c1_elastic_cube(n=2000)
    wl.accum = aemh.ACCUM_Q32_32
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    eng.upload(wl.pos, mass=np.full(wl.n, 1e9, np.float32), mat=wl.mat)
    eng.step_async(wl.dt)
    with pytest.raises(aemh.AemhError) as ei:
        eng.sync()
    assert ei.value.status == aemh.ERR_OVERFLOW
    eng.close()
