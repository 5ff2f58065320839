"""aemh_conservation_f64_noalloc (the entry point that allocates nothing)
against aemh_conservation_f64, through the C ABI: bit for bit, on two calls
in a row, the second over a smaller working set, and from a second context
on the same device (which shares the device's static scratch)."""
import numpy as np
import pytest

import aemh
from aemh import workloads as W

pytestmark = pytest.mark.gpu


def _engine(n, seed):
    wl = W.c1_elastic_cube(n=n)
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    rng = np.random.default_rng(seed)
    vel = rng.normal(0, 0.3, (n, 3)).astype(np.float32)
    eng.upload(wl.pos, vel=vel, mass=wl.mass, mat=wl.mat)
    for _ in range(3):
        eng.step(wl.dt)
    return eng, wl


def test_noalloc_conservation_matches_allocating_entry_point():
    eng, wl = _engine(8192, 1)
    a = eng.conservation_f64()
    b = eng.conservation_f64_noalloc()
    assert np.array_equal(a, b), (a, b)
    assert a[0] > 0 and a[4] > 0
    # second call, smaller input: a subset of the particles
    idx = np.arange(0, wl.n, 3, dtype=np.uint32)
    eng.mpm_step_parallel_with_grid(idx, wl.dt)
    c = eng.conservation_f64_noalloc()
    d = eng.conservation_f64()
    assert np.array_equal(c, d), (c, d)
    assert c[0] < a[0]
    # a second context on the same device shares the same scratch
    eng2, _ = _engine(1000, 2)
    e = eng2.conservation_f64_noalloc()
    f = eng2.conservation_f64()
    assert np.array_equal(e, f), (e, f)
    g = eng.conservation_f64_noalloc()
    assert np.array_equal(g, d)
    eng2.close()
    eng.close()


def test_noalloc_conservation_refuses_without_working_set():
    wl = W.c1_elastic_cube(n=512)
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    with pytest.raises(aemh.AemhError) as ei:
        eng.conservation_f64_noalloc()
    assert ei.value.status == aemh.ERR_STATE
    eng.step(wl.dt)
    assert np.array_equal(eng.conservation_f64_noalloc(), eng.conservation_f64())
    eng.close()
