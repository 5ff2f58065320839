"""The next substep's sort count made by G2P, which lets that substep skip
the count kernel.

Bars:
* the deterministic mode is bit-exact against the same steps with the
  count kernel run on every substep;
* the count kernel runs only when G2P could not count: on the first
  substep, and after a change between steps;
* a batch whose scene boundaries split waves between a velocity-only scene
  and full ones is covered, as is a host edit of positions between steps.

Re-applying the current launch shape drops the counts G2P made, so the
next substep runs the count kernel.
"""
import numpy as np
import pytest

import aemh
from aemh import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "C", "F", "J")


def _drop_counts(eng):
    s = eng.launch_shape()
    eng.set_launch_shape(s["p2g_tiles"], s["p2g_partials"], s["g2p_wg"], s["upd_wgs"], s["p2g_pairs"])


def _assert_same(got, want):
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f


@pytest.mark.parametrize("resort_interval", [0, 5])
def test_fused_count_q32_bitexact(resort_interval):
    """30 steps of a 32,768-particle snow dam; interval 5 adds in-place steps
    between re-sorts."""
    wl = W.c2_snow_dam(n=1 << 15, steps=30, res=48)
    wl.accum = aemh.ACCUM_Q32_32
    out = []
    for force in (False, True):
        eng = aemh.Engine()
        W.load_engine(eng, wl)
        cfg = wl.config()
        cfg.resort_interval = resort_interval
        eng.set_config(cfg)
        for _ in range(30):
            eng.step(wl.dt)
            if force:
                _drop_counts(eng)
        out.append(eng.download())
        eng.close()
    _assert_same(out[0], out[1])


def test_count_kernel_runs_only_when_g2p_did_not_count():
    wl = W.c1_elastic_cube(n=4000)
    wl.accum = aemh.ACCUM_Q32_32
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    eng.profile(True)
    for _ in range(6):
        eng.step(wl.dt)
    assert eng.profile_get("g2p")[1] == 6
    assert eng.profile_get("count")[1] == 1
    # a host edit of the positions: the next step counts afresh
    d = eng.download()
    eng.upload(d["pos"] + np.float32(0.01), vel=d["vel"], mass=wl.mass, mat=wl.mat)
    eng.step(wl.dt)
    assert eng.profile_get("count")[1] == 2
    eng.step(wl.dt)
    assert eng.profile_get("count")[1] == 2
    eng.close()


def test_fused_count_after_host_edits_q32_bitexact():
    """Positions moved across blocks by the host between steps (as the
    engine's boundary clamp and spawns do) are counted from the new values."""
    wl = W.c1_elastic_cube(n=6000)
    wl.accum = aemh.ACCUM_Q32_32
    rng = np.random.default_rng(5)
    picks = {s: rng.choice(wl.n, 500, replace=False) for s in (3, 7, 11)}
    out = []
    for force in (False, True):
        eng = aemh.Engine()
        W.load_engine(eng, wl)
        for s in range(12):
            eng.step(wl.dt)
            if s in picks:
                d = eng.download()
                pos = d["pos"].copy()
                k = picks[s]
                pos[k] = np.float32(0.5) + np.float32(0.2) * np.sin(pos[k] * np.float32(40.0))
                eng.upload(pos, vel=d["vel"], C9=d["C"], F9=d["F"], J=d["J"], mass=wl.mass, mat=wl.mat)
            if force:
                _drop_counts(eng)
        out.append(eng.download())
        eng.close()
    _assert_same(out[0], out[1])


def test_fused_count_batched_scenes_split_waves_q32_bitexact():
    """Scene boundaries at 1,000 and 2,700 particles split waves between a
    full scene, a velocity-only scene and a full scene of another grid."""
    rng = np.random.default_rng(12)
    specs = [(1000, 32, (0.0, 0.0, 0.0), 0), (1700, 24, (0.05, 0.0, 0.0), 1),
             (2300, 40, (0.0, 0.02, -0.05), 0)]
    mats = [aemh.Material(1e4, 0.2, 0, 0, 10.0, 0, 0, aemh.NEO_HOOKEAN)]
    pos_l, scenes, first = [], [], 0
    for n, res, o, vo in specs:
        lo = np.array(o) + 0.3
        pos_l.append(rng.uniform(lo, lo + 0.3, (n, 3)).astype(np.float32))
        sc = aemh.Scene()
        sc.first, sc.count = first, n
        sc.nx = sc.ny = sc.nz = res
        sc.dx = 1.0 / res
        sc.origin[:] = o
        sc.dt = 1e-4
        sc.velocity_only = vo
        sc.clamp_mode = aemh.CLAMP_AUTO
        scenes.append(sc)
        first += n
    pos = np.concatenate(pos_l)
    vel = rng.normal(0, 0.5, pos.shape).astype(np.float32)
    mass = np.full(len(pos), 1e-6, np.float32)
    mat = np.zeros(len(pos), np.uint16)
    out = []
    for force in (False, True):
        eng = aemh.Engine()
        cfg = aemh.default_config()
        cfg.accum = aemh.ACCUM_Q32_32
        cfg.gravity[:] = (0.0, -9.8, 0.0)
        eng.set_config(cfg)
        eng.set_grid(32, 32, 32, 1.0 / 32)
        eng.set_materials(mats)
        eng.upload(pos, vel=vel, mass=mass, mat=mat)
        eng.set_scenes(scenes)
        for _ in range(15):
            eng.batch_step(1)
            if force:
                _drop_counts(eng)
        eng.batch_step(4)  # four substeps in one call
        out.append(eng.download())
        eng.close()
    _assert_same(out[0], out[1])
