"""GPU: the full-size workloads (1,048,576-particle snow dam break on 128^3,
16,777,216 particles in three boxes on a sparse 256^3 grid, 256 batched
scenes of 15.9M particles in all, and the all-integer MpmSimulation on 1M
and 16M particles), checked through properties that hold at any size.

The oracle needs minutes per step at these sizes, so these tests do not
compare against it; the bit-exact comparisons with the oracle run at scaled
sizes in test_gpu_parity.py.  What is checked here is what the deterministic
(Q32.32) mode promises whatever the particle count: its integer grid sums do
not depend on the order in which particles are stored, sorted or grouped
into workgroups, nor on which P2G path (LDS tile or direct HBM atomic) a
particle takes, so a run from a permuted upload under a different launch
shape must reproduce every field of every particle bit for bit.
"""
import numpy as np
import pytest

import aemh
from aemh import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "C", "F", "J", "Jp")


def run(wl, steps, shape=None, perm=None):
    eng = aemh.Engine()
    if perm is not None:
        wl = W.Workload(wl.name, wl.pos[perm], wl.mass[perm], wl.mat[perm], wl.materials, wl.grid,
                        wl.dt, wl.steps, wl.gravity, wl.bc_margin, wl.accum, wl.storage)
    W.load_engine(eng, wl)
    if shape is not None:
        eng.set_launch_shape(*shape)
    for _ in range(steps):
        eng.step(wl.dt)
    d = eng.download()
    eng.close()
    return d


def assert_same(a, b, perm, label):
    for f in FIELDS:
        x = np.asarray(a[f]).reshape(len(perm), -1)[perm]
        y = np.asarray(b[f]).reshape(len(perm), -1)
        nd = int((x.view(np.uint32) != y.view(np.uint32)).sum())
        assert nd == 0, f"{label} {f}: {nd} values differ"


def test_c2_full_size_q32_order_and_shape_invariant():
    wl = W.c2_snow_dam(accum=aemh.ACCUM_Q32_32)
    assert wl.n == 1 << 20
    perm = np.random.default_rng(5).permutation(wl.n)
    a = run(wl, 4, shape=(3, 1, 256, 2048))
    b = run(wl, 4, shape=(2, 0, 512, 8192), perm=perm)
    assert np.isfinite(a["pos"]).all()
    # The block moved: gravity acted on every particle (v_y < 0 away from the floor).
    assert float(np.asarray(a["vel"]).reshape(-1, 3)[:, 1].mean()) < 0.0
    assert_same(a, b, perm, "c2 1M")


def test_c3_full_size_q32_order_and_shape_invariant():
    wl = W.c3_three_boxes()
    wl.accum = aemh.ACCUM_Q32_32
    assert wl.n == 1 << 24
    perm = np.random.default_rng(6).permutation(wl.n)
    a = run(wl, 2, shape=(4, 0, 256, 8192))
    b = run(wl, 2, shape=(2, 1, 128, 4096), perm=perm)
    assert np.isfinite(a["pos"]).all()
    assert_same(a, b, perm, "c3 16.8M")


def fixed_run(sc, steps, perm=None):
    pos, mass = (sc.pos, sc.mass) if perm is None else (sc.pos[perm], sc.mass[perm])
    sim = aemh.FixedSim(sc.config)
    sim.add_particles(pos, np.zeros_like(pos), mass)
    sim.step(steps)
    d, st = sim.read(), sim.state()
    sim.close()
    return d, st


@pytest.mark.parametrize("n,res,steps", [(1 << 20, 64, 6), (1 << 24, 160, 5)])
def test_fixed_point_full_size_order_invariant(n, res, steps):
    """The all-integer MpmSimulation at 1M and 16M particles.  Its P2G adds
    saturating Q32.32 sums; while no cell can clip (the scatter bound, which
    the run reports), they are plain integer sums, so permuting the particles
    before they are added leaves every raw field of every particle unchanged.
    Both runs cross a re-sort of the cell order (every 4th substep)."""
    sc = W.fixed_cloud(n, res)
    perm = np.random.default_rng(7).permutation(n)
    a, sa = fixed_run(sc, steps)
    b, sb = fixed_run(sc, steps, perm)
    assert sa["serial_substeps"] == 0 and sb["serial_substeps"] == 0
    assert sa["step_count"] == sb["step_count"] == steps
    nd = int((a["q"][perm] != b["q"]).sum())
    assert nd == 0, f"{nd} raw values differ"
    assert np.array_equal(a["flags"][perm], b["flags"])
    # The cloud moved: positions changed from the upload.
    assert int((a["q"][:, 0:3] != sc.pos).sum()) > 0


def test_c4_full_size_batched_equals_plain_q32():
    """All 256 scenes (log-uniform 4K..256K particles, grids 32^3..128^3,
    15.9M particles) advanced together by the batched step equal, scene by
    scene, 256 plain single-scene engines run on the same particles."""
    wl = W.c4_batched_scenes()
    wl.accum = aemh.ACCUM_Q32_32
    assert len(wl.scenes) == 256
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    eng.batch_step(3)
    d = eng.download()
    eng.close()
    for k, sc in enumerate(wl.scenes):
        sl = slice(sc.first, sc.first + sc.count)
        e = aemh.Engine()
        e.set_config(wl.config())
        e.set_grid(sc.nx, sc.ny, sc.nz, sc.dx, tuple(sc.origin))
        e.set_materials(wl.materials)
        e.upload(wl.pos[sl], mass=wl.mass[sl], mat=wl.mat[sl])
        for _ in range(3):
            e.step(sc.dt)
        p = e.download()
        e.close()
        for f in FIELDS:
            assert np.array_equal(d[f][sl].view(np.uint32), p[f].view(np.uint32)), (k, f)
