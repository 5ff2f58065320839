"""GPU parity: the HIP step (through the C ABI) against the oracle on the same
seeded inputs.

Deterministic mode (Q32.32 accumulation, reference evaluation order, portable
ln/exp) must be bit-exact for whole trajectories.  FP32 mode differs from the
oracle only by the order of the P2G float sums; its bound is stated per test.
"""
import numpy as np
import pytest

import aemh
from aemh import workloads as W
from oracle import refcpu as R

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "C", "F", "J", "Jp")


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_cfg(wl, accum=None, clamp=R.CLAMP_AUTO, substeps=1, damping=0.995):
    nx, ny, nz, dx, o = wl.grid
    acc = wl.accum if accum is None else accum
    return R.make_config(nx, ny, nz, dx, o, substeps, wl.gravity, damping, wl.bc_margin, 0.0,
                         R.ACCUM_Q32 if acc == aemh.ACCUM_Q32_32 else R.ACCUM_F32, clamp)


def oracle_mats(mats):
    p, m, e = aemh.material_arrays(mats)
    return R.Materials(p, m, e)


def oracle_state(wl):
    return R.State(wl.pos, mass=wl.mass, mat=wl.mat)


def gpu_engine(wl):
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    return eng


def assert_bitexact(d, st, fields=FIELDS):
    for f in fields:
        got = d[f] if f != "pos" else d["pos"]
        want = getattr(st, f if f != "pos" else "pos")
        if f == "vel":
            want = st.vel
        assert np.array_equal(got.reshape(want.shape), want), f"{f}: {int((got.reshape(want.shape) != want).sum())} values differ"


def state_fields(st):
    return dict(pos=st.pos, vel=st.vel, C=st.C, F=st.F, J=st.J, Jp=st.Jp)


# ---------------------------------------------------------------------------
# Deterministic mode: bit-exact trajectories
# ---------------------------------------------------------------------------
def test_c1_deterministic_bitexact_100_steps():
    wl = W.c1_elastic_cube(steps=100)
    wl.accum = aemh.ACCUM_Q32_32
    eng = gpu_engine(wl)
    st = oracle_state(wl)
    cfg = oracle_cfg(wl)
    M = oracle_mats(wl.materials)
    for s in range(100):
        eng.step(wl.dt)
        g = R.step(cfg, M, st, wl.dt, want_grid=(s == 99))
    assert np.array_equal(eng.download_grid_q32(0, 32 ** 3), g["q"])
    assert_bitexact(eng.download(), st)


def _all_models_workload(n=6000, accum=aemh.ACCUM_Q32_32):
    rng = np.random.default_rng(11)
    pos = rng.uniform(0.25, 0.75, (n, 3)).astype(np.float32)
    mats = [
        aemh.Material(1e4, 0.2, 1e3, 0.01, 10.0, 35.0, 0.0, aemh.NEO_HOOKEAN),
        aemh.Material(1e4, 0.2, 1e3, 0.005, 16.0, 35.0, 10.0, aemh.DRUCKER_PRAGER),
        aemh.Material(2e4, 0.3, 50.0, 0.01, 78.0, 0.0, 0.0, aemh.VON_MISES),
        aemh.Material(720.0, 0.2, 0.0, 0.0, 1.0, 0.0, 0.0, aemh.NEWTONIAN_FLUID),
        aemh.Material(1e4, 0.25, 0.0, 0.0, 20.0, 0.0, 0.0, aemh.COROTATED),
        aemh.Material(1e4, 0.3, 0.0, 0.0, 20.0, 0.0, 0.0, aemh.FIXED_COROTATED),
        aemh.Material(1e4, 0.3, 0.0, 0.0, 20.0, 0.0, 0.0, aemh.RIGID),
        W.snow_material(),
    ]
    # material index 9 is out of range: falls back to material_table[0] as NeoHookean
    mat = rng.integers(0, 10, n).astype(np.uint16)
    rho = np.array([m.density for m in mats] + [mats[0].density] * 2)
    mass = (rho[mat] * 0.125 / n).astype(np.float32)
    wl = W.Workload("all_models", pos, mass, mat, mats, (48, 48, 48, 1 / 48, (0.0, 0.0, 0.0)),
                    5e-5, 10, accum=accum)
    F = np.tile(np.eye(3).reshape(9), (n, 1)) + rng.normal(0, 0.02, (n, 9))
    vel = rng.normal(0, 0.05, (n, 3))
    return wl, F.astype(np.float32), vel.astype(np.float32)


def test_all_models_deterministic_bitexact():
    wl, F, vel = _all_models_workload()
    J = np.linalg.det(F.reshape(-1, 3, 3)).astype(np.float32)
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    eng.upload(wl.pos, vel=vel, F9=F, J=J, mass=wl.mass, mat=wl.mat)
    st = R.State(wl.pos, vel=vel, F9=F, J=J, mass=wl.mass, mat=wl.mat)
    cfg = oracle_cfg(wl)
    M = oracle_mats(wl.materials)
    for _ in range(10):
        eng.step(wl.dt)
        R.step(cfg, M, st, wl.dt)
    assert_bitexact(eng.download(), st)


def test_snow_scaled_dam_deterministic_bitexact():
    wl = W.c2_snow_dam(n=1 << 16, steps=20, accum=aemh.ACCUM_Q32_32, res=64)
    eng = gpu_engine(wl)
    st = oracle_state(wl)
    cfg = oracle_cfg(wl)
    M = oracle_mats(wl.materials)
    for _ in range(20):
        eng.step(wl.dt)
        R.step(cfg, M, st, wl.dt)
    d = eng.download()
    assert_bitexact(d, st)
    assert np.any(d["Jp"] != 1.0)  # the snow clamp was exercised


def test_three_boxes_scaled_deterministic_bitexact():
    """The 16M three-material workload (fluid, jelly, snow falling onto the
    floor) at its particle density, scaled: 196,608 particles on 64^3, 6
    steps; the step picks its launch shape with the autotuner first, as the
    benchmark does."""
    wl = W.c3_three_boxes(n=3 * 65536, steps=6, res=64)
    wl.accum = aemh.ACCUM_Q32_32
    eng = gpu_engine(wl)
    eng.autotune()
    cfg = oracle_cfg(wl)
    cfg_threads = R.make_config(*wl.grid[:4], wl.grid[4], 1, wl.gravity, 0.995, wl.bc_margin, 0.0,
                                R.ACCUM_Q32, R.CLAMP_AUTO, threads=8)
    st = oracle_state(wl)
    M = oracle_mats(wl.materials)
    for k in range(6):
        eng.step(wl.dt)
        # Q32.32 sums do not depend on order: the threaded oracle is exact too
        R.step(cfg_threads if k % 2 else cfg, M, st, wl.dt)
    d = eng.download()
    assert_bitexact(d, st)


def test_velocity_only_deterministic_bitexact():
    wl = W.c1_elastic_cube(n=20000)
    wl.accum = aemh.ACCUM_Q32_32
    rng = np.random.default_rng(12)
    vel = rng.normal(0, 0.5, (wl.n, 3)).astype(np.float32)
    eng = gpu_engine(wl)
    eng.upload(wl.pos, vel=vel, mass=wl.mass, mat=wl.mat)
    st = R.State(wl.pos, vel=vel, mass=wl.mass, mat=wl.mat)
    cfg = oracle_cfg(wl)
    for _ in range(10):
        eng.step_velocity_only(wl.dt * 12)
        R.step_velocity_only(cfg, st, wl.dt * 12)
    d = eng.download()
    assert np.array_equal(d["pos"], st.pos) and np.array_equal(d["vel"], st.vel)
    assert np.array_equal(d["F"], st.F) and np.array_equal(d["C"], st.C)


def test_moving_window_origin_and_window_clamp():
    wl = W.c1_elastic_cube(n=4000)
    wl.accum = aemh.ACCUM_Q32_32
    origin = (-0.25, 0.0, 0.125)
    nx, ny, nz, dx, _ = wl.grid
    wl.grid = (nx, ny, nz, dx, origin)
    rng = np.random.default_rng(13)
    vel = rng.normal(0, 0.3, (wl.n, 3)).astype(np.float32)
    vel[:200] += np.float32([-40.0, 0.0, 40.0])   # these reach the x/z clamps
    for clamp in (aemh.CLAMP_SERIAL, aemh.CLAMP_WINDOW):
        eng = aemh.Engine()
        cfg = wl.config()
        cfg.clamp_mode = clamp
        eng.set_config(cfg)
        eng.set_grid(nx, ny, nz, dx, origin)
        eng.set_materials(wl.materials)
        eng.upload(wl.pos, vel=vel, mass=wl.mass, mat=wl.mat)
        st = R.State(wl.pos, vel=vel, mass=wl.mass, mat=wl.mat)
        ocfg = oracle_cfg(wl, clamp=R.CLAMP_SERIAL if clamp == aemh.CLAMP_SERIAL else R.CLAMP_WINDOW)
        M = oracle_mats(wl.materials)
        for _ in range(30):
            eng.step(wl.dt * 2)
            R.step(ocfg, M, st, wl.dt * 2)
        assert_bitexact(eng.download(), st)


def test_index_subset_and_reference_entry_point():
    wl = W.c1_elastic_cube(n=6000)
    wl.accum = aemh.ACCUM_Q32_32
    rng = np.random.default_rng(14)
    idx = rng.permutation(6000)[:4000].astype(np.uint32)
    eng = gpu_engine(wl)
    st = oracle_state(wl)
    cfg = oracle_cfg(wl)
    M = oracle_mats(wl.materials)
    for s in range(6):
        use = idx if s < 4 else idx[:3000]  # list changes: active set re-gathered
        eng.mpm_step_parallel_with_grid(use, wl.dt)
        R.step(cfg, M, st, wl.dt, idx=use)
    d = eng.download()
    assert_bitexact(d, st)
    untouched = np.setdiff1d(np.arange(6000), idx)
    assert np.array_equal(d["pos"][untouched], wl.pos[untouched])
    eng.mpm_step_parallel_with_grid(np.zeros(0, np.uint32), wl.dt)  # empty list: no-op
    assert_bitexact(eng.download(), st)
    with pytest.raises(aemh.AemhError):
        eng.set_indices(np.array([1, 2, 2], np.uint32))
    with pytest.raises(aemh.AemhError):
        eng.set_indices(np.array([6000], np.uint32))


def test_guards_nan_outside_and_invalid_f_bitexact():
    wl = W.c1_elastic_cube(n=3000)
    wl.accum = aemh.ACCUM_Q32_32
    pos = wl.pos.copy()
    vel = np.zeros_like(pos)
    F = np.tile(np.eye(3, dtype=np.float32).reshape(9), (wl.n, 1))
    pos[0] = [np.nan, 0.4, 0.4]
    pos[1] = [0.4, np.inf, 0.4]
    pos[2] = [-0.2, 0.4, 0.4]        # outside, negative side: base clamped to 0
    pos[3] = [1.5, 0.4, 0.4]         # outside, beyond the grid
    pos[4] = [0.4, 2.0, 0.4]         # above the grid (no upper y clamp)
    pos[5] = [-1e10, 0.5, 0.5]
    vel[6] = [np.nan, 0, 0]
    vel[7] = [80.0, -70.0, 0.0]      # beyond the +-50 clamp
    F[8, 1] = 2e3                    # |F| > 1e3 -> reset
    F[9] *= 3.0                      # det 27 -> J clamped to 10
    eng = gpu_engine(wl)
    eng.upload(pos, vel=vel, F9=F, mass=wl.mass, mat=wl.mat)
    st = R.State(pos, vel=vel, F9=F, mass=wl.mass, mat=wl.mat)
    cfg = oracle_cfg(wl)
    M = oracle_mats(wl.materials)
    for _ in range(5):
        eng.step(wl.dt)
        R.step(cfg, M, st, wl.dt)
    assert_bitexact(eng.download(), st)


@pytest.mark.parametrize("case", ["sandpile", "water", "steel"])
def test_reference_behaviour_tests_on_gpu(case):
    """The reference's own mpm_solver tests (mpm_solver.rs:1612-1745) through
    the GPU in deterministic mode, bit-exact against the oracle, plus their
    assertions."""
    if case == "sandpile":
        pos = np.array([[x * .05, y * .05 + .5, z * .05] for x in range(10) for y in range(10)
                        for z in range(10)], np.float32)
        mats = [aemh.Material(1.0e4, 0.2, 1.0e3, 0.005, 1600.0, 35.0, 0.0, aemh.DRUCKER_PRAGER)]
        grid, frames, J0 = (16, 16, 16, 1.6), 60, None
    elif case == "water":
        pos = np.array([[x * .05, y * .05 + .1, z * .05] for y in range(10) for x in range(5)
                        for z in range(5)], np.float32)
        mats = [aemh.Material(1.0e3, 0.499, 0.0, 0.0, 1000.0, 0.0, 0.0, aemh.NEWTONIAN_FLUID)]
        grid, frames, J0 = (32, 32, 32, 0.05), 1, 0.95
    else:
        pos = np.array([[0.2, 0.5, 0.2]], np.float32)
        mats = [aemh.Material(1.0e5, 0.3, 1.0e4, 0.01, 7850.0, 0.0, 0.0, aemh.VON_MISES)]
        grid, frames, J0 = (32, 32, 32, 0.02), 30, None
    n = pos.shape[0]
    mass = np.full(n, 0.01 if case == "steel" else 0.001, np.float32)
    J = None if J0 is None else np.full(n, J0, np.float32)
    eng = aemh.Engine()
    cfg = aemh.default_config()
    cfg.substeps = 8
    # The sandpile packs 1,000 particles into one 1.6 m cell: its FixedPoint
    # sums saturate from the second frame on (the oracle counts ~2e4
    # saturating adds), which makes a Q32.32 result depend on summation order.
    # There the deterministic mode must refuse (AEMH_ERR_OVERFLOW) and the
    # reference's own assertion is checked in FP32 mode.
    det = case != "sandpile"
    cfg.accum = aemh.ACCUM_Q32_32 if det else aemh.ACCUM_F32
    cfg.clamp_mode = aemh.CLAMP_SERIAL   # mpm_step -> mpm_substep clamp
    eng.set_config(cfg)
    eng.set_grid(*grid)
    eng.set_materials(mats)
    eng.upload(pos, J=J, mass=mass)
    p, m, e = aemh.material_arrays(mats)
    st = R.State(pos, J=J, mass=mass)
    ocfg = R.make_config(*grid, substeps=8, accum=R.ACCUM_Q32 if det else R.ACCUM_F32,
                         clamp_mode=R.CLAMP_SERIAL)
    for _ in range(frames):
        eng.step(0.016)
        R.step(ocfg, R.Materials(p, m, e), st, 0.016)
    d = eng.download()
    if det:
        assert_bitexact(d, st)
    else:
        assert rel(d["pos"], st.pos) < 1e-2
        qe = aemh.Engine()
        cfg.accum = aemh.ACCUM_Q32_32
        qe.set_config(cfg)
        qe.set_grid(*grid)
        qe.set_materials(mats)
        qe.upload(pos, J=J, mass=mass)
        with pytest.raises(aemh.AemhError) as ei:
            for _ in range(3):
                qe.step(0.016)
        assert ei.value.status == aemh.ERR_OVERFLOW
    if case == "sandpile":
        assert np.all(d["pos"][:, 1] >= 0)
    elif case == "water":
        assert np.all(np.isfinite(d["J"])) and np.all((d["J"] > 0) & (d["J"] < 100))
    else:
        assert 0.0 <= d["pos"][0, 1] < 0.5


def test_batched_scenes_match_plain_runs_bitexact():
    """Scenes of different sizes, grids, origins, dt, and one velocity-only
    tier, advanced together: every scene equals its own plain run."""
    rng = np.random.default_rng(15)
    specs = [(3000, 32, (0, 0, 0), 1e-4, 0), (9000, 48, (0.1, 0, -0.1), 5e-5, 0),
             (500, 16, (0, 0, 0), 2e-4, 0), (7000, 40, (0, 0.05, 0), 6e-4, 1),
             (4000, 36, (0.2, 0.0, 0.3), 1e-4, 0)]
    mats = [aemh.Material(1e4, 0.2, 0, 0, 10.0, 0, 0, aemh.NEO_HOOKEAN), W.snow_material()]
    pos_l, mass_l, mat_l, vel_l, scenes = [], [], [], [], []
    first = 0
    for k, (n, res, o, dt, vo) in enumerate(specs):
        lo = np.array(o) + 0.3
        pos_l.append(aemh.uniform_box(100 + k, n, lo, lo + 0.3))
        mass_l.append(np.full(n, 10.0 * 0.027 / n, np.float32))
        mat_l.append(np.full(n, k % 2, np.uint16))
        vel_l.append(rng.normal(0, 0.2, (n, 3)).astype(np.float32))
        sc = aemh.Scene()
        sc.first, sc.count = first, n
        sc.nx = sc.ny = sc.nz = res
        sc.dx = 1.0 / res
        sc.origin[:] = o
        sc.dt = dt
        sc.velocity_only = vo
        sc.clamp_mode = aemh.CLAMP_AUTO
        scenes.append(sc)
        first += n
    pos, mass, mat, vel = map(np.concatenate, (pos_l, mass_l, mat_l, vel_l))
    eng = aemh.Engine()
    cfg = aemh.default_config()
    cfg.accum = aemh.ACCUM_Q32_32
    cfg.bc_margin = 3
    eng.set_config(cfg)
    eng.set_grid(32, 32, 32, 1 / 32)
    eng.set_materials(mats)
    eng.upload(pos, vel=vel, mass=mass, mat=mat)
    eng.set_scenes(scenes)
    eng.batch_step(7)
    d = eng.download()
    p, m, e = aemh.material_arrays(mats)
    M = R.Materials(p, m, e)
    for k, (n, res, o, dt, vo) in enumerate(specs):
        sl = slice(scenes[k].first, scenes[k].first + n)
        st = R.State(pos[sl], vel=vel[sl], mass=mass[sl], mat=mat[sl])
        oc = R.make_config(res, res, res, np.float32(1.0 / res), o, 1, (0, -9.81, 0), 0.995, 3, 0.0,
                           R.ACCUM_Q32, R.CLAMP_AUTO)
        for _ in range(7):
            if vo:
                R.step_velocity_only(oc, st, dt)
            else:
                R.step(oc, M, st, dt)
        for f in ("pos", "vel", "C", "F", "J"):
            assert np.array_equal(d[f][sl], getattr(st, f)), (k, f)


def test_q32_overflow_is_reported():
    wl = W.c1_elastic_cube(n=2000)
    wl.accum = aemh.ACCUM_Q32_32
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    eng.upload(wl.pos, mass=np.full(wl.n, 1e9, np.float32), mat=wl.mat)
    with pytest.raises(aemh.AemhError) as ei:
        eng.step(wl.dt)
    assert ei.value.status == aemh.ERR_OVERFLOW


def test_graph_replays_with_eager_work_between():
    """Substeps replay a captured hipGraph; eager kernels and host syncs in
    between must not change any replay (deterministic mode, bit-exact)."""
    wl = W.c1_elastic_cube(n=5000)
    wl.accum = aemh.ACCUM_Q32_32
    eng = gpu_engine(wl)
    st = oracle_state(wl)
    cfg = oracle_cfg(wl)
    M = oracle_mats(wl.materials)
    for s in range(6):
        eng.step(wl.dt)
        R.step(cfg, M, st, wl.dt)
        cons = eng.conservation_f64()
        assert abs(cons[0] - float(np.sum(wl.mass, dtype=np.float64))) < 1e-12
        assert_bitexact(eng.download(), st)


# ---------------------------------------------------------------------------
# FP32 mode
# ---------------------------------------------------------------------------
def test_c1_fp32_one_step_grid_and_particles():
    """One substep.  The FP32 mode accumulates the momentum after the force
    impulse, p + f dt, per contribution (its force plane reads zero); grid
    sums differ from the oracle's by float summation order and that regrouping."""
    wl = W.c1_elastic_cube()
    eng = gpu_engine(wl)
    st = oracle_state(wl)
    g = R.step(oracle_cfg(wl), oracle_mats(wl.materials), st, wl.dt, want_grid=True)
    eng.step(wl.dt)
    gg = eng.download_grid(0, 32 ** 3)
    impulse = g["p"].astype(np.float64) + g["f"].astype(np.float64) * np.float32(wl.dt)
    assert rel(gg["m"], g["m"]) < 1e-6
    assert rel(gg["p"], impulse) < 1e-6
    assert np.all(gg["f"] == 0)
    assert rel(gg["v"], g["v"]) < 1e-6
    d = eng.download()
    assert rel(d["pos"], st.pos) < 1e-7
    assert rel(d["vel"], st.vel) < 1e-6
    assert rel(d["C"], st.C) < 1e-5
    assert rel(d["F"], st.F) < 1e-7


def _fp32_vs_oracle(wl, steps, n_floor=2, check_at=None):
    """GPU FP32 trajectory, the serial oracle, and n_floor oracle reruns whose
    P2G summation order is reshuffled every step (the GPU's float sums have no
    fixed order either).  Returns (gpu, oracle, reruns, x0)."""
    eng = gpu_engine(wl)
    base = oracle_state(wl)
    reruns = [oracle_state(wl) for _ in range(n_floor)]
    cfg = oracle_cfg(wl)
    M = oracle_mats(wl.materials)
    rngs = [np.random.default_rng(100 + k) for k in range(n_floor)]
    early = None
    for s in range(steps):
        eng.step(wl.dt)
        R.step(cfg, M, base, wl.dt)
        for st, rng in zip(reruns, rngs):
            R.step(cfg, M, st, wl.dt, idx=rng.permutation(wl.n).astype(np.uint32))
        if check_at is not None and s == check_at:
            early = (eng.download(), base.vel.copy(), [st.vel.copy() for st in reruns])
    return eng.download(), base, reruns, wl.pos.astype(np.float64), early


def _floor(reruns, base, field, x0=None):
    if x0 is None:
        return max(rel(getattr(st, field), getattr(base, field)) for st in reruns)
    return max(rel(st.pos - x0, base.pos - x0) for st in reruns)


def test_c1_fp32_100_steps_within_reference_reorder_noise():
    """100 FP32 steps of the 8,192-particle cube in the default FP32 mode,
    whose float sums have no fixed order (the reference's own >= 50,000-index
    path, Q12).  The reference's serial order is the AEMH_ACCUM_F32_SERIAL
    mode, bit-exact in tests/test_gpu_serial.py.

    Bound: positions rel-L2 <= 1e-4.  Velocities and displacements: the
    reference algorithm itself moves by more than 1e-4 under a change of the
    P2G summation order (its own rayon path sums per-thread grids; here the
    oracle reruns with the order reshuffled every step measure that floor),
    so the GPU must stay within 3x that floor or 1e-4, whichever is larger.
    After 10 steps, before the cube's dynamics amplify rounding, 1e-5."""
    wl = W.c1_elastic_cube(steps=100)
    d, base, reruns, x0, early = _fp32_vs_oracle(wl, 100, check_at=9)
    assert rel(early[0]["vel"], early[1]) < 1e-5
    floor_v = _floor(reruns, base, "vel")
    floor_dx = _floor(reruns, base, "pos", x0)
    print(f"C1 FP32 (atomic order), 100 steps vs the serial oracle: rel-L2 x {rel(d['pos'], base.pos):.3e}, "
          f"v {rel(d['vel'], base.vel):.3e} (reorder floor {floor_v:.3e}), "
          f"x - x0 {rel(d['pos'] - x0, base.pos - x0):.3e} (floor {floor_dx:.3e}); "
          f"the reference-order mode meets 1e-4 bit for bit (tests/test_gpu_serial.py)")
    assert rel(d["pos"], base.pos) < 1e-4
    assert rel(d["vel"], base.vel) <= max(1e-4, 3 * floor_v), (rel(d["vel"], base.vel), floor_v)
    assert rel(d["pos"] - x0, base.pos - x0) <= max(1e-4, 3 * floor_dx)


def test_snow_fp32_short_horizon():
    """Snow dam, 65,536 particles on 64^3, 10 FP32 steps: same criterion."""
    wl = W.c2_snow_dam(n=1 << 16, steps=10, res=64)
    d, base, reruns, x0, _ = _fp32_vs_oracle(wl, 10)
    assert rel(d["pos"], base.pos) < 1e-6
    for f in ("vel", "F", "Jp"):
        assert rel(d[f], getattr(base, f)) <= max(1e-5, 3 * _floor(reruns, base, f)), f


def test_conservation_f64_matches_numpy():
    wl = W.c1_elastic_cube()
    eng = gpu_engine(wl)
    for _ in range(3):
        eng.step(wl.dt)
    d = eng.download()
    out = eng.conservation_f64()
    m = wl.mass.astype(np.float64)
    v = d["vel"].astype(np.float64)
    assert abs(out[0] - m.sum()) <= 1e-12 * m.sum()
    assert np.allclose(out[1:4], (m[:, None] * v).sum(0), rtol=1e-9, atol=1e-18)
    assert abs(out[4] - 0.5 * (m * (v * v).sum(1)).sum()) <= 1e-9 * abs(out[4]) + 1e-20


def test_fp32_nan_velocity_next_to_zero_weight_node():
    """A particle with a NaN velocity makes its nodes' momenta NaN.  A second
    particle whose stencil reaches one of those nodes with a weight below
    1e-10 skips it (mpm_solver.rs:313-314), so its velocity and C stay finite
    and equal the oracle's; a zero weight multiplied into the NaN would have
    zeroed them through the non-finite guard (Q7)."""
    dx = np.float32(1.0 / 32.0)
    b, c = 10, 12
    # B: x-axis fraction 0.5 + 1e-5, so its third x node (b + 2) has weight
    # 0.5 (1e-5)^2 times the y/z weights, far below 1e-10.  A: base b + 2.
    xb = (b + 0.5 + 0.5 + 1e-5) * dx
    xa = (b + 2 + 0.5 + 0.2) * dx
    yz = (c + 0.5 + 0.3) * dx
    pos = np.array([[xa, yz, yz], [xb, yz, yz], [xb - 0.3 * dx, yz, yz + 0.2 * dx]], np.float32)
    vel = np.array([[np.nan, 0.0, 0.0], [0.1, -0.2, 0.05], [0.0, 0.1, 0.0]], np.float32)
    mass = np.full(3, 1e-3, np.float32)
    mats = [aemh.Material(1e3, 0.2, 0, 0, 1.0, 0, 0, aemh.NEWTONIAN_FLUID)]
    eng = aemh.Engine()
    cfg = aemh.default_config()
    cfg.accum = aemh.ACCUM_F32
    eng.set_config(cfg)
    eng.set_grid(32, 32, 32, float(dx))
    eng.set_materials(mats)
    eng.upload(pos, vel=vel, mass=mass)
    eng.step(1e-4)
    d = eng.download()
    eng.close()
    st = R.State(pos, vel=vel, mass=mass)
    p, m, e = aemh.material_arrays(mats)
    R.step(R.make_config(32, 32, 32, float(dx)), R.Materials(p, m, e), st, 1e-4)
    # the lone NaN particle: v and C zeroed by Q7 on both sides
    assert np.all(d["vel"][0] == 0) and np.all(st.vel[0] == 0)
    for k in (1, 2):
        assert np.all(np.isfinite(d["vel"][k])) and np.any(d["vel"][k] != 0)
        assert np.allclose(d["vel"][k], st.vel[k], rtol=1e-5, atol=1e-9), (k, d["vel"][k], st.vel[k])
        assert np.allclose(d["C"][k], st.C[k], rtol=1e-4, atol=1e-6), k
