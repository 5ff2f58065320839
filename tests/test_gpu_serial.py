"""The reference-order FP32 mode (AEMH_ACCUM_F32_SERIAL) against the oracle's
serial FP32 substep (one thread: the path the reference takes below 50,000
indices, mpm_solver.rs:1077-1081, summing P2G in list order :831-903, with
the host C library's logf as Rust's f32::ln calls it).

Bar: bit-exact in every particle field and in the grid, for whole
trajectories.  Each test prints the rel-L2 of x and v it measured.
"""
import numpy as np
import pytest

import aemh
from aemh import workloads as W
from oracle import refcpu as R

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "C", "F", "J")


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _cfg(wl, threads=1):
    nx, ny, nz, dx, o = wl.grid
    return R.make_config(nx, ny, nz, dx, o, 1, wl.gravity, 0.995, wl.bc_margin, 0.0, R.ACCUM_F32,
                         R.CLAMP_AUTO, det_math=0, threads=threads)


def _mats(mats):
    p, m, e = aemh.material_arrays(mats)
    return R.Materials(p, m, e)


def _assert_exact(d, st, where):
    for f in FIELDS:
        want = getattr(st, f)
        got = d[f].reshape(want.shape)
        assert np.array_equal(got, want), \
            f"{where}: {f} differs in {int((got != want).sum())} values, rel-L2 {rel(got, want):.3e}"


def test_c1_serial_fp32_100_steps_bitexact():
    """The 8,192-particle Neo-Hookean cube, 100 steps: every particle field
    bit for bit, and the grid (sums after P2G, velocity after the update)
    of the last substep."""
    wl = W.c1_elastic_cube(steps=100)
    wl.accum = aemh.ACCUM_F32_SERIAL
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    st = R.State(wl.pos, mass=wl.mass, mat=wl.mat)
    cfg, M = _cfg(wl), _mats(wl.materials)
    for s in range(100):
        eng.step(wl.dt)
        g = R.step(cfg, M, st, wl.dt, want_grid=(s == 99))
    d = eng.download()
    x0 = wl.pos.astype(np.float64)
    print(f"C1 serial FP32, 100 steps: rel-L2 x {rel(d['pos'], st.pos):.3e}, "
          f"v {rel(d['vel'], st.vel):.3e}, x - x0 {rel(d['pos'] - x0, st.pos - x0):.3e}")
    _assert_exact(d, st, "C1 step 100")
    gg = eng.download_grid(0, 32 ** 3)
    for k in ("m", "p", "f", "v"):
        assert np.array_equal(gg[k], g[k]), f"grid {k}: {int((gg[k] != g[k]).sum())} values differ"
    eng.close()


def test_serial_fp32_all_reference_models_bitexact():
    """Every reference constitutive model (Neo-Hookean, Drucker-Prager, von
    Mises, fluid, both corotated, rigid) and the out-of-range material index
    falling back to entry 0 as Neo-Hookean, from a perturbed F with its
    determinant as J, 10 steps."""
    rng = np.random.default_rng(21)
    n = 6000
    pos = rng.uniform(0.25, 0.75, (n, 3)).astype(np.float32)
    mats = [
        aemh.Material(1e4, 0.2, 1e3, 0.01, 10.0, 35.0, 0.0, aemh.NEO_HOOKEAN),
        aemh.Material(1e4, 0.2, 1e3, 0.005, 16.0, 35.0, 10.0, aemh.DRUCKER_PRAGER),
        aemh.Material(2e4, 0.3, 50.0, 0.01, 78.0, 0.0, 0.0, aemh.VON_MISES),
        aemh.Material(720.0, 0.2, 0.0, 0.0, 1.0, 0.0, 0.0, aemh.NEWTONIAN_FLUID),
        aemh.Material(1e4, 0.25, 0.0, 0.0, 20.0, 0.0, 0.0, aemh.COROTATED),
        aemh.Material(1e4, 0.3, 0.0, 0.0, 20.0, 0.0, 0.0, aemh.FIXED_COROTATED),
        aemh.Material(1e4, 0.3, 0.0, 0.0, 20.0, 0.0, 0.0, aemh.RIGID),
    ]
    mat = rng.integers(0, 8, n).astype(np.uint16)  # 7: out of range
    rho = np.array([m.density for m in mats] + [mats[0].density])
    mass = (rho[mat] * 0.125 / n).astype(np.float32)
    F = (np.tile(np.eye(3).reshape(9), (n, 1)) + rng.normal(0, 0.02, (n, 9))).astype(np.float32)
    J = np.linalg.det(F.reshape(-1, 3, 3)).astype(np.float32)
    vel = rng.normal(0, 0.05, (n, 3)).astype(np.float32)
    wl = W.Workload("models", pos, mass, mat, mats, (48, 48, 48, 1 / 48, (0.0, 0.0, 0.0)), 5e-5,
                    10, accum=aemh.ACCUM_F32_SERIAL)
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    eng.upload(pos, vel=vel, F9=F, J=J, mass=mass, mat=mat)
    st = R.State(pos, vel=vel, F9=F, J=J, mass=mass, mat=mat)
    cfg, M = _cfg(wl), _mats(mats)
    for s in range(10):
        eng.step(wl.dt)
        R.step(cfg, M, st, wl.dt)
    d = eng.download()
    print(f"all models, 10 steps: rel-L2 x {rel(d['pos'], st.pos):.3e}, v {rel(d['vel'], st.vel):.3e}")
    _assert_exact(d, st, "models")
    eng.close()


def test_serial_fp32_permuted_index_list():
    """A non-ascending index list (and a shorter one later): the P2G sums run
    in list order, not slot order, as the reference's loop over
    particle_indices does."""
    wl = W.c1_elastic_cube(n=6000)
    wl.accum = aemh.ACCUM_F32_SERIAL
    rng = np.random.default_rng(22)
    idx = rng.permutation(6000)[:4500].astype(np.uint32)
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    vel = rng.normal(0, 0.3, (6000, 3)).astype(np.float32)
    eng.upload(wl.pos, vel=vel, mass=wl.mass, mat=wl.mat)
    st = R.State(wl.pos, vel=vel, mass=wl.mass, mat=wl.mat)
    cfg, M = _cfg(wl), _mats(wl.materials)
    for s in range(12):
        use = idx if s < 8 else idx[:3000]
        eng.mpm_step_parallel_with_grid(use, wl.dt)
        R.step(cfg, M, st, wl.dt, idx=use)
    _assert_exact(eng.download(), st, "permuted list")
    eng.close()


def test_serial_fp32_batched_scenes_with_velocity_only_tier():
    """Scenes of different grids, origins and dt, one of them velocity-only,
    in one batched step: each equals the serial oracle on its own."""
    rng = np.random.default_rng(23)
    specs = [(3000, 32, (0, 0, 0), 1e-4, 0), (5000, 40, (0.1, 0.0, -0.1), 5e-5, 0),
             (4000, 24, (0, 0.05, 0), 6e-4, 1)]
    mats = [aemh.Material(1e4, 0.2, 0, 0, 10.0, 0, 0, aemh.NEO_HOOKEAN)]
    pos_l, vel_l, scenes, first = [], [], [], 0
    for k, (n, res, o, dt, vo) in enumerate(specs):
        lo = np.array(o) + 0.3
        pos_l.append(aemh.uniform_box(200 + k, n, lo, lo + 0.3))
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
    pos, vel = np.concatenate(pos_l), np.concatenate(vel_l)
    mass = np.full(pos.shape[0], 10.0 * 0.027 / 3000, np.float32)
    eng = aemh.Engine()
    cfg = aemh.default_config()
    cfg.accum = aemh.ACCUM_F32_SERIAL
    cfg.bc_margin = 3
    eng.set_config(cfg)
    eng.set_grid(32, 32, 32, 1 / 32)
    eng.set_materials(mats)
    eng.upload(pos, vel=vel, mass=mass)
    eng.set_scenes(scenes)
    eng.batch_step(6)
    d = eng.download()
    M = _mats(mats)
    for k, (n, res, o, dt, vo) in enumerate(specs):
        sl = slice(scenes[k].first, scenes[k].first + n)
        st = R.State(pos[sl], vel=vel[sl], mass=mass[sl])
        oc = R.make_config(res, res, res, np.float32(1.0 / res), o, 1, (0, -9.81, 0), 0.995, 3, 0.0,
                           R.ACCUM_F32, R.CLAMP_AUTO, det_math=0)
        for _ in range(6):
            if vo:
                R.step_velocity_only(oc, st, dt)
            else:
                R.step(oc, M, st, dt)
        got = {f: d[f][sl] for f in ("pos", "vel")}
        for f in ("pos", "vel"):
            assert np.array_equal(got[f], getattr(st, f)), (k, f)
        if not vo:
            _assert_exact({f: d[f][sl] for f in FIELDS}, st, f"scene {k}")
    eng.close()
