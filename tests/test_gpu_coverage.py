"""Device paths the default configurations reach only at full size, and the
configurations' full horizons, against the oracle at scaled sizes.

* In-place G2P between re-sorts.  Below 256 MB of working set the step
  re-sorts every substep; the in-place G2P (which C3 runs at 16.8M
  particles) is forced here with aemh_config.resort_interval and compared
  bit for bit in Q32.32 mode and in the reference-order FP32 mode.
* The snow clamp's FP64 SVD (cond(F) > 1e4): ill-conditioned snow particles,
  bit for bit, with the device's count of FP64 solves.
* The configurations' own horizons: the C2 dam break (200 steps) and the C3
  three-material boxes (100 steps) at scaled particle counts, bit for bit
  against the serial FP32 oracle in the reference-order mode.
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


def _assert_exact(d, st, where):
    for f in FIELDS:
        want = getattr(st, f)
        got = d[f].reshape(want.shape)
        assert np.array_equal(got, want), \
            f"{where}: {f} differs in {int((got != want).sum())} values, rel-L2 {rel(got, want):.3e}"


def _oracle_cfg(wl, accum, det_math, threads=1):
    nx, ny, nz, dx, o = wl.grid
    return R.make_config(nx, ny, nz, dx, o, 1, wl.gravity, 0.995, wl.bc_margin, 0.0, accum,
                         R.CLAMP_AUTO, det_math=det_math, threads=threads)


def _mats(mats):
    p, m, e = aemh.material_arrays(mats)
    return R.Materials(p, m, e)


def _engine(wl, accum, resort_interval=0):
    eng = aemh.Engine()
    wl.accum = accum
    W.load_engine(eng, wl)
    cfg = wl.config()
    cfg.resort_interval = resort_interval
    eng.set_config(cfg)
    return eng


@pytest.mark.parametrize("accum", [aemh.ACCUM_Q32_32, aemh.ACCUM_F32_SERIAL], ids=["q32", "serial"])
@pytest.mark.parametrize("case", ["c2", "c3"])
def test_inplace_g2p_between_resorts_bitexact(case, accum):
    """Re-sort every 8th substep, in place in between, 24 steps."""
    if case == "c2":
        wl = W.c2_snow_dam(n=1 << 16, steps=24, res=64)
    else:
        wl = W.c3_three_boxes(n=3 * 32768, steps=24, res=64)
    eng = _engine(wl, accum, resort_interval=8)
    q32 = accum == aemh.ACCUM_Q32_32
    cfg = _oracle_cfg(wl, R.ACCUM_Q32 if q32 else R.ACCUM_F32, 1 if q32 else 2,
                      threads=8 if q32 else 1)
    M = _mats(wl.materials)
    st = R.State(wl.pos, mass=wl.mass, mat=wl.mat)
    for s in range(24):
        eng.step(wl.dt)
        R.step(cfg, M, st, wl.dt)
        if s in (3, 11):  # mid-interval: in-place steps since the last re-sort
            _assert_exact(eng.download(), st, f"{case} step {s + 1}")
    d = eng.download()
    _assert_exact(d, st, f"{case} step 24")
    if case == "c2":  # the dam break's snow has been clamped (the boxes still fall)
        assert np.any(d["Jp"] != 1.0)
    eng.close()


def test_snow_fp64_svd_fallback_counted_and_bitexact():
    """Snow particles with cond(F) 1e4, 1e6 and 1e8 among well-conditioned
    ones, Q32.32 mode, 5 steps: bit for bit with the oracle (whose SVD takes
    the same FP64 path), and the device counted its FP64 solves."""
    n = 6000
    wl = W.c2_snow_dam(n=n, steps=5, res=32)
    rng = np.random.default_rng(31)
    F = np.tile(np.eye(3), (n, 1, 1))
    ill = np.arange(0, n, 3)
    for k, i in enumerate(ill):
        c = (1e4, 1e6, 1e8)[k % 3]
        q1, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        q2, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        F[i] = q1 @ np.diag([1.5, 1.0, 1.5 / c]) @ q2.T
    F = F.reshape(n, 9).astype(np.float32)
    J = np.ones(n, np.float32)  # stored J: the stress stays bounded
    eng = _engine(wl, aemh.ACCUM_Q32_32)
    eng.upload(wl.pos, F9=F, J=J, mass=wl.mass, mat=wl.mat)
    st = R.State(wl.pos, F9=F, J=J, mass=wl.mass, mat=wl.mat)
    cfg = _oracle_cfg(wl, R.ACCUM_Q32, 1)
    M = _mats(wl.materials)
    for s in range(5):
        eng.step(wl.dt)
        R.step(cfg, M, st, wl.dt)
    d = eng.download()
    _assert_exact(d, st, "ill-conditioned snow")
    count = eng.svd_fp64_count()
    print(f"FP64 SVD solves on the device: {count} ({len(ill)} ill-conditioned particles)")
    # every particle of cond 1e6 and 1e8 goes through it on the first step
    assert count >= (2 * len(ill)) // 3
    # the clamp moved them: J_p far from 1
    assert np.all(np.abs(d["Jp"][ill[1::3]] - 1.0) > 0.5)
    eng.close()


def test_serial_fp32_c2_dam_break_200_steps_bitexact():
    """The C2 dam break (snow, dt 5e-5) at 65,536 particles on 64^3, its
    full 200 steps, in the reference-order FP32 mode."""
    wl = W.c2_snow_dam(n=1 << 16, steps=200, res=64)
    eng = _engine(wl, aemh.ACCUM_F32_SERIAL)
    cfg = _oracle_cfg(wl, R.ACCUM_F32, 2)
    M = _mats(wl.materials)
    st = R.State(wl.pos, mass=wl.mass, mat=wl.mat)
    for s in range(200):
        eng.step(wl.dt)
        R.step(cfg, M, st, wl.dt)
    d = eng.download()
    print(f"C2 shape, 200 serial FP32 steps: rel-L2 x {rel(d['pos'], st.pos):.3e}, "
          f"v {rel(d['vel'], st.vel):.3e}")
    _assert_exact(d, st, "c2 step 200")
    eng.close()


def test_serial_fp32_c3_three_boxes_100_steps_bitexact():
    """The C3 three-material boxes (fluid, jelly, snow) at 98,304 particles on
    64^3, their full 100 steps, in the reference-order FP32 mode."""
    wl = W.c3_three_boxes(n=3 * 32768, steps=100, res=64)
    eng = _engine(wl, aemh.ACCUM_F32_SERIAL)
    cfg = _oracle_cfg(wl, R.ACCUM_F32, 2)
    M = _mats(wl.materials)
    st = R.State(wl.pos, mass=wl.mass, mat=wl.mat)
    for s in range(100):
        eng.step(wl.dt)
        R.step(cfg, M, st, wl.dt)
    d = eng.download()
    print(f"C3 shape, 100 serial FP32 steps: rel-L2 x {rel(d['pos'], st.pos):.3e}, "
          f"v {rel(d['vel'], st.vel):.3e}")
    _assert_exact(d, st, "c3 step 100")
    eng.close()


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def test_fp32_snow_clamp_eigen_route_matches_svd_route():
    """The FP32 mode clamps snow through the eigen-decomposition of F^T F
    when cond(F) <= 2 and det F > 0, and through the SVD otherwise; the
    oracle always through the SVD.  One step of dt 1e-9 (F' = F to rounding,
    so only the clamp differs) from snow F with singular values spread across
    [0.85, 1.2] (at least 0.03 apart) or inside the window (eigen route),
    with det F < 0 or cond(F) 2.5 to 11 (SVD route): F within 1e-6 and J_p
    within 2e-6 (relative) of the oracle's, per particle; the SVD route's
    particles bit for bit."""
    n = 6000
    wl = W.c2_snow_dam(n=n, steps=1, res=32)
    rng = np.random.default_rng(77)
    F = np.empty((n, 3, 3))
    kind = np.arange(n) % 6
    for i in range(n):
        while True:
            s = rng.uniform(0.85, 1.2, 3)
            if np.min(np.diff(np.sort(s))) > 0.03:
                break
        if kind[i] == 0:
            s = rng.uniform(0.98, 1.005, 3)
        elif kind[i] == 1:
            s[np.argmin(s)] *= -1.0
        elif kind[i] == 2:
            s = np.array([rng.uniform(0.1, 0.4), 1.0, 1.1])
        F[i] = _rotation(rng) @ np.diag(s) @ _rotation(rng).T
    F = F.reshape(n, 9).astype(np.float32)
    J = np.ones(n, np.float32)
    dt = 1e-9
    eng = _engine(wl, aemh.ACCUM_F32)
    eng.upload(wl.pos, F9=F, J=J, mass=wl.mass, mat=wl.mat)
    st = R.State(wl.pos, F9=F, J=J, mass=wl.mass, mat=wl.mat)
    eng.step(dt)
    R.step(_oracle_cfg(wl, R.ACCUM_F32, 0), _mats(wl.materials), st, dt)
    d = eng.download()
    dF = np.abs(d["F"].reshape(n, 9) - st.F.reshape(n, 9)).max(axis=1)
    dJp = np.abs(d["Jp"] / st.Jp - 1.0)
    for k, name in ((0, "inside"), (1, "det < 0"), (2, "cond > 2"), (3, "spread")):
        sel = kind == k if k < 3 else kind >= 3
        print(f"{name:9s}: max |dF| {dF[sel].max():.3e}, max |dJp/Jp| {dJp[sel].max():.3e}, "
              f"clamped {int(np.sum(st.Jp[sel] != 1.0))} of {int(sel.sum())}")
    assert np.all(np.isfinite(d["F"])) and np.all(np.isfinite(d["Jp"]))
    assert dF.max() <= 1e-6 and dJp.max() <= 2e-6, (dF.max(), dJp.max())
    svd = (kind == 1) | (kind == 2)
    assert np.array_equal(d["F"].reshape(n, 9)[svd], st.F.reshape(n, 9)[svd])
    assert np.array_equal(d["Jp"][svd], st.Jp[svd])
    # the routes did clamp: every spread, det < 0 and ill-conditioned F
    assert np.all(st.Jp[kind >= 1] != 1.0)
    eng.close()
