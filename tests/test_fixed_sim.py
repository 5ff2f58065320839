"""All-integer MLS-MPM solver on FixedPoint Q32.32 (ae_physics/src/mpm.rs
MpmSimulation).

CPU: the oracle (oracle/fixedsim.cpp) against the reference's own unit tests
of FixedPoint (fixed_point.rs:889-962) and MpmSimulation (mpm.rs:957-1148),
and against exact closed forms for what those leave open (Python integers
restate mul / div / sqrt / the table; the first P2G deposit of a resting
particle; the free-fall velocity after one substep).
GPU: aemh_fixed_* through the C ABI against the oracle, every field of every
particle bit-exact, for each material model, substeps, added and removed
particles and fracture.
"""
import glob
import math
import os

import numpy as np
import pytest

from oracle import fixedsim as X

ONE = 1 << 32
I64MAX, I64MIN = (1 << 63) - 1, -(1 << 63)


def trunc_div(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def wrap64(v):
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >= 1 << 63 else v


# ---------------------------------------------------------------------------
# FixedPoint: the reference's unit tests (fixed_point.rs:889-962)
# ---------------------------------------------------------------------------
def f32(r):
    return X.to_f32(r)


def test_fixed_trig():
    pi = 13493037704
    half_pi = X.op("div", pi, 2 * ONE)
    assert abs(f32(X.op("sin", 0))) < 1e-3
    assert abs(f32(X.op("sin", half_pi)) - 1.0) < 1e-3
    assert abs(f32(X.op("cos", 0)) - 1.0) < 1e-3
    assert abs(f32(X.op("cos", half_pi))) < 1e-3


def test_fixed_sqrt_and_arithmetic():
    assert abs(f32(X.op("sqrt", 4 * ONE)) - 2.0) < 1e-3
    assert abs(f32(X.op("sqrt", 2 * ONE)) - 1.414) < 1e-2
    a, b = X.from_f32(3.5), X.from_f32(2.0)
    assert abs(f32(X.op("add", a, b)) - 5.5) < 1e-3
    assert abs(f32(X.op("sub", a, b)) - 1.5) < 1e-3
    assert abs(f32(X.op("mul", a, b)) - 7.0) < 1e-3
    assert abs(f32(X.op("div", a, b)) - 1.75) < 1e-3
    third = X.from_f32(1.0 / 3.0)
    assert abs(f32(X.op("mul", third, 3 * ONE)) - 1.0) < 1e-3  # determinism test
    assert X.op("add", I64MAX, ONE) == I64MAX  # overflow saturates (Add, :187-189)


# ---------------------------------------------------------------------------
# FixedPoint: exact forms
# ---------------------------------------------------------------------------
def test_fixed_mul_div_sqrt_exact():
    rng = np.random.default_rng(1)
    vals = [int(v) for v in rng.integers(-(1 << 50), 1 << 50, 200)] + [0, 1, -1, ONE, -ONE, I64MAX,
                                                                         I64MIN]
    for a in vals:
        for b in vals[:40]:
            assert X.op("mul", a, b) == wrap64(trunc_div(a * b, 1 << 32))
            want = (0 if a == 0 else (I64MAX if a > 0 else I64MIN)) if b == 0 else \
                wrap64(trunc_div(a * (1 << 32), b))
            assert X.op("div", a, b) == want
        want = 0 if a <= 0 else math.isqrt(a << 32)
        assert X.op("sqrt", a) == wrap64(want)
        assert X.op("neg", a) == (I64MAX if a == I64MIN else -a)


def test_fixed_sin_table_and_interpolation():
    # table entries: (sin(i/1024 pi/2) 2^32) truncated; within 1 raw of the
    # correctly rounded value computed with exact Python arithmetic
    for i in (0, 1, 100, 511, 512, 1000, 1023):
        want = math.sin(i / 1024 * math.pi / 2) * 2 ** 32
        assert abs(X.lib().fx_sin_table(i) - want) <= 1.0
    # between two table nodes sin interpolates linearly (fixed_point.rs:126-128)
    step = 1 << 22
    for i in (3, 200, 700):
        lo, hi = X.lib().fx_sin_table(i), X.lib().fx_sin_table(i + 1)
        ang = i * step + step // 4
        assert X.op("sin", ang) == lo + trunc_div((hi - lo) * (step // 4), step)
        assert X.op("sin", -ang) == -X.op("sin", ang)  # odd


def test_fixed_from_to_f32():
    for v in (0.0, 1.5, -2.25, 1e-9, 3.0e8, -7.0e8):
        assert X.from_f32(v) == int(np.float64(np.float32(v)) * 2 ** 32)
    assert X.from_f32(float("nan")) == 0
    assert X.from_f32(3.0e9) == I64MAX and X.from_f32(-1e30) == I64MIN
    assert X.to_f32(3 * ONE + ONE // 2) == 3.5


def test_fixed_ln_exp_against_math():
    # ln is 2 artanh((x-1)/(x+1)) cut after the t^9 term (fixed_point.rs:254-271):
    # the cut series in float, and math.log near 1 where the cut is negligible
    for v in (0.1, 0.5, 1.0, 2.0, 5.0, 40.0):
        t = (v - 1.0) / (v + 1.0)
        cut = 2.0 * sum(t ** k / k for k in (1, 3, 5, 7, 9))
        assert abs(f32(X.op("ln", X.from_f32(v))) - cut) < 1e-6 * max(1.0, abs(cut))
    for v in (0.9, 1.1, 1.5):
        assert abs(f32(X.op("ln", X.from_f32(v))) - math.log(v)) < 1e-5
    for v in (-1.0, -0.1, 0.3, 1.0):
        assert abs(f32(X.op("exp", X.from_f32(v))) - math.exp(v)) < 1e-5 * math.exp(v) + 1e-6
    assert X.op("ln", 0) == I64MIN and X.op("exp", 0) == ONE


def test_fixed_det_inverse_svd():
    m = np.array([2, 0, 0, 0, 3, 0, 0, 0, 4], np.int64) * ONE  # columns
    assert X.mat_op(0, m)[0] == 24 * ONE
    inv = X.mat_op(1, m)[:9]
    # adjugate entry times ONE / det, each product truncated (fixed_point.rs:730-762)
    inv_det = (ONE << 32) // (24 * ONE)
    assert list(inv) == [12 * inv_det, 0, 0, 0, 8 * inv_det, 0, 0, 0, 6 * inv_det]
    rng = np.random.default_rng(3)
    for _ in range(20):
        a = rng.uniform(-2, 2, (3, 3)) + 3 * np.eye(3)
        raw = np.array([X.from_f32(v) for v in a.T.reshape(9)], np.int64)  # columns
        det = X.mat_op(0, raw)[0] / 2 ** 32
        assert abs(det - np.linalg.det(a.astype(np.float32).astype(np.float64))) < 1e-6 * 64
        inv = X.mat_op(1, raw)[:9].reshape(3, 3).T / 2 ** 32
        assert np.abs(inv - np.linalg.inv(a.astype(np.float32).astype(np.float64))).max() < 1e-6
    u, s, v = X.mat_op(2, m).reshape(3, 9)
    # diagonal input: no rotation, singular values are the diagonal, and V is
    # the reference's rotated A^T A (diag of squares), not its rotations
    assert list(s[[0, 4, 8]]) == [2 * ONE, 3 * ONE, 4 * ONE]
    assert list(v[[0, 4, 8]]) == [4 * ONE, 9 * ONE, 16 * ONE]
    big = np.full(9, I64MAX, np.int64)
    assert X.mat_op(0, big)[0] == 0  # overflowing i128 pieces become 0 (:711-729)


# ---------------------------------------------------------------------------
# MpmSimulation: the reference's unit tests (mpm.rs:957-1148)
# ---------------------------------------------------------------------------
def test_sim_initialization_and_snapshots():
    s = X.Sim(X.small_config())
    d = s.read()
    assert s.count() == 8 and s.active_count() == 8
    assert (d["q"][:, 24] > 0).all() and (d["q"][:, 25] > 0).all()  # mass, volume


def test_sim_substeps_and_models_run():
    s = X.Sim(X.small_config())
    s.step(10)
    assert s.state()["step_count"] == 10 and s.active_count() == 8
    for model, param in ((X.FLUID, X.from_f32(0.01)), (X.GRANULAR, X.from_f32(30.0)),
                         (X.SNOW, X.from_f32(0.05)), (X.BRITTLE, X.from_f32(0.02))):
        t = X.Sim(X.small_config(model=model, model_param=param))
        t.step(5)
        assert t.active_count() > 0


def test_sim_elastic_kinetic_energy_rises():
    s = X.Sim(X.small_config(model=X.ELASTIC))
    e0 = s.diag()["kinetic_energy"]
    s.step(5)
    assert s.active_count() > 0 and s.diag()["kinetic_energy"] > e0


def test_sim_mass_conservation_and_fracture_enabled():
    s = X.Sim(X.small_config())
    m0 = s.diag()["total_mass"]
    s.step(5)
    assert abs(m0 - s.diag()["total_mass"]) < X.op("mul", m0, X.from_f32(0.001))
    f = X.Sim(X.small_config(enable_fracture=1, fracture_strain=X.from_f32(0.01)))
    f.step(5)
    assert f.active_count() > 0


def test_sim_add_remove_and_switch_models():
    s = X.Sim(X.small_config())
    pid = s.add_particle((X.from_f32(0.5),) * 3, (0, 0, 0), X.from_f32(0.1))
    assert s.active_count() == 9
    assert s.remove_particle(pid) and s.active_count() == 8
    assert not s.remove_particle(99999)
    for m in (X.ELASTIC, X.ELASTOPLASTIC, X.FLUID):
        s.set_model(m, X.from_f32(0.01))
        s.step()
    assert s.state()["step_count"] == 3
    z = X.Sim(X.default_config(particle_count=0, res=(4, 4, 4)))
    z.step()
    assert z.count() == 0


# ---------------------------------------------------------------------------
# MpmSimulation: closed forms
# ---------------------------------------------------------------------------
def test_sim_free_fall_one_substep_closed_form():
    """Elastic model, one particle at rest at a cell corner with F = I: stress
    is zero, P2G deposits m w per node, the grid force adds g dt m per node,
    G2P gives v = sum_w (g dt m w)/(m w)... computed here with Python
    integers from the same rules, so a dropped term or a wrong order fails."""
    cfg = X.small_config(model=X.ELASTIC, particle_count=0)
    s = X.Sim(cfg)
    pos = (X.from_f32(1.0), X.from_f32(1.25), X.from_f32(1.0))
    mass = X.from_f32(0.5)
    s.add_particle(pos, (0, 0, 0), mass)
    s.step(1)
    q = s.read()["q"][0]
    # expected, with exact integer restatements of the FixedPoint rules
    mul = lambda a, b: wrap64(trunc_div(a * b, 1 << 32))  # noqa: E731
    div = lambda a, b: wrap64(trunc_div(a << 32 if a >= 0 else -((-a) << 32), b))  # noqa: E731
    cell, dt, g = cfg.cell, cfg.dt, cfg.gravity[1]
    inv = div(ONE, cell)

    def bsp(f):
        a = abs(f)
        if a < ONE // 2:
            return (3 * ONE // 4) - mul(f, f)
        if a < 3 * ONE // 2:
            x = 3 * ONE // 2 - a
            return mul(mul(ONE // 2, x), x)
        return 0
    fx = [mul(p, inv) - (mul(p, inv) & ~0xFFFFFFFF) for p in pos]
    base = [(mul(p, inv) & ~0xFFFFFFFF) >> 32 for p in pos]
    vy = 0
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                gx, gy, gz = base[0] + dx, base[1] + dy, base[2] + dz
                if not all(0 <= c < 4 for c in (gx, gy, gz)):
                    continue
                w = mul(mul(bsp(fx[0] - dx * ONE), bsp(fx[1] - dy * ONE)), bsp(fx[2] - dz * ONE))
                if w < 10:
                    continue
                gm = mul(mass, w)
                if gm < 10:
                    continue
                gvel = div(mul(mul(g, dt), gm), gm)  # momentum 0 + g dt m, then / m
                vy += mul(gvel, w)
    assert q[4] == vy and q[3] == 0 and q[5] == 0
    assert q[1] == pos[1] + mul(vy, dt)


# ---------------------------------------------------------------------------
# Material models and fracture: invariants and closed forms on one particle
# ---------------------------------------------------------------------------
def qmul(a, b):
    return wrap64(trunc_div(a * b, 1 << 32))


def qdiv(a, b):
    return wrap64(trunc_div(a * (1 << 32), b))


def cols(m):
    """3x3 nested list (rows) of raw values -> column-major 9 raw values."""
    return [m[r][c] for c in range(3) for r in range(3)]


def rows_f(v9):
    """column-major raw 9 -> float rows."""
    return np.array(v9, np.float64).reshape(3, 3).T / 2 ** 32


def _stress(diag, shear):
    q = [[X.from_f32(diag[0]), X.from_f32(shear), 0], [X.from_f32(shear), X.from_f32(diag[1]), 0],
         [0, 0, X.from_f32(diag[2])]]
    return cols(q)


IDENT = cols([[ONE, 0, 0], [0, ONE, 0], [0, 0, ONE]])


def test_plasticity_return_mapping():
    """handle_plasticity (mpm.rs:474-502): above the limit yield + hardening
    * plastic the deviator is scaled by limit / vm with vm = sqrt(1.5 |dev|)
    (the reference's measure), the trace kept and plastic grows by
    (vm - limit) / E; below it nothing changes."""
    cfg = X.default_config(model=X.ELASTOPLASTIC, yield_=X.from_f32(1.0))
    S = _stress((3.0, -1.0, 0.5), 2.0)
    F, S2, _, pl, *_ = X.constitutive(cfg, IDENT, S, plastic=X.from_f32(2.0))
    s, s2 = rows_f(S), rows_f(S2)
    h = np.trace(s) / 3
    dev = s - h * np.eye(3)
    dn = math.sqrt((np.diag(dev) ** 2).sum() + 2 * (dev[0, 1] ** 2 + dev[0, 2] ** 2 + dev[1, 2] ** 2))
    vm = math.sqrt(1.5 * dn)
    limit = 1.0 + 0.1 * 2.0
    assert vm > limit
    assert abs(np.trace(s2) - np.trace(s)) < 1e-8
    assert np.abs((s2 - h * np.eye(3)) - dev * (limit / vm)).max() < 1e-7
    assert abs((pl - X.from_f32(2.0)) / 2 ** 32 - (vm - limit) / 1e6) < 1e-9
    assert list(F) == IDENT
    small = _stress((0.1, 0.0, 0.0), 0.0)
    _, S3, _, pl3, *_ = X.constitutive(cfg, IDENT, small)
    assert list(S3) == small and pl3 == 0


def test_granular_cap_and_tension():
    """handle_granular (mpm.rs:504-540): tension clears the stress; in
    compression the deviator norm is capped at tan(phi) p + yield."""
    cfg = X.default_config(model=X.GRANULAR, model_param=X.from_f32(30.0), yield_=X.from_f32(1.0))
    _, S, _, pl, *_ = X.constitutive(cfg, IDENT, _stress((1.0, 1.0, 1.0), 0.0))
    assert not any(S) and pl == 0
    Sin = _stress((-3.0, -1.0, -2.0), 4.0)
    _, S2, _, pl2, *_ = X.constitutive(cfg, IDENT, Sin)
    s, s2 = rows_f(Sin), rows_f(S2)
    h = np.trace(s) / 3
    dev = s - h * np.eye(3)
    dn = math.sqrt((np.diag(dev) ** 2).sum() + 2 * (dev[0, 1] ** 2 + dev[0, 2] ** 2 + dev[1, 2] ** 2))
    cap = math.tan(math.radians(30.0)) * (-h) + 1.0
    assert dn > cap
    assert np.abs((s2 - h * np.eye(3)) - dev * (cap / dn)).max() < 1e-6
    assert abs(pl2 / 2 ** 32 - (dn - cap) / 1e6) < 1e-9


def test_fluid_pressure_and_damping():
    """handle_fluid (mpm.rs:542-562): stress -p I with p = -tr / 3 (zero in
    tension); velocity times exp(-viscosity |v|), exp being the ten-term
    Taylor sum of FixedPoint::exp."""
    cfg = X.default_config(model=X.FLUID, model_param=X.from_f32(0.5))
    Sin = _stress((-3.0, -1.0, -2.0), 4.0)
    v = [X.from_f32(3.0), X.from_f32(4.0), 0]
    _, S, v2, *_ = X.constitutive(cfg, IDENT, Sin, vel3=v)
    tr = Sin[0] + Sin[4] + Sin[8]
    p = trunc_div(-tr, 3)
    assert list(S) == cols([[-p, 0, 0], [0, -p, 0], [0, 0, -p]])
    x = -0.5 * 5.0
    taylor = sum(x ** k / math.factorial(k) for k in range(12))
    assert np.allclose(np.array(v2[:2]) / 2 ** 32, np.array([3.0, 4.0]) * taylor, rtol=1e-6, atol=1e-8)
    _, S3, *_ = X.constitutive(cfg, IDENT, _stress((1.0, 2.0, 3.0), 0.0))
    assert not any(S3)


def test_svd_quirk_and_snow_clamp():
    """fixed_mat3_svd (mpm.rs:794-888) on A = R D (R a quarter turn about z):
    A^T A is already diagonal, so no Jacobi sweep runs and the closed form
    holds: S = sqrt(diag(A^T A)), V = A^T A itself (the reference hands back
    the rotated matrix, not the rotations), U = A V / S.  Snow
    (mpm.rs:565-639) clamps S to [1 - 2.5e-2, 1 + 7.5e-3] and rebuilds
    F = U S V^T; hardening takes the model's value."""
    d = [X.from_f32(1.1), X.from_f32(0.9), ONE]
    A = [[0, -d[1], 0], [d[0], 0, 0], [0, 0, d[2]]]
    u, sv, v = X.mat_op(2, np.array(cols(A), np.int64)).reshape(3, 9)
    g = [qmul(di, di) for di in d]
    s = [math.isqrt(gi << 32) for gi in g]
    assert list(sv[[0, 4, 8]]) == s
    assert list(v) == cols([[g[0], 0, 0], [0, g[1], 0], [0, 0, g[2]]])
    U = [[qmul(qmul(A[i][j], g[j]), qdiv(ONE, s[j])) for j in range(3)] for i in range(3)]
    assert list(u) == cols(U)
    cfg = X.default_config(model=X.SNOW, model_param=X.from_f32(0.05))
    F2, _, _, _, hard, *_ = X.constitutive(cfg, cols(A), [0] * 9)
    lo, hi = ONE - X.from_f32(2.5e-2), ONE + X.from_f32(7.5e-3)
    sc = [min(max(si, lo), hi) if not (lo <= si <= hi) else si for si in s]

    def mm(P, Q):
        return [[wrap64(sum(qmul(P[r][k], Q[k][c]) for k in range(3))) for c in range(3)] for r in range(3)]
    US = mm(U, [[sc[0], 0, 0], [0, sc[1], 0], [0, 0, sc[2]]])
    Vt = [[g[0], 0, 0], [0, g[1], 0], [0, 0, g[2]]]
    assert list(F2) == cols(mm(US, Vt))
    assert hard == X.from_f32(0.05)
    # brittle: plasticity, then the model's threshold switches fracture on
    cfg = X.default_config(model=X.BRITTLE, model_param=X.from_f32(0.02))
    *_, fs, on = X.constitutive(cfg, IDENT, [0] * 9)
    assert fs == X.from_f32(0.02) and on == 1


def test_fracture_closed_form():
    """handle_fracture (mpm.rs:641-716) on one particle with F = 1.2 I:
    strain = tr(F^T F) / 3 - 1 = 0.44 against 0.15 gives
    min(0.44 / 0.15, 8) -> 2 + 1 = 3 fragments on a ring of half a cell (y
    flattened to 0.2 cell, z from sin(1.7 i)), each with mass / 4, ids and
    connections in order, velocity + 2 offset, and one event."""
    cfg = X.small_config(particle_count=0)
    o = X.Sim(cfg)
    pos = [X.from_f32(1.0), X.from_f32(1.0), X.from_f32(1.0)]
    pid = o.add_particle(pos, (X.from_f32(0.5), 0, 0), X.from_f32(0.2))
    d = o.read()
    q = d["q"].copy()
    f = X.from_f32(1.2)
    q[0, 6:15] = cols([[f, 0, 0], [0, f, 0], [0, 0, f]])
    o.write(0, q, d["flags"], d["ids"], d["color"])
    thr = X.from_f32(0.15)
    o.set_fracture(True, thr)
    o.fracture_only()
    strain = trunc_div(3 * qmul(f, f), 3) - ONE
    nf = int(min(np.float32(qdiv(strain, thr)) / np.float32(2 ** 32), np.float32(8.0))) + 1
    assert nf == 3
    d = o.read()
    assert o.count() == 1 + nf and d["flags"][0].tolist() == [1, 1, 0]
    assert d["q"][0, 24] == X.from_f32(0.2) // (nf + 1)
    cell = 0.5
    for k in range(nf):
        r = 1 + k
        ang = np.float32(np.float32(k) * np.float32(2 * math.pi)) / np.float32(nf)
        off = np.array([math.cos(ang) * cell / 2, math.sin(ang) * cell * 0.2,
                        math.sin(float(np.float32(np.float32(k) * np.float32(1.7)))) * cell / 2])
        assert np.abs(d["q"][r, 0:3] / 2 ** 32 - (1.0 + off)).max() < 1e-6
        assert np.abs(d["q"][r, 3:6] / 2 ** 32 - (np.array([0.5, 0, 0]) + 2 * off)).max() < 1e-6
        assert d["q"][r, 24] == X.from_f32(0.2) // (nf + 1)
        assert d["ids"][r, 0] == pid + 1 + k
        assert d["ids"][r, 1:].tolist() == [pid + 1 + c for c in range(min(nf, 7))] + [0] * (8 - min(nf, 7))
        assert d["flags"][r].tolist() == [1, 0, min(nf, 8)]
        assert d["q"][r, 6:15].tolist() == IDENT
    ev = o.events()
    assert ev["frags"].tolist() == [nf] and ev["strain"][0] == np.float32(strain / 2 ** 32)
    assert ev["pos"][0].tolist() == [1.0, 1.0, 1.0]
    o.fracture_only()  # a fractured parent does not break again; fragments have F = I
    assert o.count() == 1 + nf


# ---------------------------------------------------------------------------
# Golden fixtures (tests/golden/fixed_*.npz, made by make_golden_fixed.py)
# ---------------------------------------------------------------------------
GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "fixed_*.npz")))


def _golden_config(z):
    return X.FxConfig.from_buffer_copy(z["config"].tobytes())


def _check_golden(sim, z, where):
    d = sim.read()
    for k in ("q", "flags", "ids", "color"):
        assert np.array_equal(d[k], z[k]), (where, k)
    st = sim.state()
    assert st["step_count"] == int(z["step_count"]) and st["next_id"] == int(z["next_id"]), where
    assert st["time"] == float(z["time"]), where
    ev = sim.events()
    for k, zk in (("time", "ev_time"), ("pos", "ev_pos"), ("strain", "ev_strain"), ("frags", "ev_frags")):
        assert np.array_equal(ev[k], z[zk]), (where, k)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_golden_fixture(path):
    z = np.load(path)
    s = X.Sim(_golden_config(z))
    s.step(int(z["steps"]))
    _check_golden(s, z, os.path.basename(path))


@pytest.mark.gpu
@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_gpu_fixed_matches_golden_fixture(path):
    import aemh
    z = np.load(path)
    g = aemh.FixedSim(_golden_config(z))
    g.step(int(z["steps"]))
    _check_golden(g, z, os.path.basename(path))
    g.close()


# ---------------------------------------------------------------------------
# GPU against the oracle
# ---------------------------------------------------------------------------
BOUND_LIMIT = 1 << 39  # sum of (|c| >> 24) + 1 below this: the atomic scatter is exact


def _gpu_pair(cfg):
    import aemh
    return aemh.FixedSim(cfg), X.Sim(cfg)


def _compare(g, o, where):
    dg, do = g.read(), o.read()
    assert dg["q"].shape == do["q"].shape, where
    bad = np.argwhere(dg["q"] != do["q"])
    assert bad.size == 0, f"{where}: {len(bad)} values differ, first {bad[:3].tolist()}"
    assert np.array_equal(dg["flags"], do["flags"]), where
    assert np.array_equal(dg["ids"], do["ids"]), where
    assert np.array_equal(dg["color"], do["color"]), where
    sg, so = g.state(), o.state()
    for k in ("step_count", "next_id", "n_events", "hardening", "fracture_strain", "fracture_on",
              "time"):
        assert sg[k] == so[k], (where, k, sg[k], so[k])
    assert g.stats() == o.stats(), where
    ev_g, ev_o = g.events(), o.events()
    for k in ev_o:
        assert np.array_equal(ev_g[k], ev_o[k]), (where, k)
    # the scatter bound is an exact integer sum below its limit
    gb, ob = sg["max_scatter_bound"], o.scatter_bound()
    assert (gb == ob) if ob < BOUND_LIMIT else (gb >= BOUND_LIMIT), (where, gb, ob)


def _compare_snapshots(g, o):
    snaps = g.snapshots()
    f, fr = o.snapshots()
    assert len(snaps) == len(f)
    got = np.array([list(s.position) + list(s.velocity) + [s.stress_magnitude, s.plastic_strain,
                    s.temperature] + list(s.color) for s in snaps], np.float32).reshape(-1, 13)
    assert np.array_equal(got, f)
    assert [s.fractured for s in snaps] == fr.tolist()


# model, parameter, steps: each run stops before the reference's own
# dynamics of that model reach the Q32.32 range (snow's SVD feeds A^T A back
# into F, brittle fragments fracture again)
MODELS = [(X.ELASTIC, 0, 8), (X.ELASTOPLASTIC, 0, 8), (X.GRANULAR, X.from_f32(30.0), 12),
          (X.FLUID, X.from_f32(0.01), 12), (X.SNOW, X.from_f32(0.05), 3),
          (X.BRITTLE, X.from_f32(0.02), 4)]


@pytest.mark.gpu
@pytest.mark.parametrize("model,param,steps", MODELS,
                         ids=["elastic", "elastoplastic", "granular", "fluid", "snow", "brittle"])
@pytest.mark.parametrize("count,substeps", [(8, 1), (125, 2)], ids=["8x1", "125x2"])
def test_gpu_fixed_models_bitexact(model, param, steps, count, substeps):
    cfg = X.small_config(model=model, model_param=param, particle_count=count, substeps=substeps)
    print(f"fixed model {model} particles {count} substeps {substeps}", flush=True)
    g, o = _gpu_pair(cfg)
    for k in range(steps):
        g.step(1)
        o.step(1)
        _compare(g, o, f"model {model} step {k + 1}")
    _compare_snapshots(g, o)
    g.close()


@pytest.mark.gpu
def test_gpu_fixed_fracture_add_remove_switch():
    cfg = X.small_config(enable_fracture=1, fracture_strain=X.from_f32(0.01))
    print("fixed fracture/add/remove", flush=True)
    g, o = _gpu_pair(cfg)
    g.step(2)
    o.step(2)
    _compare(g, o, "fracture 2")
    for s in (g, o):
        s.add_particle((X.from_f32(0.7), X.from_f32(1.1), X.from_f32(0.6)), (0, X.from_f32(-0.5), 0),
                       X.from_f32(0.05))
    g.step(3)
    o.step(3)
    _compare(g, o, "fracture 5")
    assert o.state()["n_events"] > 0 and o.count() > 9
    pid = int(o.read()["ids"][2, 0])
    assert g.remove_particle(pid) is True and o.remove_particle(pid) is True
    assert g.remove_particle(10 ** 9) is False
    g.set_model(X.SNOW, X.from_f32(0.05))
    o.set_model(X.SNOW, X.from_f32(0.05))
    g.set_fracture(False)
    o.set_fracture(False)
    g.step(1)
    o.step(1)
    _compare(g, o, "after remove + snow")
    _compare_snapshots(g, o)
    g.close()


@pytest.mark.gpu
def test_gpu_fixed_default_config():
    """MpmConfig::default: 22^3 = 10,648 particles in one cell's span of a
    64^3 grid, 4 substeps, all scattering into the same few nodes."""
    cfg = X.default_config()
    print("fixed default config", flush=True)
    g, o = _gpu_pair(cfg)
    g.step(2)
    o.step(2)
    _compare(g, o, "default config")
    g.close()


def _heavy(sims, stress, mass, shear=None):
    """Assigns every particle a mass and a diagonal stress (and an
    alternating xy shear), as a caller writing the public fields."""
    d = sims[-1].read()
    q = d["q"].copy()
    q[:, 24] = X.from_f32(mass)
    for k in (15, 19, 23):
        q[:, k] = X.from_f32(stress)
    if shear is not None:
        q[::2, 16] = X.from_f32(shear)
    for s_ in sims:
        s_.write(0, q, d["flags"], d["ids"], d["color"])


@pytest.mark.gpu
def test_gpu_fixed_scatter_check_paths():
    """Scatter exactness past the cheap bound.  Spread: 1,000 unit-mass
    particles at stress 4e6 over a 16^3 grid, whose substep total passes the
    limit while every cell stays far below it (the per-cell check alone).
    Heavy: 27 particles of mass 1000 at stress 1e6 in one cell span, whose
    cells can saturate (the serial rebuild, saturating adds in order)."""
    print("fixed scatter check: spread", flush=True)
    cfg = X.small_config(model=X.ELASTIC, particle_count=0, res=(16, 16, 16))
    g, o = _gpu_pair(cfg)
    rng = np.random.default_rng(11)
    pos = np.array([[X.from_f32(v) for v in p] for p in rng.uniform(0.75, 7.25, (1000, 3))], np.int64)
    zero = np.zeros((1000, 3), np.int64)
    g.add_particles(pos, zero, np.full(1000, X.ONE, np.int64))
    for i in range(1000):
        o.add_particle(pos[i], zero[i], X.ONE)
    _heavy((g, o), 4.0e6, 1.0)
    for k in range(2):
        g.step(1)
        o.step(1)
        _compare(g, o, f"spread step {k + 1}")
    st = g.state()
    assert o.scatter_bound() >= BOUND_LIMIT and st["checked_substeps"] > 0 and st["serial_substeps"] == 0
    g.close()
    print("fixed scatter check: heavy", flush=True)
    cfg = X.small_config(model=X.ELASTIC, particle_count=27)
    g, o = _gpu_pair(cfg)
    _heavy((g, o), 1.0e6, 1000.0, shear=-3.0e5)
    for k in range(3):
        g.step(1)
        o.step(1)
        _compare(g, o, f"heavy step {k + 1}")
    assert g.state()["serial_substeps"] > 0
    g.close()


@pytest.mark.gpu
def test_gpu_fixed_cloud_reused_order():
    """20,000 particles in random order (the benchmark cloud, scaled down)
    for 10 steps with a re-sort interval of 4: P2G sorts them by base cell
    and reuses that order for the next substeps while particles cross cells,
    so equal cells end up in several runs of a wave.  Every step bit-exact."""
    import aemh
    from aemh import workloads as W
    sc = W.fixed_cloud(20000, 32)
    print("fixed cloud 20k, reused P2G order", flush=True)
    g, o = _gpu_pair(sc.config)
    for bad in (0, 65537):
        with pytest.raises(aemh.AemhError):
            g.set_sort_interval(bad)
    g.set_sort_interval(4)
    zero = np.zeros_like(sc.pos)
    g.add_particles(sc.pos, zero, sc.mass)
    o.add_particles(sc.pos, zero, sc.mass)
    for k in range(10):
        g.step(1)
        o.step(1)
        _compare(g, o, f"cloud step {k + 1}")
    g.close()
    assert isinstance(sc.config, aemh.FixedConfig)


@pytest.mark.gpu
def test_gpu_fixed_brittle_cascade_growth():
    """Brittle fragments break again: 125 particles become 9,588 with 4,043
    events in 10 steps, growing the particle store, the event array and the
    P2G order buffers past their first sizes; every step bit-exact."""
    cfg = X.small_config(model=X.BRITTLE, model_param=X.from_f32(0.02), particle_count=125, substeps=2)
    print("fixed brittle cascade", flush=True)
    g, o = _gpu_pair(cfg)
    for k in range(10):
        g.step(1)
        o.step(1)
        _compare(g, o, f"brittle step {k + 1}")
    assert o.count() > 4096 and o.state()["n_events"] > 1024
    _compare_snapshots(g, o)
    g.close()


@pytest.mark.gpu
def test_gpu_fixed_random_api_sequence():
    """A seeded random sequence of the API between steps -- add one or many
    particles, remove (first, middle, last, missing), write records (move,
    push, deactivate, reactivate), switch models, toggle fracture, step one
    or several times -- on a 2,000-particle cloud.  The P2G order the device
    keeps across substeps must follow every change of the particle set;
    every field is compared after every operation."""
    from aemh import workloads as W
    rng = np.random.default_rng(2024)
    sc = W.fixed_cloud(2000, 32)
    g, o = _gpu_pair(sc.config)
    zero = np.zeros_like(sc.pos)
    g.add_particles(sc.pos, zero, sc.mass)
    o.add_particles(sc.pos, zero, sc.mass)
    models = [(X.ELASTOPLASTIC, 0), (X.ELASTIC, 0), (X.FLUID, X.from_f32(0.01)),
              (X.GRANULAR, X.from_f32(30.0)), (X.SNOW, X.from_f32(0.05))]
    for it in range(40):
        op = rng.integers(0, 7)
        n = o.count()
        if op == 0:  # add one
            p = [X.from_f32(v) for v in rng.uniform(4.0, 12.0, 3)]
            v = [X.from_f32(v) for v in rng.uniform(-1.0, 1.0, 3)]
            m = X.from_f32(float(rng.uniform(0.5, 2.0)))
            assert g.add_particle(p, v, m) == o.add_particle(p, v, m)
        elif op == 1:  # add many
            k = int(rng.integers(1, 300))
            p = np.array([[X.from_f32(v) for v in q] for q in rng.uniform(4.0, 12.0, (k, 3))], np.int64)
            m = np.full(k, X.from_f32(1.5), np.int64)
            col = rng.uniform(0, 1, (k, 4)).astype(np.float32)
            g.add_particles(p, np.zeros_like(p), m, col)
            for i in range(k):
                o.add_particle(p[i], (0, 0, 0), m[i], tuple(col[i]))
        elif op == 2 and n:  # remove
            ids = o.read()["ids"][:, 0]
            pick = [ids[0], ids[n // 2], ids[-1], 10 ** 12][int(rng.integers(0, 4))]
            assert g.remove_particle(int(pick)) == o.remove_particle(int(pick))
        elif op == 3 and n:  # write a slice of records
            first = int(rng.integers(0, n))
            cnt = int(rng.integers(1, min(200, n - first) + 1))
            d = o.read()
            q, fl = d["q"][first:first + cnt].copy(), d["flags"][first:first + cnt].copy()
            q[:, 0:3] += rng.integers(-(1 << 31), 1 << 31, (cnt, 3))
            q[:, 3:6] = rng.integers(-(1 << 32), 1 << 32, (cnt, 3))
            fl[:, 0] = rng.integers(0, 2, cnt)
            for s_ in (g, o):
                s_.write(first, q, fl, d["ids"][first:first + cnt], d["color"][first:first + cnt])
        elif op == 4:  # switch model
            m, prm = models[int(rng.integers(0, len(models)))]
            g.set_model(m, prm)
            o.set_model(m, prm)
        elif op == 5:  # one step with fracture on (fragments of this exploding
            # scene would break again and again if it stayed on)
            thr = X.from_f32(float(rng.uniform(0.05, 0.5)))
            for s_ in (g, o):
                s_.set_fracture(True, thr)
                s_.step(1)
                s_.set_fracture(False)
            _compare(g, o, f"op {it} (fracture step)")
        k = int(rng.integers(1, 4))
        g.step(k)
        o.step(k)
        print(f"random api op {it}: kind {op}, {o.count()} particles", flush=True)
        _compare(g, o, f"op {it} ({op})")
    _compare_snapshots(g, o)
    g.close()


@pytest.mark.gpu
def test_gpu_fixed_snow_long():
    """Snow for 12 steps: the reference's SVD returns the rotated A^T A as V,
    so F = U S (A^T A)^T grows each substep and the stresses reach ~2e9."""
    cfg = X.small_config(model=X.SNOW, model_param=X.from_f32(0.05), particle_count=125)
    g, o = _gpu_pair(cfg)
    for k in range(12):
        g.step(1)
        o.step(1)
        _compare(g, o, f"snow step {k + 1}")
    g.close()


@pytest.mark.gpu
def test_gpu_fixed_batch_add_write_and_empty():
    """add_particles equals add_particle in a loop; writing records (a
    caller assigning the public fields, e.g. deactivating) is honoured; an
    empty simulation steps."""
    import aemh
    cfg = X.small_config(particle_count=0, res=(8, 8, 8), cell=X.from_f32(0.25))
    g, o = _gpu_pair(cfg)
    g.step(1)
    o.step(1)
    _compare(g, o, "empty")
    rng = np.random.default_rng(5)
    n = 300
    pos = np.array([[X.from_f32(v) for v in p] for p in rng.uniform(0.3, 1.7, (n, 3))], np.int64)
    vel = np.array([[X.from_f32(v) for v in p] for p in rng.uniform(-0.5, 0.5, (n, 3))], np.int64)
    mass = np.array([X.from_f32(v) for v in rng.uniform(0.001, 0.01, n)], np.int64)
    col = rng.uniform(0, 1, (n, 4)).astype(np.float32)
    first = g.add_particles(pos, vel, mass, col)
    ids = [o.add_particle(pos[i], vel[i], mass[i], tuple(col[i])) for i in range(n)]
    assert first == ids[0] and ids == list(range(ids[0], ids[0] + n))
    _compare(g, o, "batch add")
    g.set_model(X.ELASTOPLASTIC, 0)
    o.set_model(X.ELASTOPLASTIC, 0)
    g.step(2)
    o.step(2)
    _compare(g, o, "batch add stepped")
    d = o.read()
    fl = d["flags"].copy()
    fl[::7, 0] = 0
    q = d["q"].copy()
    q[::5, 3:6] = 0
    for s in (g, o):
        s.write(0, q, fl, d["ids"], d["color"])
    g.step(2)
    o.step(2)
    _compare(g, o, "after write")
    _compare_snapshots(g, o)
    g.close()
    assert isinstance(aemh.fixed_config_default(), aemh.FixedConfig)
