"""Pinning the oracle (oracle/refcpu.cpp) before trusting it.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign or index, a transposed operand) fails: the reference's own
known-answer tests, closed forms, exact integer arithmetic, and invariants.
CPU only.
"""
import math

import numpy as np
import pytest

from oracle import refcpu as R

f32 = np.float32


# ---------------------------------------------------------------------------
# B-spline: ae_particle/src/mpm_solver.rs:1596-1609 (reference unit tests)
# ---------------------------------------------------------------------------
def test_bspline_reference_known_answers():
    L = R.lib()
    assert abs(L.ref_b_spline(0.0) - 0.75) < 1e-6
    assert abs(L.ref_b_spline(0.5) - 0.5) < 1e-6
    assert abs(L.ref_b_spline(1.0) - 0.125) < 1e-6
    assert abs(L.ref_b_spline(1.5)) < 1e-6
    assert abs(L.ref_b_spline(2.0)) < 1e-6
    assert abs(L.ref_b_spline_deriv(0.0)) < 1e-6
    assert abs(L.ref_b_spline_deriv(0.5) + 1.0) < 1e-6
    assert abs(L.ref_b_spline_deriv(1.0) + 0.5) < 1e-6


def test_bspline_closed_form_and_derivative():
    L = R.lib()
    xs = np.linspace(-2.2, 2.2, 881, dtype=np.float64)
    for x in xs:
        ax = abs(x)
        n = 0.75 - ax * ax if ax < 0.5 else (0.5 * (1.5 - ax) ** 2 if ax < 1.5 else 0.0)
        assert abs(L.ref_b_spline(float(x)) - n) < 2e-7
        if abs(ax - 0.5) > 1e-3 and abs(ax - 1.5) > 1e-3:
            h = 1e-4
            fd = (0.75 - (x + h) ** 2 if abs(x + h) < .5 else (.5 * (1.5 - abs(x + h)) ** 2 if abs(x + h) < 1.5 else 0)) - \
                 (0.75 - (x - h) ** 2 if abs(x - h) < .5 else (.5 * (1.5 - abs(x - h)) ** 2 if abs(x - h) < 1.5 else 0))
            assert abs(L.ref_b_spline_deriv(float(x)) - fd / (2 * h)) < 1e-3


def test_stencil_partition_deficit():
    """Q1: the reference stencil {0,1,2} from floor(X-0.5) sums to
    1 - (0.5-fx)^2/2 for fx < 0.5 and to 1 otherwise."""
    L = R.lib()
    for fx in np.linspace(0.0, 0.999, 200):
        s = sum(L.ref_b_spline(float(f32(d) - f32(fx))) for d in range(3))
        expect = 1.0 - 0.5 * (0.5 - fx) ** 2 if fx < 0.5 else 1.0
        assert abs(s - expect) < 2e-6


# ---------------------------------------------------------------------------
# FixedPoint Q32.32: ae_physics/src/fixed_point.rs
# ---------------------------------------------------------------------------
def _trunc_div(a, b):
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def _wrap64(v):
    v &= (1 << 64) - 1
    return v - (1 << 64) if v >= (1 << 63) else v


def test_fixed_from_f32_exact_truncation():
    L = R.lib()
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.normal(0, 1, 2000), rng.normal(0, 1e-6, 500), rng.normal(0, 1e5, 500),
                           [0.0, -0.0, 1.0, -1.0, 0.5, 2.5e-10, -2.5e-10]]).astype(f32)
    from fractions import Fraction
    for v in vals:
        exact = Fraction(float(v)) * (1 << 32)
        t = int(exact)  # truncation toward zero
        assert L.ref_fp_from_f32(float(v)) == t
    # saturation and NaN (Rust `as i64` semantics)
    assert L.ref_fp_from_f32(3e9) == (1 << 63) - 1
    assert L.ref_fp_from_f32(-3e9) == -(1 << 63)
    assert L.ref_fp_from_f32(float("nan")) == 0
    assert L.ref_fp_from_f32(float("inf")) == (1 << 63) - 1


def test_fixed_to_f32_correctly_rounded():
    L = R.lib()
    rng = np.random.default_rng(2)
    raws = rng.integers(-(1 << 62), 1 << 62, 3000, dtype=np.int64)
    for r in raws:
        expect = f32(np.int64(r).astype(f32) / f32(4294967296.0))
        assert L.ref_fp_to_f32(int(r)) == expect


def test_fixed_reference_unit_tests():
    """fixed_point.rs:889-962 (sqrt, arithmetic, determinism, overflow)."""
    L = R.lib()
    one = 1 << 32
    from_i = lambda i: i << 32
    to_f = lambda r: r / 2 ** 32
    assert abs(to_f(L.ref_fp_sqrt(from_i(4))) - 2.0) < 1e-3
    assert abs(to_f(L.ref_fp_sqrt(from_i(2))) - 1.414) < 1e-2
    a, b = L.ref_fp_from_f32(3.5), L.ref_fp_from_f32(2.0)
    assert abs(to_f(L.ref_fp_add(a, b)) - 5.5) < 1e-3
    assert abs(to_f(L.ref_fp_mul(a, b)) - 7.0) < 1e-3
    assert abs(to_f(L.ref_fp_div(a, b)) - 1.75) < 1e-3
    third = L.ref_fp_from_f32(1.0 / 3.0)
    assert abs(to_f(L.ref_fp_mul(third, from_i(3))) - 1.0) < 1e-3
    # MAX + ONE saturates (checked_add is None -> `+` saturates)
    assert L.ref_fp_add((1 << 63) - 1, one) == (1 << 63) - 1
    assert L.ref_fp_add(-(1 << 63), -one) == -(1 << 63)


def test_fixed_mul_div_sqrt_exact_i128():
    L = R.lib()
    rng = np.random.default_rng(3)
    for _ in range(2000):
        a = int(rng.integers(-(1 << 40), 1 << 40))
        b = int(rng.integers(-(1 << 40), 1 << 40)) or 1
        assert L.ref_fp_mul(a, b) == _wrap64(_trunc_div(a * b, 1 << 32))
        assert L.ref_fp_div(a, b) == _wrap64(_trunc_div(a << 32, b))
        if a > 0:
            assert L.ref_fp_sqrt(a) == math.isqrt(a << 32)
    assert L.ref_fp_div(5, 0) == (1 << 63) - 1 and L.ref_fp_div(-5, 0) == -(1 << 63)
    assert L.ref_fp_div(0, 0) == 0


# ---------------------------------------------------------------------------
# Portable ln / exp of the deterministic mode (no reference counterpart)
# ---------------------------------------------------------------------------
def _ulp_diff(a, b):
    ia = np.array([a], f32).view(np.int32)[0]
    ib = np.array([b], f32).view(np.int32)[0]
    return abs(int(ia) - int(ib))


def test_det_logf_matches_libm_double():
    L = R.lib()
    rng = np.random.default_rng(4)
    xs = np.concatenate([np.exp(rng.uniform(-80, 80, 20000)), rng.uniform(0.9, 1.1, 20000),
                         [1.0, 2.0, 0.5, 1e-38, 3e38, 1.0000001]]).astype(f32)
    worst, exact = 0, 0
    for x in xs:
        ref = f32(math.log(float(x)))
        got = L.ref_det_logf(float(x))
        d = _ulp_diff(got, ref)
        worst = max(worst, d)
        exact += d == 0
    assert worst <= 1
    assert exact / len(xs) > 0.999
    assert L.ref_det_logf(1.0) == 0.0


def test_det_expf_matches_libm_double():
    L = R.lib()
    rng = np.random.default_rng(5)
    xs = np.concatenate([rng.uniform(-87, 88, 20000), rng.uniform(-1, 1, 20000), [0.0, 1.0, -1.0]]).astype(f32)
    worst, exact = 0, 0
    for x in xs:
        ref = f32(math.exp(float(x)))
        got = L.ref_det_expf(float(x))
        d = _ulp_diff(got, ref)
        worst = max(worst, d)
        exact += d == 0
    assert worst <= 1
    assert exact / len(xs) > 0.999


# ---------------------------------------------------------------------------
# Constitutive models: ae_particle/src/constitutive.rs
# ---------------------------------------------------------------------------
DEFAULT = [1e6, 0.3, 1e4, 0.01, 1000.0, 35.0, 0.0]       # MaterialParams::default
SAND = [5.0e7, 0.2, 1.0e4, 0.005, 1600.0, 35.0, 0.0]       # presets::SAND
STEEL = [2.0e11, 0.3, 2.5e8, 0.01, 7850.0, 0.0, 0.0]       # presets::STEEL
I3 = np.eye(3, dtype=f32)


def lame(E, nu):
    nu = min(max(nu, -0.999), 0.4999)
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu)), E / (3 * (1 - 2 * nu))


def test_reference_constitutive_unit_tests():
    """constitutive.rs:623-665."""
    s, _, _ = R.compute_stress(R.NEO_HOOKEAN, DEFAULT, I3, 1.0)
    assert np.all(np.abs(s) < 1e-4)
    s, _, _ = R.compute_stress(R.NEWTONIAN_FLUID, DEFAULT, I3, 1.0)
    assert abs(s[0, 0]) < 1e-4 and abs(s[1, 1]) < 1e-4 and abs(s[2, 2]) < 1e-4
    s, _, _ = R.compute_stress(R.NEWTONIAN_FLUID, DEFAULT, I3, 0.9)
    assert s[0, 0] < 0
    shear = I3.copy()
    shear[0, 1] = 0.5
    _, _, plastic = R.compute_stress(R.DRUCKER_PRAGER, SAND, shear, 1.0)
    assert plastic
    _, _, plastic = R.compute_stress(R.VON_MISES, STEEL, I3, 1.0)
    assert not plastic


@pytest.mark.parametrize("diag,J", [((1.1, 0.95, 1.02), None), ((0.9, 0.9, 0.9), 0.8), ((1.3, 1.0, 0.7), None)])
def test_neo_hookean_closed_form(diag, J):
    mu, lam, _ = lame(1e4, 0.2)
    F = np.diag(diag).astype(f32)
    Jv = float(np.prod(diag)) if J is None else J
    s, nf, pl = R.compute_stress(R.NEO_HOOKEAN, [1e4, 0.2, 0, 0, 1, 0, 0], F, Jv)
    expect = np.diag([mu / Jv * (d * d - 1) + lam * math.log(Jv) / Jv for d in diag])
    assert np.allclose(s, expect, rtol=2e-5, atol=1e-3)
    assert np.array_equal(nf, F) and not pl
    # off-diagonal F: sigma = mu/J (F F^T - I) + lambda ln J / J I with stored J
    F2 = np.array([[1.0, 0.1, 0], [0, 1.0, 0.05], [0.02, 0, 1.0]], f32)
    s2, _, _ = R.compute_stress(R.NEO_HOOKEAN, [1e4, 0.2, 0, 0, 1, 0, 0], F2, 1.01)
    b = F2.astype(np.float64) @ F2.T.astype(np.float64)
    e2 = mu / 1.01 * (b - np.eye(3)) + lam * math.log(1.01) / 1.01 * np.eye(3)
    assert np.allclose(s2, e2, rtol=1e-4, atol=1e-2)
    # J <= 0: pressure term dropped, scale uses max(J, 1e-10)
    s3, _, _ = R.compute_stress(R.NEO_HOOKEAN, [1e4, 0.2, 0, 0, 1, 0, 0], F, -0.5)
    assert np.allclose(s3, mu / 1e-10 * (F.astype(np.float64) @ F.T - np.eye(3)), rtol=1e-5)


def test_fluid_closed_form_and_nu_clamp():
    for nu in (0.2, 0.45, 0.5, 0.7):
        # K = E / (3 (1 - 2 nu)) in f32 (at the 0.4999 clamp 1 - 2 nu cancels)
        nuc = min(max(f32(nu), f32(-0.999)), f32(0.4999))
        K = f32(2e6) / (f32(3.0) * (f32(1.0) - f32(2.0) * nuc))
        s, _, _ = R.compute_stress(R.NEWTONIAN_FLUID, [2e6, nu, 0, 0, 1000, 0, 0], I3, 0.97)
        assert np.allclose(np.diag(s), K * (f32(0.97) - f32(1)), rtol=1e-6)
        assert np.all(s[~np.eye(3, dtype=bool)] == 0)


def test_rigid_identity_stress():
    s, _, _ = R.compute_stress(R.RIGID, DEFAULT, np.diag([1.2, 0.8, 1.0]).astype(f32), 0.96)
    assert np.array_equal(s, np.eye(3, dtype=f32))


def _rot(axis, deg):
    t = math.radians(deg)
    c, s = math.cos(t), math.sin(t)
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + s * K + (1 - c) * K @ K


def test_corotated_polar_iteration_as_written():
    """polar_decomp averages r with the INVERSE (constitutive.rs:306-324 returns
    adjugate/det), i.e. the matrix-sign iteration: r -> I for an SPD F and for a
    rotation under 90 degrees.  A transpose-inverse polar would give r = R and
    zero stress for a pure rotation; the reference gives 2 mu (R - I)."""
    mu, lam, _ = lame(1e6, 0.3)
    Rm = _rot((0, 0, 1), 30).astype(f32)
    s, _, _ = R.compute_stress(R.COROTATED, DEFAULT, Rm, 1.0)
    assert np.allclose(s, 2 * mu * (Rm.astype(np.float64) - np.eye(3)), rtol=1e-4, atol=1.0)
    Fs = np.array([[1.05, 0.02, 0], [0.02, 0.97, 0.01], [0, 0.01, 1.01]], f32)   # SPD
    J = float(np.linalg.det(Fs))
    s, _, _ = R.compute_stress(R.COROTATED, DEFAULT, Fs, J)
    assert np.allclose(s, 2 * mu * (Fs - np.eye(3)) + lam * (J - 1) * np.eye(3), rtol=1e-4, atol=2.0)
    s, _, _ = R.compute_stress(R.FIXED_COROTATED, DEFAULT, Fs, J)
    assert np.allclose(s, 2 * mu * (Fs - np.eye(3)) + lam * (J - 1) * J * np.eye(3), rtol=1e-4, atol=2.0)


def test_drucker_prager_elastic_closed_form():
    """Small symmetric strain, cohesive material: r = I, s = F, elastic branch
    sigma = 2 mu dev(eps) + lambda tr(eps) I (constitutive.rs:443-530)."""
    p = [1e6, 0.25, 1e4, 0.0, 1800.0, 30.0, 5e3]
    mu, lam, _ = lame(1e6, 0.25)
    eps = np.diag([1e-4, -5e-5, 2e-5])
    F = (np.eye(3) + eps).astype(f32)
    s, nf, pl = R.compute_stress(R.DRUCKER_PRAGER, p, F, float(np.linalg.det(F)))
    e = F.astype(np.float64) - np.eye(3)
    tr = np.trace(e)
    expect = 2 * mu * (e - tr / 3 * np.eye(3)) + lam * tr * np.eye(3)
    assert not pl
    assert np.allclose(s, expect, rtol=1e-3, atol=1e-2)


def test_von_mises_return_mapping_closed_form():
    """Deviatoric stress scaled onto sigma_y/sqrt(3) (constitutive.rs:566-591)."""
    p = [1e6, 0.3, 1e3, 0.0, 7850.0, 0.0, 0.0]
    mu, lam, _ = lame(1e6, 0.3)
    F = np.diag([1.01, 0.995, 1.0]).astype(f32)
    s, nf, pl = R.compute_stress(R.VON_MISES, p, F, float(np.prod(np.diag(F))))
    assert pl
    e = F.astype(np.float64) - np.eye(3)
    tr = np.trace(e)
    sd = 2 * mu * (e - tr / 3 * np.eye(3))
    norm = math.sqrt(0.5 * np.sum(sd * sd))
    expect = sd * (1e3 / math.sqrt(3) / norm) + lam * tr * np.eye(3)
    assert np.allclose(s, expect, rtol=1e-3, atol=1e-2)


def test_snow_stress_hardening():
    mu, lam, _ = lame(1.4e5, 0.2)
    F = np.diag([1.002, 0.998, 1.001]).astype(f32)
    J = float(np.prod(np.diag(F)))
    for Jp in (1.0, 0.98, 1.05, 0.5, 3.0):
        h = min(max(math.exp(10 * (1 - Jp)), 0.1), 5.0)
        s, _, _ = R.compute_stress(R.SNOW, [1.4e5, 0.2, 0, 0, 400, 0, 0], F, J, Jp=Jp)
        expect = np.diag([mu * h / J * (d * d - 1) + lam * h * math.log(J) / J for d in np.diag(F)])
        assert np.allclose(s, expect, rtol=1e-4, atol=1e-2)


# ---------------------------------------------------------------------------
# SVD and snow clamp (no reference counterpart)
# ---------------------------------------------------------------------------
def test_svd_reconstruction_and_rotations():
    rng = np.random.default_rng(6)
    for _ in range(300):
        F = (np.eye(3) + rng.normal(0, 0.3, (3, 3))).astype(f32)
        U, s, V, used64 = R.svd3(F)
        rec = U.astype(np.float64) @ np.diag(s) @ V.T
        assert np.allclose(rec, F, atol=3e-5 * max(1, np.abs(F).max()))
        assert np.allclose(U.T @ U, np.eye(3), atol=2e-5)
        assert np.allclose(V.T @ V, np.eye(3), atol=2e-5)
        assert np.linalg.det(U) > 0 and np.linalg.det(V) > 0
        assert np.allclose(np.sort(np.abs(s)), np.sort(np.linalg.svd(F.astype(np.float64))[1]), rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("cond", [1e2, 1e4 * 3, 1e6, 1e8])
def test_svd_known_singular_values_and_fp64_fallback(cond):
    R1 = _rot((1, 2, 3), 40)
    R2 = _rot((-2, 1, 0.5), 75)
    sv = np.array([1.0, 0.5, 1.0 / cond])
    F = (R1 @ np.diag(sv) @ R2.T).astype(f32)
    U, s, V, used64 = R.svd3(F)
    assert used64 == (cond > 1e4)
    ref = np.linalg.svd(F.astype(np.float64))[1]
    assert np.allclose(np.sort(np.abs(s)), np.sort(ref), rtol=1e-5, atol=1e-7 if cond > 1e4 else 1e-5)
    Ud, sd, Vd = R.svd3_d(F.astype(np.float64))
    assert np.allclose(Ud @ np.diag(sd) @ Vd.T, F.astype(np.float64), atol=1e-12)


def test_snow_clamp_closed_form():
    R1 = _rot((1, 0, 1), 20)
    R2 = _rot((0, 1, 1), -35)
    for sv, jp0 in (((1.2, 0.9, 1.0), 1.0), ((1.004, 0.99, 0.97), 0.8), ((1.0, 1.0, 1.0), 1.3)):
        F = (R1 @ np.diag(sv) @ R2.T).astype(f32)
        Fc, jp = R.snow_clamp(F, jp0)
        sh = np.clip(sv, 1 - 2.5e-2, 1 + 7.5e-3)
        expect = R1 @ np.diag(sh) @ R2.T
        assert np.allclose(Fc, expect, atol=2e-5)
        assert abs(jp - jp0 * np.prod(sv) / np.prod(sh)) < 2e-5 * max(1, jp)


def test_snow_clamp_skip_rule_against_numpy_svd():
    """F and J_p stay bit-identical exactly when no singular value leaves
    [1 - theta_c, 1 + theta_s] (the Gershgorin shortcut must never skip a
    clamp); otherwise F = U clamp(S) V^T and J_p scales by prod(s / s_hat).
    Random F around the yield surface, checked against numpy's FP64 SVD;
    cases within 1e-5 of a bound are left out (float rounding decides)."""
    rng = np.random.default_rng(5)
    lo, hi = 1 - 2.5e-2, 1 + 7.5e-3
    n_skip = n_clamp = 0
    for _ in range(400):
        sv = rng.uniform(0.96, 1.02, 3)
        F = (_rot(rng.normal(size=3), rng.uniform(0, 180)) @ np.diag(sv)
             @ _rot(rng.normal(size=3), rng.uniform(0, 180)).T).astype(f32)
        s = np.linalg.svd(F.astype(np.float64))[1]
        if np.min(np.abs(np.concatenate([s - lo, s - hi]))) < 1e-5:
            continue
        Fc, jp = R.snow_clamp(F, 1.0)
        inside = bool(np.all((s >= lo) & (s <= hi)))
        if inside:
            n_skip += 1
            assert np.array_equal(Fc, F) and jp == 1.0, sv
        else:
            n_clamp += 1
            U, s2, Vt = np.linalg.svd(F.astype(np.float64))
            sh = np.clip(s2, lo, hi)
            assert np.allclose(Fc, U @ np.diag(sh) @ Vt, atol=3e-6), sv
            assert abs(jp - np.prod(s2) / np.prod(sh)) < 1e-5
    assert n_skip > 40 and n_clamp > 40, (n_skip, n_clamp)


# ---------------------------------------------------------------------------
# Step-level closed forms and invariants
# ---------------------------------------------------------------------------
def _fluid(E=1e3):
    return R.Materials([[E, 0.2, 0, 0, 1000.0, 0, 0]], [R.NEWTONIAN_FLUID])


def _S(fx):
    return 1.0 - 0.5 * (0.5 - fx) ** 2 if fx < 0.5 else 1.0


def test_free_fall_single_particle_closed_form():
    """sigma = 0 (fluid at J = 1): grid v = g dt on every node with mass, so
    G2P gives v = g dt * S(fx) S(fy) S(fz) (Q1 deficit), x += v dt, stored
    v *= damping (mpm_solver.rs:112-139, 914-1053)."""
    dx = 0.125  # dyadic grid and positions: node - particle distances are exact
    cfg = R.make_config(16, 16, 16, dx, substeps=1, gravity=(0, -9.81, 0))
    for pos in ([0.828125, 0.703125, 0.625], [0.5625, 0.90625, 0.765625], [0.796875, 0.546875, 0.8125]):
        st = R.State([pos], mass=[0.01])
        dt = 1e-3
        R.step(cfg, _fluid(), st, dt)
        X = np.array(pos, np.float64) / dx - 0.5
        fx = X - np.floor(X)
        vy = -9.81 * dt * np.prod([_S(f) for f in fx])
        assert np.allclose(st.vel[0], [0, vy * 0.995, 0], rtol=1e-5, atol=1e-9)
        assert abs(st.pos[0][1] - (pos[1] + vy * dt)) < 1e-6
        # C = 4/dx^2 sum w (g dt) (x) d with d = (k - fx) dx: row y only,
        # column c carries the first moment of the truncated stencil on axis c
        S = [_S(f) for f in fx]
        M1 = [sum(float(R.lib().ref_b_spline(k - f)) * (k - f) for k in range(3)) for f in fx]
        Cy = [4 / dx * (-9.81 * dt) * M1[c] * np.prod([S[a] for a in range(3) if a != c]) for c in range(3)]
        expect = np.zeros(9)
        expect[3:6] = Cy
        assert np.allclose(st.C[0], expect, rtol=1e-4, atol=1e-8)


def test_grid_mass_is_deficit_weighted_particle_mass():
    rng = np.random.default_rng(7)
    n = 500
    pos = rng.uniform(0.35, 0.65, (n, 3)).astype(f32)
    mass = rng.uniform(0.5, 1.5, n).astype(f32)
    cfg = R.make_config(32, 32, 32, 1 / 32)
    st = R.State(pos, mass=mass)
    g = R.step(cfg, _fluid(), st, 1e-4, want_grid=True)
    X = pos.astype(np.float64) * 32 - 0.5
    fx = X - np.floor(X)
    S = np.where(fx < 0.5, 1 - 0.5 * (0.5 - fx) ** 2, 1.0).prod(1)
    assert abs(g["m"].sum() - (mass * S).sum()) < 1e-4 * mass.sum()


def test_apic_linear_field_exact_where_stencil_complete():
    """v = A x and C = A on every particle: P2G makes node velocities A x_i
    (+ g dt); G2P with a complete stencil (fx >= 0.5 on all axes) returns
    C = A and v = A x_p + g dt exactly (Jiang et al. 2015)."""
    dx = 1 / 32
    rng = np.random.default_rng(8)
    A = rng.normal(0, 0.5, (3, 3))
    # dense cloud so every node near the probes is fully supported
    cloud = rng.uniform(0.3, 0.7, (4000, 3))
    probes = []
    for _ in range(20):
        c = rng.integers(12, 20, 3)
        f = rng.uniform(0.55, 0.95, 3)
        probes.append((c + 0.5 + f) * dx)
    pos = np.vstack([cloud, np.array(probes)]).astype(f32)
    vel = (pos.astype(np.float64) @ A.T).astype(f32)
    Cm = np.tile(A.reshape(9), (pos.shape[0], 1)).astype(f32)
    cfg = R.make_config(32, 32, 32, dx, gravity=(0, 0, 0), damping=1.0)
    st = R.State(pos, vel=vel, C9=Cm, mass=np.ones(pos.shape[0]))
    dt = 1e-6
    R.step(cfg, _fluid(), st, dt)
    k = slice(4000, None)
    Xp = pos[k].astype(np.float64)
    assert np.allclose(st.C[k], np.tile(A.reshape(9), (20, 1)), atol=2e-3 * np.abs(A).max())
    assert np.allclose(st.vel[k], Xp @ A.T, atol=2e-5)


def test_guards_nan_velocity_position_and_f_validity():
    cfg = R.make_config(16, 16, 16, 0.1)
    # NaN velocity on a lone particle: its node momenta are NaN -> G2P zeroes v, C
    st = R.State([[0.8, 0.8, 0.8]], vel=[[np.nan, 0, 0]], mass=[0.01])
    R.step(cfg, _fluid(), st, 1e-3)
    assert np.all(st.vel == 0) and np.all(st.C == 0)
    # NaN position -> (6, 5, 6), v = 0 (Q8), then the serial x/z clamp to
    # [0, (n-1) dx] = [0, 1.5] (mpm_solver.rs:1002-1015)
    st = R.State([[np.nan, 0.5, 0.5]], mass=[0.01])
    R.step(cfg, _fluid(), st, 1e-3)
    assert np.array_equal(st.pos[0], np.array([1.5, 5, 1.5], f32)) and np.all(st.vel == 0)
    # |F_ij| > 1e3 -> F = I, J = 1 (Q9)
    F = np.eye(3).reshape(1, 9).astype(f32)
    F[0, 1] = 2e3
    st = R.State([[0.8, 0.8, 0.8]], F9=F, J=[1.0], mass=[0.01])
    R.step(cfg, _fluid(), st, 1e-3)
    assert np.array_equal(st.F[0], np.eye(3, dtype=f32).reshape(9)) and st.J[0] == 1.0
    # J clamp [0.1, 10] (det of a valid F outside the range)
    F = (np.eye(3) * 3.0).reshape(1, 9).astype(f32)
    st = R.State([[0.8, 0.8, 0.8]], F9=F, J=[1.0], mass=[0.01])
    R.step(cfg, _fluid(), st, 1e-3)
    assert st.J[0] == f32(10.0)


def test_floor_and_side_clamps():
    cfg = R.make_config(16, 16, 16, 0.1, clamp_mode=R.CLAMP_SERIAL, gravity=(0, 0, 0))
    # below the floor: y -> 0, v_y = max(v_y, 0) * 0.5 then damping
    st = R.State([[0.8, 0.0001, 0.8]], mass=[0.01])
    st.pos[0, 1] = -0.05
    R.step(cfg, _fluid(), st, 1e-3)
    assert st.pos[0, 1] == 0.0 and st.vel[0, 1] == 0.0
    # serial clamp uses [0, (n-1) dx] regardless of origin; window uses origin
    o = (1.0, 0.0, 0.0)
    for mode, lo in ((R.CLAMP_SERIAL, 0.0), (R.CLAMP_WINDOW, 1.0)):
        c = R.make_config(16, 16, 16, 0.1, origin=o, clamp_mode=mode, gravity=(0, 0, 0))
        st = R.State([[0.5, 0.5, 0.5]], mass=[0.01])
        R.step(c, _fluid(), st, 1e-3)
        assert st.pos[0, 0] == f32(lo) if mode == R.CLAMP_WINDOW else st.pos[0, 0] == f32(0.5)


def test_q32_accumulation_is_order_independent_and_close_to_f32():
    rng = np.random.default_rng(9)
    n = 3000
    pos = rng.uniform(0.3, 0.6, (n, 3)).astype(f32)
    vel = rng.normal(0, 0.1, (n, 3)).astype(f32)
    mats = R.Materials([[1e4, 0.2, 0, 0, 1.0, 0, 0]], [R.NEO_HOOKEAN])
    mass = np.full(n, 0.027 / n, f32)
    cq = R.make_config(32, 32, 32, 1 / 32, accum=R.ACCUM_Q32)
    a = R.State(pos, vel=vel, mass=mass)
    b = a.copy()
    ga = R.step(cq, mats, a, 1e-4, want_grid=True)
    gb = R.step(cq, mats, b, 1e-4, idx=rng.permutation(n).astype(np.uint32), want_grid=True)
    assert np.array_equal(ga["q"], gb["q"])
    for f in ("pos", "vel", "C", "F", "J"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    cf = R.make_config(32, 32, 32, 1 / 32)
    c = R.State(pos, vel=vel, mass=mass)
    gc = R.step(cf, mats, c, 1e-4, want_grid=True)
    # per node, truncation error <= one raw unit (2^-32) per contribution
    assert np.all(np.abs(ga["m"] - gc["m"]) <= 27 * 40 * 2.0 ** -32 + 1e-6 * np.abs(gc["m"]))


def test_velocity_only_free_fall_closed_form():
    cfg = R.make_config(16, 16, 16, 0.1, gravity=(0, -9.81, 0))
    pos = [0.83, 0.71, 0.62]
    st = R.State([pos], mass=[0.01])
    R.step_velocity_only(cfg, st, 1e-3)
    X = np.array(pos) / 0.1 - 0.5
    vy = -9.81e-3 * np.prod([_S(f) for f in X - np.floor(X)])
    assert np.allclose(st.vel[0], [0, vy * 0.995, 0], rtol=1e-5, atol=1e-9)


# ---------------------------------------------------------------------------
# The reference's own behavioural tests, run on the oracle
# ---------------------------------------------------------------------------
def test_reference_sandpile_stays_above_floor():
    """mpm_solver.rs:1612-1657 (mpm_step, DP sand, substeps 8, 60 frames)."""
    pos = np.array([[x * .05, y * .05 + .5, z * .05] for x in range(10) for y in range(10) for z in range(10)], f32)
    st = R.State(pos, mass=np.full(1000, 0.001))
    mats = R.Materials([[1.0e4, 0.2, 1.0e3, 0.005, 1600.0, 35.0, 0.0]], [R.DRUCKER_PRAGER])
    cfg = R.make_config(16, 16, 16, 1.6, substeps=8, clamp_mode=R.CLAMP_SERIAL)
    for _ in range(60):
        R.step(cfg, mats, st, 0.016)
    assert np.all(st.pos[:, 1] >= 0.0)


def test_reference_water_pressure_j_bounded():
    """mpm_solver.rs:1660-1711."""
    pos = np.array([[x * .05, y * .05 + .1, z * .05] for y in range(10) for x in range(5) for z in range(5)], f32)
    st = R.State(pos, mass=np.full(250, 0.001), J=np.full(250, 0.95))
    mats = R.Materials([[1.0e3, 0.499, 0, 0, 1000.0, 0, 0]], [R.NEWTONIAN_FLUID])
    cfg = R.make_config(32, 32, 32, 0.05, substeps=8, clamp_mode=R.CLAMP_SERIAL)
    R.step(cfg, mats, st, 0.016)
    assert np.all(np.isfinite(st.J)) and np.all((st.J > 0) & (st.J < 100))


def test_reference_steel_falls():
    """mpm_solver.rs:1714-1745."""
    st = R.State([[0.2, 0.5, 0.2]], mass=[0.01])
    mats = R.Materials([[1.0e5, 0.3, 1.0e4, 0.01, 7850.0, 0, 0]], [R.VON_MISES])
    cfg = R.make_config(32, 32, 32, 0.02, substeps=8, clamp_mode=R.CLAMP_SERIAL)
    for _ in range(30):
        R.step(cfg, mats, st, 0.016)
    assert st.pos[0, 1] < 0.5 and st.pos[0, 1] >= 0.0
