"""Closed-form pins of the reference's own MPM formulas -- the ones that
differ from textbook MLS-MPM -- on single particles, against the oracles.

FP32 solver (ae_particle/src/mpm_solver.rs, oracle/refcpu.cpp):
  * P2G force with the 1-D derivative only, f_r += vol sigma[r][c] N'(d_c - f_c)
    inv_dx (:276-285), and the grid update v = p/m + (g + f/m) dt (:121-129);
  * G2P grad v[r][c] -= w v_r N'(d_c - f_c) inv_dx (full w times the 1-D
    derivative, :343-354) and F' = (I + dt grad v) F (:401-429).
FixedPoint solver (ae_physics/src/mpm.rs, oracle/fixedsim.cpp):
  * the momentum's stress term m (v + S rel / cell^2) (:316-324);
  * the velocity gradient accumulated into column a = velocity component,
    row b = offset direction (:384-402), and F + (grad v F) dt (:416-419);
  * update_deformation's Kirchhoff stress (mu (F - F^-T) + lambda ln J F^-T)
    F^T / J (:429-454).

Expected values come from the B-spline's definition and the mathematics of
one particle (the stencil sums factor per axis), in float64.  The last test
shows the pins bite: each pin runs again in a subprocess against a
throw-away oracle build with one planted mistake (a textbook gradient, a
transposed gradient, F multiplied from the wrong side, a dropped update
term, a transposed stress or deformation product) and must fail there.
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import fixedsim as X
from oracle import refcpu as R

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
f32 = np.float32


def N(x):
    """Quadratic B-spline (mpm_solver.rs:442-452; the reference's known
    answers N(0) = .75, N(.5) = .5, N(1) = .125, N(1.5) = 0)."""
    a = abs(x)
    return 0.75 - a * a if a < 0.5 else (0.5 * (1.5 - a) ** 2 if a < 1.5 else 0.0)


def Np(x):
    """Its derivative (N'(0) = 0, N'(.5) = -1, N'(1) = -.5)."""
    a = abs(x)
    return -2.0 * x if a < 0.5 else (-(1.5 - a) * math.copysign(1.0, x) if a < 1.5 else 0.0)


# ---------------------------------------------------------------------------
# FP32 solver
# ---------------------------------------------------------------------------
DX = 1.0 / 32.0


def _particle(base, fx):
    """Position of a particle with the given base node and fraction per axis
    (cell-centred nodes: x / dx - 0.5 = base + fx), exact in float32."""
    return (np.asarray(base, np.float64) + 0.5 + np.asarray(fx, np.float64)) * DX


def _fluid(E=1e3, nu=0.2, rho=1000.0):
    return R.Materials([[E, nu, 0, 0, rho, 0, 0]], [R.NEWTONIAN_FLUID])


def _gid(i, j, k):
    return i + j * 32 + k * 32 * 32


def test_q3_force_is_one_dimensional_derivative_and_update_term():
    """Fluid at J = 0.9: sigma = K (J - 1) I, so the grid force at stencil
    node (i, j, k) is vol K (J - 1) (N'(i - fx), N'(j - fy), N'(k - fz)) / dx
    -- no transverse weights -- and the updated node velocity is
    (g + f / (w m)) dt."""
    E, nu, rho, mass, J, dt = 1e3, 0.2, 1000.0, 0.01, 0.9, 1e-4
    K = E / (3.0 * (1.0 - 2.0 * nu))
    g = (0.0, -9.81, 0.0)
    cfg = R.make_config(32, 32, 32, DX, gravity=g, damping=1.0)
    for base, fx in (((12, 15, 9), (0.3, 0.7, 0.55)), ((8, 20, 14), (0.62, 0.18, 0.41))):
        pos = _particle(base, fx).astype(f32)
        fxe = pos.astype(np.float64) / DX - 0.5 - np.asarray(base)
        st = R.State([pos], J=[J], mass=[mass])
        grid = R.step(cfg, _fluid(E, nu, rho), st, dt, want_grid=True)
        vsig = mass / rho * K * (J - 1.0)
        want_f = np.zeros((32 ** 3, 3))
        want_v = np.zeros((32 ** 3, 3))
        for dk in range(3):
            for dj in range(3):
                for di in range(3):
                    d = (di, dj, dk)
                    w = np.prod([N(d[a] - fxe[a]) for a in range(3)])
                    if w < 1e-10:
                        continue
                    gid = _gid(base[0] + di, base[1] + dj, base[2] + dk)
                    want_f[gid] = [vsig * Np(d[a] - fxe[a]) / DX for a in range(3)]
                    want_v[gid] = [(g[a] + want_f[gid][a] / (w * mass)) * dt for a in range(3)]
        got_f = grid["f"].astype(np.float64)
        assert np.abs(want_f).max() > 0
        assert np.allclose(got_f, want_f, rtol=2e-5, atol=1e-6 * np.abs(want_f).max()), \
            np.abs(got_f - want_f).max()
        assert np.allclose(grid["v"], want_v, rtol=2e-5, atol=1e-6 * np.abs(want_v).max())


def test_q4_grad_v_and_q9_f_update():
    """sigma = 0 (fluid with stored J = 1) and velocity u: every stencil node
    gets v = u, so G2P returns v = u S_x S_y S_z and
    grad v[r][c] = -u_r / dx * sum_d N(d - f_c) N'(d - f_c) * prod_{a != c} S_a
    (S_a = sum_d N(d - f_a), the truncated stencil's partition sum); then
    F' = (I + dt grad v) F on a non-symmetric F, J' = det F'."""
    dt, mass = 1e-3, 0.01
    u = np.array([0.3, -0.1, 0.2])
    F0 = np.array([[1.1, 0.2, 0.0], [0.0, 0.9, 0.1], [0.05, 0.0, 1.0]])
    cfg = R.make_config(32, 32, 32, DX, gravity=(0, 0, 0), damping=1.0)
    for base, fx in (((12, 15, 9), (0.3, 0.7, 0.55)), ((10, 11, 17), (0.8, 0.25, 0.6))):
        pos = _particle(base, fx).astype(f32)
        fxe = pos.astype(np.float64) / DX - 0.5 - np.asarray(base)
        st = R.State([pos], vel=[u.astype(f32)], F9=[F0.reshape(9).astype(f32)], J=[1.0],
                     mass=[mass])
        R.step(cfg, _fluid(), st, dt)
        S = np.array([sum(N(d - f) for d in range(3)) for f in fxe])
        P = np.array([sum(N(d - f) * Np(d - f) for d in range(3)) for f in fxe])
        G = np.array([P[c] / DX * np.prod([S[a] for a in range(3) if a != c]) for c in range(3)])
        gv = -np.outer(u, G)
        F1 = (np.eye(3) + dt * gv) @ F0
        v1 = u * np.prod(S)
        assert np.allclose(st.vel[0], v1, rtol=1e-5, atol=1e-7)
        assert np.abs(st.F[0].reshape(3, 3) - F1).max() < 5e-7, (st.F[0].reshape(3, 3), F1)
        assert abs(st.J[0] - np.linalg.det(F1)) < 1e-6
        assert np.allclose(st.pos[0], pos + v1 * dt, atol=1e-7)


# ---------------------------------------------------------------------------
# FixedPoint solver
# ---------------------------------------------------------------------------
CELL = 0.5
Q_POS, Q_VEL, Q_F, Q_S, Q_MASS, Q_ACC = slice(0, 3), slice(3, 6), slice(6, 15), slice(15, 24), 24, 28


def _fx_config():
    return X.small_config(res=(8, 8, 8), particle_count=0, model=X.ELASTIC, E=X.from_f32(50.0),
                          nu=X.from_f32(0.3), gravity=(0, 0, 0))


def _cols(m):
    """3x3 float rows -> column-major raw Q32.32."""
    return [X.from_f32(m[r][c]) for c in range(3) for r in range(3)]


def _rows(q9):
    return np.array(q9, np.float64).reshape(3, 3).T / 2 ** 32


def _fx_one(pos, vel=(0.0, 0.0, 0.0), F=None, S=None, mass=0.5):
    """One substep of one particle whose F and stress were written first;
    returns (position raw, the particle's record after the substep)."""
    sim = X.Sim(_fx_config())
    p = [X.from_f32(v) for v in pos]
    sim.add_particle(p, [X.from_f32(v) for v in vel], X.from_f32(mass))
    r = sim.read()
    q = r["q"].copy()
    if F is not None:
        q[0, Q_F] = _cols(F)
    if S is not None:
        q[0, Q_S] = _cols(S)
    sim.write(0, q, r["flags"], r["ids"], r["color"])
    sim.step(1)
    out = sim.read()["q"][0]
    sim.close()
    return p, out


def _fx_moments(p_raw):
    """fx per axis from the raw position (cell 0.5: x / cell is exact), and the
    stencil sums of the {-1, 0, 1} stencil (mpm.rs:295-312): S_a = sum N(f - d),
    M_a = sum N(f - d) (d - f) prod_{b != a} S_b."""
    fx = np.array([((2 * r) & 0xFFFFFFFF) / 2 ** 32 for r in p_raw])
    S = np.array([sum(N(f - d) for d in (-1, 0, 1)) for f in fx])
    m1 = np.array([sum(N(f - d) * (d - f) for d in (-1, 0, 1)) for f in fx])
    M = np.array([m1[a] * np.prod([S[b] for b in range(3) if b != a]) for a in range(3)])
    return fx, S, M


POS = (1.65, 2.35, 1.775)  # base (3, 4, 3), fractions 0.3, 0.7, 0.55


def test_fx_momentum_stress_term_non_symmetric():
    """A resting particle with a non-symmetric stress S: node momentum
    m w S rel / cell^2 with rel = (d - f) cell, so each node's velocity is
    S (d - f) / cell and G2P returns v = S M / cell."""
    Sm = np.array([[2.0, 1.5, 0.0], [-0.5, 1.0, 0.25], [0.75, 0.0, -1.0]])
    p, q = _fx_one(POS, S=Sm)
    _, _, M = _fx_moments(p)
    want = Sm @ M / CELL
    got = np.array(q[Q_VEL], np.float64) / 2 ** 32
    assert np.abs(want).max() > 0.05
    assert np.abs(got - want).max() < 1e-6, (got, want)


def test_fx_velocity_gradient_column_layout_and_f_update():
    """Velocity u, stress 0: every node's velocity is u, so v = u W and the
    gradient has vg[b][a] = u_a M_b / cell (column a = velocity component);
    F' = I + dt vg, and the accumulated strain grows by |1 - det vg| = 1
    (vg has rank 1)."""
    u = np.array([0.25, -0.1, 0.15])
    p, q = _fx_one(POS, vel=u)
    fx, S, M = _fx_moments(p)
    dt = 1.0 / 60.0
    F1 = np.eye(3) + dt * np.outer(M, u) / CELL
    got_F = _rows(q[Q_F])
    assert np.abs(got_F - F1).max() < 1e-6, (got_F, F1)
    assert np.abs(np.array(q[Q_VEL]) / 2 ** 32 - u * np.prod(S)).max() < 1e-7
    assert abs(q[Q_ACC] / 2 ** 32 - 1.0) < 1e-6


def _kirchhoff(F, E=50.0, nu=0.3):
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    J = np.linalg.det(F)
    jc = min(max(J, 0.01), 100.0)
    Fit = np.linalg.inv(F).T
    return (mu * (F - Fit) + lam * math.log(jc) * Fit) @ F.T / jc


@pytest.mark.parametrize("F", [np.diag([1.2, 0.9, 1.05]),
                               np.array([[1.1, 0.15, 0.0], [0.05, 0.95, 0.1], [0.0, 0.02, 1.02]])],
                         ids=["diagonal", "general"])
def test_fx_update_deformation_kirchhoff(F):
    """A resting particle with zero stress and a given F: the grid stays at
    rest, F is unchanged, and update_deformation leaves
    sigma = (mu (F - F^-T) + lambda ln J F^-T) F^T / J; on diag(s) that is
    diag(mu (s^2 - 1) + lambda ln J) / J."""
    _, q = _fx_one(POS, F=F)
    F = _rows(_cols(F))  # the values written (from_f32 of float32)
    assert np.array_equal(_rows(q[Q_F]), F)
    want = _kirchhoff(F)
    if np.count_nonzero(F - np.diag(np.diag(F))) == 0:
        s = np.diag(F)
        J = float(np.prod(s))
        mu, lam = 50.0 / 2.6, 50.0 * 0.3 / (1.3 * 0.4)
        assert np.allclose(want, np.diag((mu * (s * s - 1) + lam * math.log(J)) / J), atol=1e-12)
    got = _rows(q[Q_S])
    assert np.abs(got - want).max() < 2e-6 * max(1.0, np.abs(want).max()), (got, want)


def test_fx_update_deformation_of_the_deformed_particle():
    """The same stress on the non-symmetric F' the moving particle above
    reaches in one substep (read back from the oracle)."""
    p, q = _fx_one(POS, vel=(0.25, -0.1, 0.15))
    F1 = _rows(q[Q_F])
    want = _kirchhoff(F1)
    assert np.abs(_rows(q[Q_S]) - want).max() < 2e-6 * max(1.0, np.abs(want).max())


# ---------------------------------------------------------------------------
# The pins bite: planted mistakes in throw-away oracle builds
# ---------------------------------------------------------------------------
MUTATIONS = [
    ("refcpu", "TEXTBOOK_P2G", "test_q3_force_is_one_dimensional_derivative_and_update_term"),
    ("refcpu", "NO_FORCE_TERM", "test_q3_force_is_one_dimensional_derivative_and_update_term"),
    ("refcpu", "TEXTBOOK_G2P", "test_q4_grad_v_and_q9_f_update"),
    ("refcpu", "TRANSPOSE_GV", "test_q4_grad_v_and_q9_f_update"),
    ("refcpu", "F_RIGHT", "test_q4_grad_v_and_q9_f_update"),
    ("fixedsim", "FX_STRESS_T", "test_fx_momentum_stress_term_non_symmetric"),
    ("fixedsim", "FX_VGRAD_ROWS", "test_fx_velocity_gradient_column_layout_and_f_update"),
    ("fixedsim", "FX_KIRCHHOFF_F", "test_fx_update_deformation_kirchhoff"),
]


@pytest.mark.parametrize("lib,flag,test", MUTATIONS, ids=[m[1] for m in MUTATIONS])
def test_pin_fails_on_planted_mistake(tmp_path, lib, flag, test):
    so = tmp_path / f"lib{lib}.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off",
                    "-march=x86-64-v3", "-fno-math-errno", f"-DORACLE_MUT_{flag}",
                    os.path.join(ROOT, "oracle", f"{lib}.cpp"), "-o", str(so)], check=True)
    env = dict(os.environ)
    env["AEMH_ORACLE_REFCPU" if lib == "refcpu" else "AEMH_ORACLE_FIXEDSIM"] = str(so)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        f"{os.path.join(ROOT, 'tests', 'test_oracle_pins.py')}::{test}"],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    # 1: the pin ran and failed (2 and above would be an error of the harness)
    assert r.returncode == 1, r.stdout[-2000:] + r.stderr[-2000:]
    assert "failed" in r.stdout
