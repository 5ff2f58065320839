"""Writes the golden fixtures of this directory from the oracle.

Each fixture is an .npz of inputs (seeded positions and the set-up),
expected outputs after a number of steps, and the configuration, made by
oracle/refcpu.cpp (single-threaded, so the float summation order is the
reference's serial order, mpm_solver.rs:831-903).  The reference itself is
Rust and cannot be run here; these vectors freeze the oracle's answers so a
later change to it, or a GPU run on a box without the build tree, is checked
against the same numbers.

    python tests/golden/make_golden.py [name ...]   (default: every fixture)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "ae-engine_amd"))

from oracle import refcpu as R  # noqa: E402

# Material rows: MaterialParams field order (constitutive.rs:31-39):
# young, poisson, yield_stress, hardening, density, friction_angle (deg), cohesion
NH = [1e4, 0.2, 0.0, 0.0, 1.0, 0.0, 0.0]
SNOW = [1.4e5, 0.2, 0.0, 0.0, 400.0, 0.0, 0.0]
SNOW_EXT = [2.5e-2, 7.5e-3, 10.0]


def uniform(seed, n, lo, hi):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, (n, 3)).astype(np.float32)


ONLY = set(sys.argv[1:])


def case(name, pos, mass, mat, params, models, ext, grid, dt, steps, accum, vel=None, F0=None,
         velocity_only=False, bc_margin=3, gravity=(0.0, -9.8, 0.0), J0=None):
    if ONLY and name not in ONLY:
        return
    nx, ny, nz, dx, origin = grid
    cfg = R.make_config(nx, ny, nz, dx, origin, 1, gravity, 0.995, bc_margin, 0.0, accum,
                        R.CLAMP_AUTO, threads=1)
    M = R.Materials(np.array(params, np.float32), np.array(models, np.uint8),
                    np.array(ext, np.float32))
    n = pos.shape[0]
    F0 = np.tile(np.eye(3, dtype=np.float32).reshape(9), (n, 1)) if F0 is None else F0
    if J0 is None:
        J0 = np.linalg.det(F0.reshape(-1, 3, 3).astype(np.float64)).astype(np.float32)
    st = R.State(pos, vel=vel, F9=F0, J=J0, mass=mass, mat=mat)
    for _ in range(steps):
        if velocity_only:
            R.step_velocity_only(cfg, st, dt)
        else:
            R.step(cfg, M, st, dt)
    np.savez_compressed(
        os.path.join(HERE, name + ".npz"),
        pos0=pos, vel0=np.zeros_like(pos) if vel is None else vel, F0=F0, J0=J0, mass=mass,
        mat=mat,
        params=np.array(params, np.float32), models=np.array(models, np.uint8),
        ext=np.array(ext, np.float32), grid_n=np.array([nx, ny, nz], np.uint32),
        dx=np.float32(dx), origin=np.array(origin, np.float32), dt=np.float32(dt),
        steps=np.uint32(steps), accum=np.uint32(accum), velocity_only=np.uint32(velocity_only),
        bc_margin=np.uint32(bc_margin), gravity=np.array(gravity, np.float32),
        pos=st.pos, vel=st.vel, C=st.C, F=st.F, J=st.J, Jp=st.Jp)
    print(name, "written")


def main():
    n = 512
    pos = uniform(1, n, 0.3, 0.6)
    mass = np.full(n, 0.3 ** 3 / n, np.float32)
    mat = np.zeros(n, np.uint16)
    grid = (16, 16, 16, 1.0 / 16, (0.0, 0.0, 0.0))
    for accum, tag in ((R.ACCUM_Q32, "q32"), (R.ACCUM_F32, "f32")):
        case(f"cube512_nh_{tag}_10", pos, mass, mat, [NH], [R.NEO_HOOKEAN], [SNOW_EXT], grid,
             1e-4, 10, accum)
    # velocity-only tier with random velocities
    rng = np.random.default_rng(3)
    vel = rng.normal(0, 0.5, (n, 3)).astype(np.float32)
    # snow block with a perturbed F (singular values outside the window) and
    # random velocities: the clamp and J_p move
    pos_s = uniform(2, n, (0.3, 0.2, 0.3), (0.55, 0.6, 0.55))
    mass_s = np.full(n, 400.0 * 0.25 * 0.4 * 0.25 / n, np.float32)
    F0 = (np.tile(np.eye(3).reshape(9), (n, 1)) + rng.normal(0, 0.02, (n, 9))).astype(np.float32)
    case("snow512_q32_20", pos_s, mass_s, mat, [SNOW], [7], [SNOW_EXT], grid, 5e-5, 20,
         R.ACCUM_Q32, vel=vel, F0=F0)
    case("velonly512_q32_10", pos, mass, mat, [NH], [R.NEO_HOOKEAN], [SNOW_EXT], grid, 1e-4, 10,
         R.ACCUM_Q32, vel=vel, velocity_only=True)
    # snow with ill-conditioned F: a third of the particles at cond(F) 1e4,
    # 1e6 and 1e8 (F = R1 diag(1.5, 1, 1.5 / cond) R2^T, stored J = 1), so
    # the snow clamp's SVD takes its FP64 path
    F0 = ill_conditioned_f(n, 4)
    case("snow_ill512_q32_5", pos_s, mass_s, mat, [SNOW], [7], [SNOW_EXT], grid, 5e-5, 5,
         R.ACCUM_Q32, F0=F0, J0=np.ones(n, np.float32))


def rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def ill_conditioned_f(n, seed):
    """Identity F, except every third particle: F = R1 diag(1.5, 1, 1.5 / c) R2^T
    with c cycling through 1e4, 1e6, 1e8."""
    rng = np.random.default_rng(seed)
    F = np.tile(np.eye(3), (n, 1, 1))
    for i in range(0, n, 3):
        c = (1e4, 1e6, 1e8)[(i // 3) % 3]
        F[i] = rotation(rng) @ np.diag([1.5, 1.0, 1.5 / c]) @ rotation(rng).T
    return F.reshape(n, 9).astype(np.float32)


if __name__ == "__main__":
    main()
