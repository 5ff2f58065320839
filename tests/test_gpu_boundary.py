"""The drop-in boundary as the engine drives it, call by call.

The engine owns an MpssBuffer on the host and, every tick, spawns and kills
particles, writes fields, steps the near tier with
mpm_step_parallel_with_grid(buffer, grid, config, material_table,
model_table, dt, particle_indices) (ae_engine/src/managers/simulation.rs:
573-581), then clamps positions with apply_boundary_conditions_par
(ae_engine/src/lib.rs:346-349, ae_particle/src/mpss.rs:449-476).

`PerCallStep` below is the ctypes mirror of the Rust binding's
MpmGpu::mpm_step_parallel_with_grid (ae-engine_amd/ae_mpm_hip/src/lib.rs):
each call writes the listed slots from the host buffer, steps, reads them
back and ages lifetimes on the host.  The same tick sequence runs on a copy
of the buffer through the serial FP32 oracle; every field must agree bit for
bit after every tick (the binding picks the reference-order mode below
50,000 indices, as the reference takes its serial path).
"""
import numpy as np
import pytest

import aemh
from oracle import refcpu as R

pytestmark = pytest.mark.gpu

CAP = 4000


class HostBuffer:
    """The MpssBuffer arrays the step touches (mpss.rs:104-157) and the
    spawn / kill / update_lifetimes bookkeeping (:189-212, 431-445)."""

    def __init__(self, cap):
        self.cap, self.count = cap, 0
        self.pos = np.zeros((cap, 3), np.float32)
        self.vel = np.zeros((cap, 3), np.float32)
        self.F = np.zeros((cap, 9), np.float32)
        self.J = np.ones(cap, np.float32)
        self.C = np.zeros((cap, 9), np.float32)
        self.mass = np.ones(cap, np.float32)
        self.mat = np.zeros(cap, np.uint16)
        self.lifetime = np.full(cap, np.finfo(np.float32).max, np.float32)
        self.age = np.zeros(cap, np.float32)
        self.active = np.zeros(cap, bool)
        self.free = []

    def spawn(self):
        if self.free:
            i = self.free.pop()
        elif self.count < self.cap:
            i = self.count
            self.count += 1
        else:
            return None
        self.pos[i] = 0
        self.vel[i] = 0
        self.F[i] = np.eye(3, dtype=np.float32).reshape(9)
        self.J[i] = 1
        self.mass[i] = 1
        self.mat[i] = 0
        self.lifetime[i] = np.finfo(np.float32).max
        self.age[i] = 0
        self.active[i] = True
        return i

    def kill(self, i):
        if i < self.cap and self.active[i]:
            self.active[i] = False
            self.free.append(i)

    def update_lifetimes(self, dt):
        for i in range(self.count):
            if self.active[i]:
                self.age[i] += np.float32(dt)
                if self.age[i] >= self.lifetime[i]:
                    self.kill(i)

    def apply_boundary_conditions(self, lo, hi):
        for d in range(3):
            a = self.active[:self.count]
            p, v = self.pos[:self.count, d], self.vel[:self.count, d]
            below, above = a & (p < lo[d]), a & (p > hi[d])
            p[below] = lo[d]
            v[below & (v < 0)] *= np.float32(-0.5)
            p[above] = hi[d]
            v[above & (v > 0)] *= np.float32(-0.5)


class PerCallStep:
    """ctypes mirror of MpmGpu (ae_mpm_hip/src/lib.rs)."""

    def __init__(self):
        self.eng = aemh.Engine()
        self.capacity = 0

    def step(self, b, grid, gravity, damping, params, models, dt, idx):
        if len(idx) == 0:
            return
        cfg = aemh.default_config()
        cfg.gravity[:] = gravity
        cfg.damping = damping
        if len(idx) < 50000:
            cfg.accum = aemh.ACCUM_F32_SERIAL
        self.eng.set_config(cfg)
        self.eng.set_grid(*grid)
        n = max(len(params), len(models))
        p7 = np.array([params[i] if i < len(params) else params[0] for i in range(n)], np.float32)
        md = np.array([models[i] if i < len(models) else aemh.NEO_HOOKEAN for i in range(n)], np.uint8)
        self.eng.set_materials([aemh.Material(*[float(x) for x in p7[i]], model=int(md[i]))
                                for i in range(n)])
        if self.capacity != b.cap:
            self.eng.mpss_reserve(b.cap)
            self.capacity = b.cap
        idx = np.asarray(idx, np.uint32)
        self.eng.mpss_write(idx, pos=b.pos[idx], vel=b.vel[idx], C=b.C[idx], F=b.F[idx], J=b.J[idx],
                            mass=b.mass[idx], mat=b.mat[idx])
        self.eng.mpm_step_parallel_with_grid(idx, dt)
        r = self.eng.mpss_read(idx, fields=("pos", "vel", "C", "F", "J"))
        b.pos[idx], b.vel[idx], b.C[idx], b.F[idx], b.J[idx] = r["pos"], r["vel"], r["C"], r["F"], r["J"]
        b.update_lifetimes(dt)

    def close(self):
        self.eng.close()


def oracle_step(b, grid, gravity, damping, params, models, dt, idx):
    nx, ny, nz, dx, o = grid
    cfg = R.make_config(nx, ny, nz, dx, o, 1, gravity, damping, 0, 0.0, R.ACCUM_F32, R.CLAMP_AUTO,
                        det_math=0, threads=1)
    n = max(len(params), len(models))
    M = R.Materials([params[i] if i < len(params) else params[0] for i in range(n)],
                    [models[i] if i < len(models) else R.NEO_HOOKEAN for i in range(n)])
    st = R.State(b.pos, vel=b.vel, C9=b.C, F9=b.F, J=b.J, mass=b.mass, mat=b.mat)
    R.step(cfg, M, st, dt, idx=np.asarray(idx, np.uint32))
    b.pos[:], b.vel[:], b.C[:], b.F[:], b.J[:] = st.pos, st.vel, st.C, st.F, st.J
    b.update_lifetimes(dt)


def game_tick_edits(b, rng):
    """What the game does to the buffer between steps: spawns with field
    writes, kills, and velocity edits of live particles."""
    for _ in range(int(rng.integers(100, 400))):
        i = b.spawn()
        if i is None:
            break
        b.pos[i] = rng.uniform((-15, 2, -15), (15, 30, 15))
        b.vel[i] = rng.normal(0, 2.0, 3)
        b.mass[i] = np.float32(rng.uniform(0.5, 3.0))
        b.mat[i] = rng.integers(0, 4)
        b.lifetime[i] = np.float32(rng.uniform(0.05, 1.0))
    live = np.flatnonzero(b.active[:b.count])
    for i in rng.choice(live, size=min(30, live.size), replace=False):
        b.kill(int(i))
    live = np.flatnonzero(b.active[:b.count])
    edit = rng.choice(live, size=min(25, live.size), replace=False)
    b.vel[edit] += rng.normal(0, 1.0, (edit.size, 3)).astype(np.float32)


def test_engine_call_pattern_matches_oracle_tick_by_tick():
    # material_table has three entries, model_table two: index 2 is
    # Neo-Hookean with its own parameters, index 3 falls back to entry 0
    params = [[1e5, 0.3, 1e4, 0.01, 1000.0, 35.0, 0.0], [2e5, 0.45, 0, 0, 1000.0, 0, 0],
              [5e4, 0.25, 0, 0, 800.0, 0, 0]]
    models = [aemh.NEO_HOOKEAN, aemh.NEWTONIAN_FLUID]
    grid = (32, 32, 32, 1.6, (-25.6, 0.0, -25.6))
    gravity, damping, dt = (0.0, -9.81, 0.0), 0.995, 1.0 / 60.0
    lo, hi = (-20.0, 0.0, -20.0), (20.0, 40.0, 20.0)
    host, ref = HostBuffer(CAP), HostBuffer(CAP)
    rng_h, rng_r = np.random.default_rng(41), np.random.default_rng(41)
    gpu = PerCallStep()
    for tick in range(14):
        game_tick_edits(host, rng_h)
        game_tick_edits(ref, rng_r)
        idx = np.flatnonzero(host.active[:host.count]).astype(np.uint32)
        gpu.step(host, grid, gravity, damping, params, models, dt, idx)
        oracle_step(ref, grid, gravity, damping, params, models, dt, idx)
        host.apply_boundary_conditions(lo, hi)
        ref.apply_boundary_conditions(lo, hi)
        for f in ("pos", "vel", "C", "F", "J", "age", "active"):
            a, b = getattr(host, f), getattr(ref, f)
            assert np.array_equal(a, b), f"tick {tick}: {f} differs in {int((a != b).sum())} values"
        assert host.free == ref.free
    # the buffer grew, particles expired and slots were reused
    assert host.count > 1000 and len(host.free) > 0
    gpu.close()
