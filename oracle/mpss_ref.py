"""CPU restatement of the reference's MpssBuffer slot lifecycle.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the resident
buffer (aemh_mpss_*); never by the product path.

Follows ae_particle/src/mpss.rs:
  new()              :161-187   all slots inactive; strain and C zero,
                                jacobian and mass 1, lifetime f32::MAX
  spawn()            :189-203   free list popped last-in first-out, then
                                slot `count`; None when full
  kill()             :205-212   only active, in-range slots
  reset_slot()       :215-232   pos/vel 0, strain identity, jacobian 1,
                                mass 1, material 0, lifetime MAX, age 0;
                                C is left as it was
  active_indices()   :240-242
  compact()          :283-297   the literal swap sequence (swap_slots,
                                :299-320, swaps every field incl. active)
  update_lifetimes() :431-445   ages active slots below `count` only;
                                kills the expired ones in ascending order
The full step wrappers close with update_lifetimes(total_dt)
(mpm_solver.rs:489, :508, :1083); the serial substep visits the active
slots below `count` (mpm_solver.rs:187-193).

J_p (the snow model's plastic ratio, an addition of this implementation)
is kept per slot like the other fields and restarts at 1 on spawn.

Pinned by the reference's own unit tests (mpss.rs:571-660,
ae_particle/tests/integration_tests.rs:80-135), restated in
tests/test_mpss.py, and by brute-force invariants there.
"""
from __future__ import annotations

import numpy as np

from . import refcpu as R

F32_MAX = np.float32(np.finfo(np.float32).max)
FIELDS = ("pos", "vel", "C", "F", "J", "Jp", "mass", "mat", "lifetime", "age", "active")


class MpssBuffer:
    def __init__(self, capacity: int):
        n = capacity
        self.capacity = n
        self.count = 0
        self.pos = np.zeros((n, 3), np.float32)
        self.vel = np.zeros((n, 3), np.float32)
        self.F = np.zeros((n, 9), np.float32)  # strain: zeros in new()
        self.J = np.ones(n, np.float32)
        self.C = np.zeros((n, 9), np.float32)
        self.mass = np.ones(n, np.float32)
        self.mat = np.zeros(n, np.uint16)
        self.Jp = np.ones(n, np.float32)
        self.lifetime = np.full(n, F32_MAX, np.float32)
        self.age = np.zeros(n, np.float32)
        self.active = np.zeros(n, bool)
        self.free_list: list[int] = []

    # mpss.rs:189-203
    def spawn(self):
        if self.free_list:
            idx = self.free_list.pop()
        elif self.count < self.capacity:
            idx = self.count
        else:
            return None
        self._reset_slot(idx)
        self.active[idx] = True
        self.count += 1
        return idx

    # mpss.rs:205-212
    def kill(self, idx: int):
        if 0 <= idx < self.capacity and self.active[idx]:
            self.active[idx] = False
            self.free_list.append(idx)
            self.count -= 1

    # mpss.rs:215-232
    def _reset_slot(self, idx: int):
        self.pos[idx] = 0.0
        self.vel[idx] = 0.0
        self.F[idx] = np.eye(3, dtype=np.float32).reshape(9)
        self.J[idx] = 1.0
        self.mass[idx] = 1.0
        self.mat[idx] = 0
        self.lifetime[idx] = F32_MAX
        self.age[idx] = 0.0
        self.Jp[idx] = 1.0

    def len(self) -> int:
        return self.count

    # mpss.rs:240-242
    def active_indices(self) -> np.ndarray:
        return np.flatnonzero(self.active).astype(np.uint32)

    # mpss.rs:299-320
    def _swap_slots(self, a: int, b: int):
        for f in FIELDS:
            arr = getattr(self, f)
            tmp = arr[a].copy()
            arr[a] = arr[b]
            arr[b] = tmp

    # mpss.rs:283-297
    def compact(self):
        write = 0
        for read in range(self.capacity):
            if self.active[read]:
                if write != read:
                    self._swap_slots(write, read)
                write += 1
        self.count = write
        self.free_list.clear()
        self.active[write:] = False

    # mpss.rs:431-445 (f32 age arithmetic)
    def update_lifetimes(self, dt: float) -> list[int]:
        expired = []
        dt32 = np.float32(dt)
        for i in range(self.count):
            if self.active[i]:
                self.age[i] = np.float32(self.age[i] + dt32)
                if self.age[i] >= self.lifetime[i]:
                    expired.append(i)
        for i in expired:
            self.kill(i)
        return expired

    # -- steps over the resident slots -------------------------------------------
    def serial_indices(self) -> np.ndarray:
        """Slots the serial substep visits (mpm_solver.rs:187-193)."""
        n = self.count
        return np.flatnonzero(self.active[:n]).astype(np.uint32)

    def _state(self) -> R.State:
        return R.State(self.pos, self.vel, self.C, self.F, self.J, self.mass, self.mat, self.Jp)

    def _take(self, st: R.State):
        for f in ("pos", "vel", "C", "F", "J", "Jp"):
            setattr(self, f, getattr(st, f))

    def step(self, cfg, mats, total_dt, idx=None):
        """mpm_step_with_grid (idx None: serial set) or
        mpm_step_parallel_with_grid (explicit idx), then update_lifetimes."""
        if idx is not None and len(idx) == 0:
            return []  # mpm_solver.rs:1072-1074 returns before ageing
        st = self._state()
        sel = self.serial_indices() if idx is None else np.asarray(idx, np.uint32)
        if sel.size:
            R.step(cfg, mats, st, total_dt, idx=sel)
        self._take(st)
        return self.update_lifetimes(total_dt)
