"""CPU restatement of the engine's LOD tier tick over an MpssBuffer.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
aemh_tier_tick / aemh_tier_lists / aemh_tier_render; never by the product.

Follows the MPSS part of SimulationManager::update_fields_and_particles
(ae_engine/src/managers/simulation.rs:474-697):
  * field_step_counter += 1 every tick (:474); nothing else when the buffer
    is empty (:692-697 visit active slots only);
  * reclassify when |player - cached|^2 > 25, the near cache is empty, or
    lod_reclassify_counter % 8 == 0 (:488-494); near / mid / far by f32
    squared distance over the active slots below count (:500-517);
  * a near list longer than max_near keeps the max_near closest, the rest
    are appended to mid (:520-531).  The reference uses
    select_nth_unstable_by, which fixes the kept SET (up to ties at the cut)
    but not its order; this restatement (and the device) use a stable sort
    by distance, i.e. ties at the cut go to the lower slot;
  * moving-window origins player - nx dx / 2 (:540-556);
  * near: mpm_step_parallel_with_grid(dt, near) (:563-575), which ends with
    update_lifetimes(dt) (mpm_solver.rs:1083);
  * mid: every MID_INTERVAL ticks, mpm_step_velocity_only_parallel with
    dt x MID_INTERVAL on the mid grid (:586-603);
  * far: every far_field_update_interval ticks, v_y -= 9.81 far_dt,
    y += v_y far_dt (:604-690; the coarse temperature blend is not part of
    the path);
  * update_lifetimes(dt) (:692).
Render readback (ae_engine/src/lib.rs:1891-1936): list entries still live
(slot < count and active), in list order, from the first `take` entries.

Pinned in tests/test_tier.py by hand-worked scenes (classification counts,
cull set against brute force, reclassification and tier cadences, the far
update's closed form, double ageing on near ticks).
"""
from __future__ import annotations

import numpy as np

from . import refcpu as R
from .mpss_ref import MpssBuffer

f32 = np.float32


def default_lod():
    near = f32(50.0)
    near_dx = f32(1.6)
    n = int(np.ceil(f32(f32(2.0) * near) / near_dx))
    p = 1
    while p < n:
        p <<= 1
    mid = f32(200.0)
    return dict(near_field_distance=near, mid_field_distance=mid, far_field_update_interval=60,
                max_near_particles=10000, mid_interval=12, near_n=p, near_dx=near_dx, mid_n=16,
                mid_dx=f32(f32(2.0) * mid) / f32(16))


class TierManager:
    def __init__(self, buf: MpssBuffer, mats, lod=None, gravity=(0, -9.81, 0), damping=0.995,
                 accum=R.ACCUM_Q32, clamp=R.CLAMP_AUTO, substeps=1):
        self.buf = buf
        self.mats = mats
        self.lod = dict(default_lod(), **(lod or {}))
        self.gravity = gravity
        self.damping = damping
        self.accum = accum
        self.clamp = clamp
        self.substeps = substeps
        self.field_counter = 0
        self.lod_counter = 0
        self.cached_player = np.zeros(3, f32)
        self.near: list[int] = []
        self.mid: list[int] = []
        self.far: list[int] = []

    def _cfg(self, n, dx, origin):
        return R.make_config(n, n, n, float(dx), tuple(float(o) for o in origin), self.substeps,
                             self.gravity, self.damping, 0, 0.0, self.accum, self.clamp)

    def classify(self, player):
        b = self.buf
        L = self.lod
        near2 = f32(L["near_field_distance"]) * f32(L["near_field_distance"])
        mid2 = f32(L["mid_field_distance"]) * f32(L["mid_field_distance"])
        near, mid, far = [], [], []
        for i in range(b.count):
            if not b.active[i]:
                continue
            dx = f32(b.pos[i, 0] - player[0])
            dy = f32(b.pos[i, 1] - player[1])
            dz = f32(b.pos[i, 2] - player[2])
            d2 = f32(f32(f32(dx * dx) + f32(dy * dy)) + f32(dz * dz))
            if d2 < near2:
                near.append((i, d2))
            elif d2 < mid2:
                mid.append(i)
            else:
                far.append(i)
        max_near = int(L["max_near_particles"])
        if len(near) > max_near:
            near = sorted(near, key=lambda t: t[1])  # stable: ties keep slot order
            mid.extend(i for i, _ in near[max_near:])
            near = near[:max_near]
        self.near = [i for i, _ in near]
        self.mid = mid
        self.far = far

    def tick(self, player, dt):
        b = self.buf
        L = self.lod
        player = np.asarray(player, f32)
        dt = f32(dt)
        st = dict(near=len(self.near), mid=len(self.mid), far=len(self.far), reclassified=0,
                  near_ran=0, mid_ran=0, far_ran=0, expired=0)
        self.field_counter = (self.field_counter + 1) & 0xFFFFFFFF
        if b.count == 0:
            return dict(st, near=0, mid=0, far=0)
        m = (player - self.cached_player).astype(f32)
        moved = f32(f32(f32(m[0] * m[0]) + f32(m[1] * m[1])) + f32(m[2] * m[2])) > f32(25.0)
        self.lod_counter = (self.lod_counter + 1) & 0xFFFFFFFF
        if moved or not self.near or self.lod_counter % 8 == 0:
            self.cached_player = player.copy()
            self.classify(player)
            st["reclassified"] = 1
        near_half = f32(f32(f32(L["near_n"]) * f32(L["near_dx"])) * f32(0.5))
        mid_half = f32(f32(f32(L["mid_n"]) * f32(L["mid_dx"])) * f32(0.5))
        near_o = (player - near_half).astype(f32)
        mid_o = (player - mid_half).astype(f32)
        expired = []
        if self.near:
            cfg = self._cfg(L["near_n"], L["near_dx"], near_o)
            expired += b.step(cfg, self.mats, dt, idx=np.array(self.near, np.uint32))
            st["near_ran"] = 1
        if self.field_counter % L["mid_interval"] == 0 and self.mid:
            cfg = self._cfg(L["mid_n"], L["mid_dx"], mid_o)
            s = b._state()
            R.step_velocity_only(cfg, s, float(f32(dt * f32(L["mid_interval"]))),
                                 idx=np.array(self.mid, np.uint32))
            b._take(s)
            st["mid_ran"] = 1
        far_interval = max(int(L["far_field_update_interval"]), 1)
        if self.field_counter % far_interval == 0:
            far_dt = f32(dt * f32(far_interval))
            g = f32(f32(9.81) * far_dt)
            for i in self.far:
                v = f32(b.vel[i, 1] - g)
                b.vel[i, 1] = v
                b.pos[i, 1] = f32(b.pos[i, 1] + f32(v * far_dt))
            st["far_ran"] = 1
        expired += b.update_lifetimes(dt)
        st.update(near=len(self.near), mid=len(self.mid), far=len(self.far), expired=len(expired),
                  near_origin=tuple(near_o), mid_origin=tuple(mid_o))
        return st

    def render(self, tier, take):
        lst = (self.near, self.mid, self.far)[tier][:take]
        b = self.buf
        keep = [i for i in lst if i < b.count and b.active[i]]
        return b.pos[keep].copy(), np.array(keep, np.uint32)
