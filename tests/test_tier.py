"""LOD tier orchestrator: near / mid / far classification, near cull, tier
cadences, moving-window origins, far update, double ageing, render readback.

CPU: the oracle (oracle/tier_ref.py) on hand-worked scenes whose answers
follow from ae_engine/src/managers/simulation.rs:474-697 directly (counts,
the cull set against brute force, reclassification rules, cadences, the far
update's closed form).  GPU: aemh_tier_tick through the C ABI against the
oracle tick by tick, every field of every slot bit-exact (Q32.32 mode), the
cached lists, stats and render readback equal.
"""
import numpy as np
import pytest

import aemh
from oracle import refcpu as R
from oracle.mpss_ref import MpssBuffer as RefBuf
from oracle.tier_ref import TierManager, default_lod

f32 = np.float32

# The engine's material kinds (simulation.rs:243-294: elastic, fluid, sand)
# with E scaled so that every tier is CFL-stable at dt = 1/60 s, dx = 1.6 m
# (wave speed <= 26 m/s).  The engine's own values (sand E = 5e7 at
# rho = 1600: 177 m/s, CFL 1.8) blow up until F hits its 1e3 guard, and the
# stresses then leave the Q32.32 range of a node sum, which the deterministic
# mode reports as an overflow instead of saturating.
ENGINE_MATS = [
    aemh.Material(young_modulus=1e5, poisson_ratio=0.3, yield_stress=1e4, hardening=0.01,
                  density=1000.0, friction_angle=35.0, cohesion=0.0, model=aemh.NEO_HOOKEAN),
    aemh.Material(young_modulus=2e5, poisson_ratio=0.45, yield_stress=0.0, hardening=0.0,
                  density=1000.0, friction_angle=0.0, cohesion=0.0, model=aemh.NEWTONIAN_FLUID),
    aemh.Material(young_modulus=1e6, poisson_ratio=0.3, yield_stress=0.0, hardening=0.0,
                  density=1600.0, friction_angle=35.0, cohesion=0.0, model=aemh.DRUCKER_PRAGER),
]


def ref_mats():
    p, m, e = aemh.material_arrays(ENGINE_MATS)
    return R.Materials(p, m, e)


def spawn_at(buf, positions, lifetimes=None):
    slots = []
    for k, p in enumerate(positions):
        s = buf.spawn()
        buf.pos[s] = p
        if lifetimes is not None:
            buf.lifetime[s] = lifetimes[k]
        slots.append(s)
    return slots


# ---------------------------------------------------------------------------
# Oracle on hand-worked scenes
# ---------------------------------------------------------------------------
def test_default_lod_matches_engine_constants():
    L = default_lod()
    assert L["near_n"] == 64 and L["near_dx"] == f32(1.6)  # simulation.rs:219-227
    assert L["mid_n"] == 16 and L["mid_dx"] == f32(25.0)  # :234-235
    assert L["mid_interval"] == 12 and L["far_field_update_interval"] == 60
    assert L["max_near_particles"] == 10000
    c = aemh.default_lod_config()
    assert (c.near_n, c.near_dx, c.mid_n, c.mid_dx) == (64, f32(1.6), 16, f32(25.0))
    assert (c.near_field_distance, c.mid_field_distance) == (50.0, 200.0)
    assert (c.max_near_particles, c.mid_interval, c.far_field_update_interval) == (10000, 12, 60)


def test_classification_counts_and_order():
    # lod_debug's layout (game/src/bin/lod_debug.rs:16): near within 5 m,
    # mid at 100 m, far at 300 m; an inactive slot is skipped
    b = RefBuf(16)
    spawn_at(b, [(0, 5, 0), (100, 5, 0), (300, 5, 0), (3, 1, 2), (0, 5, 15), (0, 5, -150),
                 (49.9, 0, 0), (50, 0, 0), (199.9, 0, 0), (200, 0, 0)])
    b.kill(3)
    t = TierManager(b, ref_mats())
    t.classify(np.zeros(3, f32))
    # 49.9 m is near, exactly 50 m is mid (strict <); slot 9 (200 m) is not
    # visited: the scan runs over slots below count, 9 after the kill (:500)
    assert t.near == [0, 4, 6]
    assert t.mid == [1, 5, 7, 8]
    assert t.far == [2]
    b.compact()
    t.classify(np.zeros(3, f32))
    assert t.far == [2, 8]  # after compaction 200 m sits at slot 8 < count


def test_near_cull_keeps_closest_set():
    rng = np.random.default_rng(3)
    b = RefBuf(200)
    pos = rng.uniform(-40, 40, (120, 3)).astype(f32) * f32(0.5)
    pos[100:] = rng.uniform(60, 150, (20, 3)).astype(f32)  # mid
    spawn_at(b, pos)
    t = TierManager(b, ref_mats(), lod=dict(max_near_particles=30))
    t.classify(np.zeros(3, f32))
    d2 = np.array([f32(f32(f32(p[0] * p[0]) + f32(p[1] * p[1])) + f32(p[2] * p[2])) for p in pos])
    near_all = [i for i in range(120) if d2[i] < f32(2500)]
    brute = sorted(near_all, key=lambda i: (d2[i], i))[:30]
    assert sorted(t.near) == sorted(brute)
    assert len(t.near) == 30
    mid_base = [i for i in range(120) if f32(2500) <= d2[i] < f32(40000)]
    assert t.mid[:len(mid_base)] == mid_base
    assert sorted(t.mid[len(mid_base):]) == sorted(set(near_all) - set(brute))


def test_reclassify_rules_and_cadences():
    b = RefBuf(8)
    spawn_at(b, [(1, 0, 0), (100, 0, 0), (300, 0, 0)])
    t = TierManager(b, ref_mats(), lod=dict(far_field_update_interval=5))
    seen = []
    player = np.zeros(3, f32)
    for k in range(1, 25):
        if k == 10:
            player = np.array([4, 0, 0], f32)  # 16 m^2: not a move
        if k == 11:
            player = np.array([9.1, 0, 0], f32)  # 5.1 m from the cached origin
        st = t.tick(player, 1.0 / 60)
        seen.append((k, st["reclassified"], st["mid_ran"], st["far_ran"]))
    recl = [k for k, r, _, _ in seen if r]
    # tick 1: empty cache; every 8th tick; tick 11: moved > 5 m
    assert recl == [1, 8, 11, 16, 24]
    assert [k for k, _, m, _ in seen if m] == [12, 24]
    assert [k for k, _, _, f in seen if f] == [5, 10, 15, 20]


def test_far_update_closed_form():
    b = RefBuf(4)
    spawn_at(b, [(500, 40, 0)])
    b.vel[0] = (0, 2, 0)
    t = TierManager(b, ref_mats(), lod=dict(far_field_update_interval=3))
    dt = f32(1 / 60)
    for _ in range(3):
        t.tick(np.zeros(3, f32), dt)
    far_dt = f32(dt * f32(3))
    v = f32(f32(2) - f32(f32(9.81) * far_dt))
    assert b.vel[0, 1] == v
    assert b.pos[0, 1] == f32(f32(40) + f32(v * far_dt))
    assert b.pos[0, 0] == 500 and b.vel[0, 0] == 0


def test_near_ticks_age_twice():
    # mpm_step_parallel_with_grid ages the whole buffer, then the tick ages
    # it again: lifetime 4 dt lasts 2 ticks with a near tier, 4 without.
    dt = f32(0.25)
    b = RefBuf(8)
    spawn_at(b, [(1, 2, 1), (300, 2, 0)], lifetimes=[f32(1.0), f32(1.0)])
    t = TierManager(b, ref_mats())
    t.tick(np.zeros(3, f32), dt)
    assert list(b.age[:2]) == [0.5, 0.5]
    st = t.tick(np.zeros(3, f32), dt)
    assert st["expired"] == 2 and b.count == 0
    b2 = RefBuf(8)
    spawn_at(b2, [(300, 2, 0)], lifetimes=[f32(1.0)])
    t2 = TierManager(b2, ref_mats())
    for _ in range(3):
        t2.tick(np.zeros(3, f32), dt)
    assert b2.count == 1 and b2.age[0] == 0.75
    assert t2.tick(np.zeros(3, f32), dt)["expired"] == 1


def test_moving_window_origins():
    b = RefBuf(4)
    spawn_at(b, [(10, 1, 10)])
    t = TierManager(b, ref_mats())
    st = t.tick(np.array([10, 5, 10], f32), 1 / 60)
    near_half = f32(f32(64) * f32(1.6)) * f32(0.5)
    assert st["near_origin"] == tuple(np.array([10, 5, 10], f32) - near_half)
    assert st["mid_origin"] == (f32(-190), f32(-195), f32(-190))


# ---------------------------------------------------------------------------
# GPU against the oracle
# ---------------------------------------------------------------------------
def _scene(rng, n):
    r = rng.uniform(0, 400, n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = (d * r[:, None]).astype(f32)
    pos[:, 1] = np.abs(pos[:, 1]) * f32(0.25)
    return pos


def _gpu_setup(lod_over, accum=aemh.ACCUM_Q32_32):
    cfg = aemh.default_config()
    cfg.accum = accum
    eng = aemh.Engine()
    eng.set_config(cfg)
    eng.set_materials(ENGINE_MATS)
    lod = aemh.default_lod_config()
    for k, v in lod_over.items():
        setattr(lod, k, v)
    eng.tier_set_config(lod)
    return eng


def _compare_all(eng, ref, where):
    cnt, cap, act = eng.mpss_state()
    assert cnt == ref.count and np.array_equal(act, ref.active), where
    got = eng.mpss_read(np.arange(cap, dtype=np.uint32))
    for f, want in (("pos", ref.pos), ("vel", ref.vel), ("C", ref.C), ("F", ref.F), ("J", ref.J),
                    ("mass", ref.mass), ("lifetime", ref.lifetime), ("age", ref.age)):
        g = got[f].reshape(want.shape)
        assert np.array_equal(g, want), f"{where}: {f} differs in {int((g != want).sum())} values"


@pytest.mark.gpu
def test_gpu_tier_ticks_bitexact():
    lod_over = dict(max_near_particles=300, far_field_update_interval=7)
    rng = np.random.default_rng(17)
    cap = 5000
    eng = _gpu_setup(lod_over)
    eng.mpss_reserve(cap)
    ref = RefBuf(cap)
    tm = TierManager(ref, ref_mats(), lod=lod_over)
    n = 4000
    pos = _scene(rng, n)
    slots = eng.mpss_spawn(n)
    want = [ref.spawn() for _ in range(n)]
    assert slots.tolist() == want
    vals = dict(pos=pos, vel=rng.uniform(-1, 1, (n, 3)).astype(f32),
                mat=rng.integers(0, 3, n).astype(np.uint16),
                lifetime=rng.choice([f32(0.2), f32(0.5), f32(1e9)], n).astype(f32))
    eng.mpss_write(slots, **vals)
    for k, s in enumerate(slots):
        for f, v in vals.items():
            getattr(ref, f)[s] = v[k]
    dt = 1.0 / 60
    player = np.zeros(3, f32)
    for tick in range(40):
        player = (player + np.array([0.6, 0.0, 0.2], f32)).astype(f32)
        if tick == 20:  # spawns between ticks land in the next classification
            extra = _scene(rng, 200) * f32(0.1)
            s2 = eng.mpss_spawn(200)
            w2 = [ref.spawn() for _ in range(200)]
            assert s2.tolist() == w2
            eng.mpss_write(s2, pos=extra)
            for k, s in enumerate(s2):
                ref.pos[s] = extra[k]
        got = eng.tier_tick(player, dt)
        exp = tm.tick(player, dt)
        for key in ("near", "mid", "far", "reclassified", "near_ran", "mid_ran", "far_ran",
                    "expired"):
            assert got[key] == exp[key], (tick, key, got[key], exp[key])
        assert got["near_origin"] == tuple(float(v) for v in exp["near_origin"])
        near, mid, far = eng.tier_lists()
        assert near.tolist() == tm.near and mid.tolist() == tm.mid and far.tolist() == tm.far
        if tick % 5 == 4 or tick == 39:
            _compare_all(eng, ref, f"tick {tick}")
    for tier, take in ((0, 10 ** 6), (1, 75000), (2, 100)):
        p, s = eng.tier_render(tier, take)
        pw, sw = tm.render(tier, take)
        assert np.array_equal(s, sw) and np.array_equal(p, pw), tier
    eng.close()


@pytest.mark.gpu
def test_gpu_tier_empty_and_cull_edges():
    eng = _gpu_setup(dict(max_near_particles=0))
    eng.mpss_reserve(64)
    st = eng.tier_tick((0, 0, 0), 1 / 60)  # empty buffer: nothing runs
    assert st["near"] == st["mid"] == st["far"] == 0 and not st["reclassified"]
    s = eng.mpss_spawn(3)
    eng.mpss_write(s, pos=np.array([[1, 1, 1], [2, 1, 2], [300, 1, 0]], f32))
    st = eng.tier_tick((0, 0, 0), 1 / 60)
    # max_near 0: every near particle is culled to mid
    assert (st["near"], st["mid"], st["far"]) == (0, 2, 1)
    near, mid, far = eng.tier_lists()
    assert near.size == 0 and sorted(mid.tolist()) == [0, 1] and far.tolist() == [2]
    eng.close()
