"""Resident MpssBuffer lifecycle: spawn / kill / compact / update_lifetimes.

CPU: the oracle (oracle/mpss_ref.py) against the reference's own unit tests
(ae_particle/src/mpss.rs:571-660, ae_particle/tests/integration_tests.rs:
80-135) and hand-worked cases for what those leave open: where compact()
leaves the vacated slots, update_lifetimes() ageing only slots below count,
free-list order.
GPU: aemh_mpss_* through the C ABI against the oracle on random operation
sequences, every field of every slot bit-exact, steps in the Q32.32 mode.
"""
import numpy as np
import pytest

import aemh
from oracle import refcpu as R
from oracle.mpss_ref import F32_MAX, MpssBuffer as RefBuf

EYE = np.eye(3, dtype=np.float32).reshape(9)


# ---------------------------------------------------------------------------
# Oracle pinned by the reference's unit tests
# ---------------------------------------------------------------------------
def test_ref_spawn_and_kill():  # mpss.rs:572-587
    b = RefBuf(100)
    assert b.len() == 0
    i0 = b.spawn()
    b.spawn()
    assert b.len() == 2
    b.kill(i0)
    assert b.len() == 1
    assert b.spawn() == i0
    assert b.len() == 2


def test_ref_compact():  # mpss.rs:611-622
    b = RefBuf(5)
    b.spawn()
    i1 = b.spawn()
    b.spawn()
    b.kill(i1)
    b.compact()
    assert b.len() == 2
    assert b.active[0] and b.active[1] and not b.active[2]


def test_ref_strain_identity():  # mpss.rs:625-635
    b = RefBuf(10)
    i = b.spawn()
    assert np.array_equal(b.F[i], EYE)
    # new() leaves strain zero in slots never spawned
    assert not b.F[i + 1].any()


def test_ref_swap_slots_all_fields():  # mpss.rs:637-660, on this buffer's fields
    b = RefBuf(4)
    i0, i1, i2 = b.spawn(), b.spawn(), b.spawn()
    b.C[i1] = [1.1, 2.2, 3.3, 4.4, 5.5, 6.6, 7.7, 8.8, 9.9]
    b.mat[i1] = 42
    b.lifetime[i1] = 5.0
    b.age[i1] = 0.25
    b.Jp[i1] = 0.75
    b.kill(i0)
    b.compact()
    assert b.len() == 2
    assert np.allclose(b.C[0], [1.1, 2.2, 3.3, 4.4, 5.5, 6.6, 7.7, 8.8, 9.9])
    assert b.mat[0] == 42 and b.lifetime[0] == 5.0 and b.age[0] == 0.25 and b.Jp[0] == 0.75


def test_ref_lifetime_expiry():  # integration_tests.rs:80-95
    b = RefBuf(100)
    for _ in range(50):
        b.lifetime[b.spawn()] = 1.0
    assert b.len() == 50
    assert len(b.update_lifetimes(2.0)) == 50
    assert b.len() == 0


def test_ref_compact_full():  # integration_tests.rs:113-135
    b = RefBuf(1000)
    for i in range(500):
        b.pos[b.spawn()] = [i, 0, 0]
    for i in range(0, 500, 2):
        b.kill(i)
    b.compact()
    assert b.len() == 250
    assert b.active[:250].all() and not b.active[250:].any()
    # the survivors keep their order: odd x positions ascending
    assert np.array_equal(b.pos[:250, 0], np.arange(1, 500, 2, dtype=np.float32))


# ---------------------------------------------------------------------------
# Oracle: hand-worked cases for what the reference tests leave open
# ---------------------------------------------------------------------------
def test_ref_compact_vacated_slots_follow_swap_sequence():
    # active [F,T,F,T,T] holding a..e: swaps (0,1), (1,3), (2,4) give
    # [b, d, e, a, c] -- the vacated slots hold a and c in that order.
    b = RefBuf(5)
    for _ in range(5):
        b.spawn()
    for k in range(5):
        b.mass[k] = 10.0 + k  # a..e
    b.kill(0)
    b.kill(2)
    b.compact()
    assert list(b.mass) == [11.0, 13.0, 14.0, 10.0, 12.0]
    assert list(b.active) == [True, True, True, False, False]
    assert b.free_list == [] and b.len() == 3


def test_ref_update_lifetimes_ages_only_below_count():
    b = RefBuf(4)
    for _ in range(3):
        b.spawn()
    b.kill(0)  # count 2, slot 2 active but not below count
    assert b.update_lifetimes(1.0) == []
    assert list(b.age[:3]) == [0.0, 1.0, 0.0]
    assert b.serial_indices().tolist() == [1]


def test_ref_free_list_lifo_and_expiry_order():
    b = RefBuf(6)
    for _ in range(6):
        b.spawn()
    b.lifetime[[1, 3, 4]] = 0.5
    assert b.update_lifetimes(1.0) == [1, 3, 4]
    assert b.spawn() == 4 and b.spawn() == 3 and b.spawn() == 1
    assert b.spawn() is None  # full
    # spawn resets every field but C
    b.C[2] = 7.0
    b.kill(2)
    assert b.spawn() == 2
    assert (b.C[2] == 7.0).all() and np.array_equal(b.F[2], EYE) and b.lifetime[2] == F32_MAX


def test_ref_random_sequences_keep_invariants():
    rng = np.random.default_rng(5)
    b = RefBuf(64)
    for _ in range(400):
        op = rng.integers(4)
        if op == 0:
            b.spawn()
        elif op == 1:
            b.kill(int(rng.integers(70)))
        elif op == 2 and rng.random() < 0.2:
            b.compact()
            assert b.active[:b.count].all() and not b.active[b.count:].any()
        else:
            b.update_lifetimes(0.1)
        assert b.count == int(b.active.sum())
        assert len(set(b.free_list)) == len(b.free_list)
        assert not any(b.active[i] for i in b.free_list)


# ---------------------------------------------------------------------------
# GPU: the resident buffer against the oracle
# ---------------------------------------------------------------------------
def _engine(accum, n_grid=24, dx=1.0 / 24):
    mats = [aemh.Material(young_modulus=2e4, poisson_ratio=0.3),
            aemh.Material(young_modulus=1e4, model=aemh.DRUCKER_PRAGER),
            aemh.Material(young_modulus=3e4, model=aemh.SNOW)]
    cfg = aemh.default_config()
    cfg.accum = accum
    eng = aemh.Engine()
    eng.set_config(cfg)
    eng.set_grid(n_grid, n_grid, n_grid, dx, (0.0, 0.0, 0.0))
    eng.set_materials(mats)
    p, m, e = aemh.material_arrays(mats)
    rcfg = R.make_config(n_grid, n_grid, n_grid, dx, (0, 0, 0), 1, (0, -9.81, 0), 0.995, 0, 0.0,
                         R.ACCUM_Q32 if accum == aemh.ACCUM_Q32_32 else R.ACCUM_F32,
                         R.CLAMP_AUTO)
    return eng, rcfg, R.Materials(p, m, e)


def _compare(eng, ref, where):
    cnt, cap, act = eng.mpss_state()
    assert cnt == ref.count and cap == ref.capacity, where
    assert np.array_equal(act, ref.active), where
    got = eng.mpss_read(np.arange(cap, dtype=np.uint32))
    for f, want in (("pos", ref.pos), ("vel", ref.vel), ("C", ref.C), ("F", ref.F), ("J", ref.J),
                    ("Jp", ref.Jp), ("mass", ref.mass), ("mat", ref.mat),
                    ("lifetime", ref.lifetime), ("age", ref.age)):
        g = got[f].reshape(want.shape)
        assert np.array_equal(g, want), f"{where}: {f} differs in {int((g != want).sum())} values"


def _fill(rng, slots):
    n = len(slots)
    return dict(pos=rng.uniform(0.3, 0.7, (n, 3)).astype(np.float32),
                vel=rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32),
                mass=rng.uniform(0.5, 2.0, n).astype(np.float32),
                mat=rng.integers(0, 3, n).astype(np.uint16),
                lifetime=rng.choice([0.004, 0.012, 1e9], n).astype(np.float32))


def _apply_fill(ref, slots, vals):
    for k, s in enumerate(slots):
        for f, v in vals.items():
            getattr(ref, f)[s] = v[k]


@pytest.mark.gpu
def test_gpu_lifecycle_ops_match_oracle():
    eng, _, _ = _engine(aemh.ACCUM_Q32_32)
    rng = np.random.default_rng(21)
    cap = 300
    ref = RefBuf(cap)
    eng.mpss_reserve(cap)
    _compare(eng, ref, "new")
    for it in range(60):
        op = rng.integers(5)
        if op == 0:
            k = int(rng.integers(1, 40))
            got = eng.mpss_spawn(k)
            want = [ref.spawn() for _ in range(k)]
            want = np.array([aemh.NO_SLOT if w is None else w for w in want], np.uint32)
            assert np.array_equal(got, want), f"spawn {it}"
            live = got[got != aemh.NO_SLOT]
            vals = _fill(rng, live)
            eng.mpss_write(live, **vals)
            _apply_fill(ref, live, vals)
        elif op == 1:
            ks = rng.integers(0, cap + 10, int(rng.integers(1, 30))).astype(np.uint32)
            eng.mpss_kill(ks)
            for s in ks:
                ref.kill(int(s))
        elif op == 2:
            eng.mpss_compact()
            ref.compact()
        else:
            dt = float(rng.choice([0.004, 0.01]))
            got = eng.mpss_update_lifetimes(dt)
            assert got.tolist() == ref.update_lifetimes(dt), f"update_lifetimes {it}"
        _compare(eng, ref, f"op {it} ({op})")
    eng.close()


@pytest.mark.gpu
def test_gpu_steps_with_lifecycle_bitexact():
    """Steps over the resident slots (serial set and explicit lists) with
    spawns, kills, compaction and expiry between them, Q32.32 mode."""
    eng, rcfg, M = _engine(aemh.ACCUM_Q32_32)
    rng = np.random.default_rng(8)
    cap = 2000
    ref = RefBuf(cap)
    eng.mpss_reserve(cap)
    dt = 0.004
    for it in range(24):
        if it % 3 == 0:
            got = eng.mpss_spawn(int(rng.integers(50, 400)))
            want = []
            for _ in range(got.size):
                w = ref.spawn()
                want.append(aemh.NO_SLOT if w is None else w)
            assert np.array_equal(got, np.array(want, np.uint32))
            live = got[got != aemh.NO_SLOT]
            vals = _fill(rng, live)
            eng.mpss_write(live, **vals)
            _apply_fill(ref, live, vals)
        if it % 4 == 1:
            ks = rng.choice(cap, 120, replace=False).astype(np.uint32)
            eng.mpss_kill(ks)
            for s in ks:
                ref.kill(int(s))
        if it % 7 == 6:
            eng.mpss_compact()
            ref.compact()
        if it % 2 == 0:
            eng.step(dt)
            ref.step(rcfg, M, dt)
        else:
            idx = ref.active_indices()
            idx = idx[rng.random(idx.size) < 0.7]
            eng.mpm_step_parallel_with_grid(idx, dt)
            ref.step(rcfg, M, dt, idx=idx)
        _compare(eng, ref, f"step {it}")
    eng.close()


@pytest.mark.gpu
def test_gpu_mpss_errors_and_reserve_resets():
    eng, _, _ = _engine(aemh.ACCUM_F32)
    with pytest.raises(aemh.AemhError):
        eng.mpss_compact()  # no resident buffer
    eng.mpss_reserve(8)
    with pytest.raises(aemh.AemhError):
        eng.mpss_write(np.array([8], np.uint32), mass=np.ones(1, np.float32))
    s = eng.mpss_spawn(10)
    assert s[:8].tolist() == list(range(8)) and (s[8:] == aemh.NO_SLOT).all()
    eng.mpss_reserve(8)
    cnt, cap, act = eng.mpss_state()
    assert cnt == 0 and cap == 8 and not act.any()
    eng.close()
