"""Call sequences through the C ABI that change the state behind a batched
step: the configuration, or the particles themselves, replaced after
aemh_set_scenes.  The batched step must advance what the caller set last
(compared bit for bit with the oracle in Q32.32 mode), or refuse with a
status code when the scene table no longer fits the particles.
"""
import numpy as np
import pytest

import aemh
from oracle import refcpu as R

pytestmark = pytest.mark.gpu

SPECS = [(2500, 32, (0.0, 0.0, 0.0), 1e-4), (4000, 40, (0.1, 0.0, -0.1), 5e-5)]
MATS = [aemh.Material(1e4, 0.2, 0, 0, 10.0, 0, 0, aemh.NEO_HOOKEAN)]


def _scenes():
    out, first = [], 0
    for n, res, o, dt in SPECS:
        sc = aemh.Scene()
        sc.first, sc.count = first, n
        sc.nx = sc.ny = sc.nz = res
        sc.dx = 1.0 / res
        sc.origin[:] = o
        sc.dt = dt
        sc.velocity_only = 0
        sc.clamp_mode = aemh.CLAMP_AUTO
        out.append(sc)
        first += n
    return out


def _particles(seed):
    rng = np.random.default_rng(seed)
    pos, vel = [], []
    for k, (n, res, o, dt) in enumerate(SPECS):
        lo = np.array(o) + 0.3
        pos.append(aemh.uniform_box(seed * 10 + k, n, lo, lo + 0.3))
        vel.append(rng.normal(0, 0.2, (n, 3)).astype(np.float32))
    pos, vel = np.concatenate(pos), np.concatenate(vel)
    mass = np.full(pos.shape[0], 10.0 * 0.027 / 3000, np.float32)
    return pos, vel, mass


def _config(accum):
    cfg = aemh.default_config()
    cfg.accum = accum
    cfg.bc_margin = 3
    return cfg


def _oracle_scenes(d, steps, accum=R.ACCUM_Q32):
    """Each scene of the downloaded state `d` advanced `steps` plain steps."""
    p, m, e = aemh.material_arrays(MATS)
    M = R.Materials(p, m, e)
    out, first = [], 0
    for n, res, o, dt in SPECS:
        sl = slice(first, first + n)
        st = R.State(d["pos"][sl], vel=d["vel"][sl], C9=d["C"][sl], F9=d["F"][sl], J=d["J"][sl],
                     mass=d["mass"][sl] if "mass" in d else None, mat=None)
        oc = R.make_config(res, res, res, np.float32(1.0 / res), o, 1, (0, -9.81, 0), 0.995, 3, 0.0,
                           accum, R.CLAMP_AUTO)
        for _ in range(steps):
            R.step(oc, M, st, dt)
        out.append(st)
        first += n
    return out


def _assert_scenes(d, states):
    first = 0
    for (n, *_), st in zip(SPECS, states):
        sl = slice(first, first + n)
        for f, want in (("pos", st.pos), ("vel", st.vel), ("C", st.C), ("F", st.F), ("J", st.J)):
            got = d[f][sl].reshape(want.shape)
            assert np.array_equal(got, want), f"{f}: {int((got != want).sum())} values differ"
        first += n


def _engine(accum, pos, vel, mass):
    eng = aemh.Engine()
    eng.set_config(_config(accum))
    eng.set_grid(32, 32, 32, 1 / 32)
    eng.set_materials(MATS)
    eng.upload(pos, vel=vel, mass=mass)
    eng.set_scenes(_scenes())
    return eng


def test_batch_step_after_accumulation_mode_switch():
    """FP32 batched steps, then aemh_set_config switches to Q32.32 with the
    same scene table: the next batched steps run in Q32.32 on the current
    particles (the grid buffers of the new mode are allocated first)."""
    pos, vel, mass = _particles(1)
    eng = _engine(aemh.ACCUM_F32, pos, vel, mass)
    eng.batch_step(2)
    d1 = eng.download()
    d1["mass"] = mass
    eng.set_config(_config(aemh.ACCUM_Q32_32))
    eng.batch_step(3)
    _assert_scenes(eng.download(), _oracle_scenes(d1, 3))
    eng.close()


def test_batch_step_after_particle_upload():
    """New particles uploaded after aemh_set_scenes (same count): the batched
    step gathers them again instead of advancing the stale working copy."""
    pos, vel, mass = _particles(2)
    eng = _engine(aemh.ACCUM_Q32_32, pos, vel, mass)
    eng.batch_step(2)
    pos2, vel2, _ = _particles(3)
    eng.upload(pos2, vel=vel2, mass=mass)
    eng.batch_step(3)
    d0 = dict(pos=pos2, vel=vel2, C=np.zeros((pos2.shape[0], 9), np.float32),
              F=np.tile(np.eye(3, dtype=np.float32).reshape(9), (pos2.shape[0], 1)),
              J=np.ones(pos2.shape[0], np.float32), mass=mass)
    _assert_scenes(eng.download(), _oracle_scenes(d0, 3))
    eng.close()


def test_batch_step_refuses_table_that_no_longer_fits():
    """More particles uploaded than the scene table covers: the batched step
    returns AEMH_ERR_ARG and launches nothing."""
    pos, vel, mass = _particles(4)
    eng = _engine(aemh.ACCUM_Q32_32, pos, vel, mass)
    eng.batch_step(1)
    extra = np.concatenate([pos, pos[:10]])
    eng.upload(extra, mass=np.concatenate([mass, mass[:10]]))
    with pytest.raises(aemh.AemhError) as ei:
        eng.batch_step(1)
    assert ei.value.status == aemh.ERR_ARG
    eng.close()
