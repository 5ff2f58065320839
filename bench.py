#!/usr/bin/env python3
"""bench.py -- MI355X MLS-MPM step throughput (particle-updates/s).

Workload of `value`: 1,048,576 snow particles (dam-break block
[0,0.4]x[0,0.6]x[0,1]) on a 128^3 grid, dx 1/128, dt 5e-5, FP32 mode, one
substep per step, inputs resident in HBM.  One step = sort by block key,
P2G, grid update, G2P.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]

Prints one JSON line.  By default W = 5 and K = 195: the timed window is
steps 6-200 of the configuration's 200-step dam break.  `value` times K
steps enqueued back to back through aemh_step_async (each replays the
captured substep hipGraph) and one aemh_sync at the end, which brackets the
region with the stream synchronized on both sides.  The particles are then
put back to their state at the start of that window twice: the same K steps
run once through aemh_step, which returns after a stream synchronize per step
(`extra.synchronized_steps_particle_updates_per_s`), and once with profiling
on -- the same kernels launched eagerly on the context stream, each stage's
HIP event pair carried by its kernels' dispatches -- to give the per-kernel
durations behind `roofline` (per-kernel events slow the step itself, so they
are kept out of the window that gives `value`; replaying the same steps
matters because P2G's cost grows as the scene evolves).  The launch shape is
autotuned once before the warm-up.  `extra` adds the profiled window's own
throughput, the 16M-particle sparse-grid workload, the deterministic Q32.32
mode, the batched-scenes speedup, the mixed-precision workload, the
engine's LOD tier tick on a 3M-capacity resident buffer and the all-integer
FixedPoint MpmSimulation on 1M particles.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ae-engine_amd"))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/s at 1M & 16M particles; %HBM peak; batched vs plain speedup"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes (counted once per step, FP32 storage): per particle
#   P2G reads x 12, v 12, C 36, F 36, J 4, J_p 4, mass 4, material 2 = 110 B
#   G2P reads x 12, F 36, J_p 4 and writes x, v, C, F, J, J_p = 104 B -> 156 B
#   whole step 206 B + 8 B (J_p read+write)                         = 214 B
# per active grid node: P2G writes 28 B (m, p, f); update reads 28, writes
#   16 (v, m); G2P reads 16 -> step 88 B; P2G 28 B; G2P 16 B.
# A G2P that also counts the next substep's sort (cache-resident working
# sets; the count kernel then does not run) writes each particle's block key
# and rank, 8 B more.
BYTES = {"p2g": (110, 28), "g2p": (156, 16), "g2p_count": (164, 16), "grid_update": (0, 44),
         "step": (214, 88)}
KERNELS = ["count", "scan", "scatter", "p2g", "grid_update", "g2p"]
# HBM bytes per dispatch from rocprofv3 FETCH_SIZE (x 2.000) and WRITE_SIZE
# (x 1.000), calibrated against a 256 MiB copy, as recorded in
# profiles/p2g_g2p_sq_counters.md (1M-particle snow workload, launch shape
# 5,1,256,8192; 16M-particle workload, early steps, 5,0,256,8192).
TRAFFIC_C2 = {"p2g": 184.2e6, "g2p": 220.5e6, "grid_update": 90.9e6}
TRAFFIC_C3 = {"p2g": 2500e6, "g2p": 3068e6}
# VALU instructions per wave (one wave = 64 particles) from rocprofv3
# SQ_INSTS_VALU / SQ_WAVES, recorded in profiles/p2g_g2p_sq_counters.md; P2G
# by its lane assignment (lanes keep their particles, layers, pairs).  The
# issue ceiling without packed FP32: one wave64 VALU instruction per 4 cycles
# per SIMD, 4 SIMDs x 256 CUs at 2.4 GHz = 614.4e9 wave-instructions/s.
VALU_PEAK_WINST = 256 * 4 * 2.4e9 / 4
VALU_PER_WAVE_C2 = {"p2g": {"own": 1544, "layers": 1588, "pairs": 1743}, "g2p": 1521}
VALU_PER_WAVE_C3 = {"p2g": {"own": 1546, "layers": 1601, "pairs": 1797}, "g2p": 1295}


def lanes_of(shape):
    """The P2G lane assignment of a launch shape."""
    return "pairs" if shape["p2g_pairs"] else ("layers" if shape["p2g_layers"] else "own")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_steps(eng, dt, k):
    """K steps, each returning after the GPU finished it (aemh_step)."""
    t0 = time.perf_counter()
    for _ in range(k):
        eng.step(dt)
    return time.perf_counter() - t0


def timed_steps_async(eng, dt, k):
    """K steps enqueued back to back (aemh_step_async), one synchronize at
    the end: the GPU does not idle between steps for the host's round trip."""
    t0 = time.perf_counter()
    for _ in range(k):
        eng.step_async(dt)
    eng.sync()
    return time.perf_counter() - t0


def snapshot(eng):
    return eng.download()


def restore(eng, wl, snap, shape):
    """Puts the particles back to a snapshot, so the profiled window replays
    the timed window's own steps (P2G's cost grows as the scene evolves).
    The upload leaves them in canonical order: the first step after it
    re-sorts, so the caller runs that step untimed on both sides."""
    eng.upload(snap["pos"], vel=snap["vel"], C9=snap["C"], F9=snap["F"], J=snap["J"],
               mass=wl.mass, mat=wl.mat, Jp=snap["Jp"])
    eng.set_launch_shape(shape["p2g_tiles"], shape["p2g_partials"], shape["g2p_wg"],
                         shape["upd_wgs"], shape["p2g_pairs"], shape["p2g_layers"])


def run_main(W, aemh, steps, warmup, accum=None):
    wl = W.c2_snow_dam()
    if accum is not None:
        wl.accum = accum
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    eng.step(wl.dt)
    shape = eng.autotune()  # one-time launch-shape choice, before the timed region
    # The last warm-up step follows the snapshot: it is repeated after the
    # restore, where it re-sorts the uploaded particles.
    pre = 1 if warmup > 0 else 0
    for _ in range(warmup - pre):
        eng.step(wl.dt)
    snap = snapshot(eng)
    for _ in range(pre):
        eng.step(wl.dt)
    el = timed_steps_async(eng, wl.dt, steps)
    # the same steps again, each synchronized, then profiled
    restore(eng, wl, snap, shape)
    for _ in range(pre):
        eng.step(wl.dt)
    el_sync = timed_steps(eng, wl.dt, steps)
    restore(eng, wl, snap, shape)
    for _ in range(pre):
        eng.step(wl.dt)
    eng.profile(True)
    eng.profile_reset()
    el_prof = timed_steps(eng, wl.dt, steps)
    kern = {k: eng.profile_get(k)[0] for k in KERNELS}
    eng.profile(False)
    blocks, npart = eng.last_substep_info()
    eng.close()
    return wl, el, el_sync, el_prof, kern, blocks, npart, shape


def roofline_of(kern, name, npart, blocks, traffic=None, valu=None, lanes="own"):
    """HBM roofline of one kernel: algorithmic bytes per launch over its average
    HIP-event duration.  With `valu` (VALU instructions per wave from the
    counter profile; P2G's by its lane assignment), also the VALU issue rate
    against its ceiling: P2G is limited by VALU issue and LDS atomics long
    before HBM.  A G2P that counted the next substep's sort (no count
    kernel in the window) is charged its key and rank writes."""
    counted = name == "g2p" and kern.get("count", 0.0) == 0.0
    per_p, per_n = BYTES["g2p_count" if counted else name]
    nbytes = npart * per_p + blocks * 64 * per_n
    ms = kern[name]
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4),
           "traffic": None if traffic is None else int(traffic.get(name, 0)) or None,
           "kernel": name, "algorithmic_bytes": int(nbytes), "avg_ms": round(ms, 4)}
    per_wave = None
    if valu and name in valu:
        per_wave = valu[name].get(lanes) if isinstance(valu[name], dict) else valu[name]
    if per_wave and ms > 0:
        rate = (npart + 63) // 64 * per_wave / (ms * 1e-3)
        out["valu_issue"] = {"insts_per_wave": per_wave, "achieved": round(rate / 1e9, 1),
                             "peak": VALU_PEAK_WINST / 1e9, "unit": "G wave-instructions/s",
                             "frac": round(rate / VALU_PEAK_WINST, 4)}
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def native_oracle():
    """The oracle built for this host the way the reference builds for its own
    (Cargo.toml:116-120: opt-level 3, fat LTO; .cargo/config.toml:
    target-cpu=native): g++ -O3 -march=native -flto, still without FMA
    contraction (rustc never fuses).  Built here, on the machine that runs
    it; returns the library path, or None if the build fails."""
    import subprocess
    import tempfile
    out = os.path.join(tempfile.mkdtemp(prefix="aemh_native_"), "librefcpu_native.so")
    cmd = ["g++", "-O3", "-march=native", "-flto", "-std=c++17", "-fPIC", "-shared", "-pthread",
           "-ffp-contract=off", "-fno-math-errno", os.path.join(ROOT, "oracle", "refcpu.cpp"),
           "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return out if r.returncode == 0 else None


def cpu_baseline(W, aemh, target_s=12.0):
    """The oracle on the same C2 workload on this host: the reference's
    >= 50,000-index structure (per-thread grids and a merge, parallel G2P),
    built -O3 -march=native -flto here, on the cores this process may use
    (at most the box's CPU share of 16), over ~target_s seconds of steps
    after one warm-up step."""
    from oracle import refcpu as R
    path = native_oracle()
    if path:
        R.use_library(path)
    wl = W.c2_snow_dam()
    p, m, e = aemh.material_arrays(wl.materials)
    M = R.Materials(p, m, e)
    nx, ny, nz, dx, o = wl.grid
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    cores = max(1, min(affinity, share))
    cfg = R.make_config(nx, ny, nz, dx, o, 1, wl.gravity, 0.995, wl.bc_margin, 0.0, R.ACCUM_F32,
                        R.CLAMP_AUTO, threads=cores)
    st = R.State(wl.pos, mass=wl.mass, mat=wl.mat)
    R.step(cfg, M, st, wl.dt)  # warm-up
    n_steps = 0
    t0 = time.perf_counter()
    while n_steps < 5 or (time.perf_counter() - t0 < target_s and n_steps < 200):
        R.step(cfg, M, st, wl.dt)
        n_steps += 1
    el = time.perf_counter() - t0
    return {"value": wl.n * n_steps / el, "unit": "particle-updates/s", "cores": cores,
            "kind": "port",
            "host": {"cpu": cpu_model(), "nproc": os.cpu_count(), "affinity_cores": affinity,
                     "cpu_share": share,
                     "build": "g++ -O3 -march=native -flto -ffp-contract=off" if path
                              else "prebuilt -O3 -march=x86-64-v3 (native build failed)"},
            "sample": f"oracle (C++ restatement, per-thread grid fold + merge like the "
                      f"reference's rayon path) on the same 1,048,576-particle 128^3 snow workload, "
                      f"steps 2-{n_steps + 1} ({el:.1f} s) after 1 warm-up, {cores} threads"}


def extra_c3(W, aemh, steps=96):
    """The configuration's 100 steps: 4 untimed (the first, the autotuner's
    state, two more), then steps 5-100 timed; the profiled replay repeats
    them from a snapshot."""
    wl = W.c3_three_boxes()
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    eng.step(wl.dt)
    shape = eng.autotune()
    for _ in range(2):
        eng.step(wl.dt)
    snap = snapshot(eng)
    eng.step(wl.dt)  # repeated after the restore, where it re-sorts the uploaded particles
    el = timed_steps(eng, wl.dt, steps)
    restore(eng, wl, snap, shape)
    eng.step(wl.dt)
    eng.profile(True)
    # per-step kernel times of the replayed window, averaged over the whole
    # window and over windows of 10 steps (P2G and G2P grow as the boxes land:
    # the snow starts to yield and fast fluid leaves its sorted order)
    per = {k: [] for k in KERNELS}
    for _ in range(steps):
        eng.profile_reset()
        eng.step(wl.dt)
        for k in KERNELS:
            per[k].append(eng.profile_get(k)[0])
    kern = {k: sum(v) / len(v) for k, v in per.items()}
    first = 101 - steps
    windows = {}
    for a in range(0, steps, 10):
        b = min(a + 10, steps)
        windows[f"steps {first + a}-{first + b - 1}"] = {
            k: round(sum(per[k][a:b]) / (b - a), 4) for k in ("p2g", "g2p", "count")}
    blocks, npart = eng.last_substep_info()
    eng.close()
    total_blocks = 64 ** 3
    step_bytes = npart * BYTES["step"][0] + blocks * 64 * BYTES["step"][1]
    return {
        "workload": "c3_three_boxes_16m_256: 16,777,216 particles (fluid / jelly / snow boxes), "
                    "256^3 grid of 4^3 blocks, dt 5e-5, FP32",
        "particle_updates_per_s": wl.n * steps / el,
        "ms_per_step": el / steps * 1e3,
        "active_block_fraction": blocks / total_blocks,
        "step_hbm_frac": round(step_bytes / (el / steps) / 1e9 / HBM_PEAK_GBS, 4),
        "g2p": roofline_of(kern, "g2p", npart, blocks, TRAFFIC_C3, VALU_PER_WAVE_C3),
        "p2g": roofline_of(kern, "p2g", npart, blocks, TRAFFIC_C3, VALU_PER_WAVE_C3,
                           lanes_of(shape)),
        "ms_by_window": windows,
        "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
        "launch_shape": shape,
    }


def extra_c5(W, aemh, steps=498):
    """4,194,304 particles on 192^3, C and F - I stored in BF16 (FP32 compute),
    FP64 energy/momentum reductions every 10 steps inside the timed loop; the
    same run with FP32 storage beside it.  The same pair on the all-Neo-Hookean
    scene (`_nh`), whose G2P is bound by HBM traffic rather than by the snow
    SVD, shows what the smaller storage buys there."""
    out = {"workload": "c5_mixed_4m_192: 4,194,304 particles (snow + Neo-Hookean boxes), 192^3 "
                       "grid, dt 5e-5, C and F-I in BF16 / FP32 storage, FP32 compute, FP64 "
                       "conservation sums every 10 steps; _nh: both boxes Neo-Hookean"}
    for name, storage, nh in (("bf16", aemh.STORE_BF16, False), ("f32", aemh.STORE_F32, False),
                              ("bf16_nh", aemh.STORE_BF16, True), ("f32_nh", aemh.STORE_F32, True)):
        wl = W.c5_mixed_precision(storage=storage, nh_only=nh)
        eng = aemh.Engine()
        W.load_engine(eng, wl)
        eng.step(wl.dt)
        eng.autotune()
        eng.step(wl.dt)
        sums = []
        t0 = time.perf_counter()
        for k in range(steps):
            eng.step(wl.dt)
            if k % 10 == 9:
                sums.append(eng.conservation_f64())
        el = time.perf_counter() - t0
        eng.profile(True)
        eng.profile_reset()
        timed_steps(eng, wl.dt, 3)
        kern = {k: eng.profile_get(k)[0] for k in KERNELS}
        eng.close()
        out[name] = {"particle_updates_per_s": wl.n * steps / el, "ms_per_step": el / steps * 1e3,
                     "steps": f"3-{steps + 2} timed (2 untimed), FP64 sums every 10 inside",
                     "p2g_ms": round(kern["p2g"], 4), "g2p_ms": round(kern["g2p"], 4),
                     "fp64_sums_last": [float(x) for x in sums[-1]]}
    out["bf16_vs_f32_speedup"] = out["bf16"]["particle_updates_per_s"] / out["f32"]["particle_updates_per_s"]
    out["bf16_vs_f32_speedup_nh"] = (out["bf16_nh"]["particle_updates_per_s"] /
                                     out["f32_nh"]["particle_updates_per_s"])
    out["g2p_bf16_vs_f32_nh"] = out["f32_nh"]["g2p_ms"] / out["bf16_nh"]["g2p_ms"]
    return out


def c5_error(W, aemh, n=1 << 16, res=48, steps=500):
    """The C5 error report at its configured horizon: the mixed-precision
    workload (snow box + Neo-Hookean box) at 65,536 particles on 48^3 (its
    particle density), 500 steps on the GPU with C and F - I stored in BF16,
    FP16 or FP32 (FP32 compute), against an all-FP64 run of the oracle on the
    same seed: rel-L2 of x, v, x - x0 and F."""
    import numpy as np
    from oracle import refcpu as R
    wl = W.c5_mixed_precision(n=n, res=res)
    nx, ny, nz, dx, o = wl.grid
    share = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    cfg = R.make_config(nx, ny, nz, dx, o, 1, wl.gravity, 0.995, wl.bc_margin, 0.0, R.ACCUM_F32,
                        R.CLAMP_AUTO, threads=max(1, min(share, os.cpu_count() or 1)))
    p, m, e = aemh.material_arrays(wl.materials)
    M = R.Materials(p, m, e)
    ref = R.State(wl.pos, mass=wl.mass, mat=wl.mat, dtype=np.float64)
    t0 = time.perf_counter()
    for _ in range(steps):
        R.step(cfg, M, ref, wl.dt)
    t_ref = time.perf_counter() - t0
    x0 = wl.pos.astype(np.float64)

    def rel(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    out = {"workload": f"c5_mixed scaled: {n} particles, {res}^3, snow + Neo-Hookean, {steps} steps, "
                       f"against the FP64 oracle ({t_ref:.1f} s on the host)"}
    for name, storage in (("bf16", aemh.STORE_BF16), ("f16", aemh.STORE_F16), ("f32", aemh.STORE_F32)):
        wl.storage = storage
        eng = aemh.Engine()
        W.load_engine(eng, wl)
        for _ in range(steps):
            eng.step(wl.dt)
        d = eng.download()
        eng.close()
        out[name] = {"x": rel(d["pos"], ref.pos), "v": rel(d["vel"], ref.vel),
                     "x_minus_x0": rel(d["pos"] - x0, ref.pos - x0), "F": rel(d["F"], ref.F)}
    return out


def extra_batched(W, aemh, steps=96):
    """The configuration's 100 steps: 4 untimed batched substeps (the first
    runs the autotuner and builds the graphs), then 96 timed; the plain form
    runs the same 4 + 96 per scene."""
    wl = W.c4_batched_scenes()
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    eng.batch_step(4)
    t0 = time.perf_counter()
    eng.batch_step(steps)
    el_b = time.perf_counter() - t0
    eng.close()
    # plain: one context per scene, one aemh_step call per scene and step
    engines = []
    for sc in wl.scenes:
        sl = slice(sc.first, sc.first + sc.count)
        e = aemh.Engine()
        cfg = wl.config()
        e.set_config(cfg)
        e.set_grid(sc.nx, sc.ny, sc.nz, sc.dx, tuple(sc.origin))
        e.set_materials(wl.materials)
        e.upload(wl.pos[sl], mass=wl.mass[sl], mat=wl.mat[sl])
        for _ in range(4):
            e.step(sc.dt)
        engines.append((e, sc.dt))
    t0 = time.perf_counter()
    for _ in range(steps):
        for e, dt in engines:
            e.step(dt)
    el_p = time.perf_counter() - t0
    for e, _ in engines:
        e.close()
    return {"workload": f"c4_batched_256_scenes: {len(wl.scenes)} scenes, {wl.n} particles, "
                        f"grids 32^3..128^3, NH + snow",
            "batched_particle_updates_per_s": wl.n * steps / el_b,
            "plain_particle_updates_per_s": wl.n * steps / el_p,
            "batched_vs_plain_speedup": el_p / el_b}


TIER_KERNELS = ("lod_classify", "lod_select", "near_cull", "count", "scan", "scatter",
                "p2g", "grid_update", "g2p", "far_update", "lifetimes")


def extra_tiers(W, aemh, ticks=60, with_cpu=True):
    """The engine's LOD tier tick (simulation.rs:474-697) on the resident
    buffer at engine scale; per-tick wall time by kind of tick, the
    classification pass against HBM, and the oracle's cost of the same tick
    on the host."""
    import numpy as np
    tw = W.engine_tiers()
    eng = aemh.Engine()
    eng.set_config(aemh.default_config())  # MpmConfig::default: FP32, 1 substep
    eng.set_materials(tw.materials)
    lod = aemh.default_lod_config()
    eng.tier_set_config(lod)
    eng.mpss_reserve(tw.capacity)
    slots = eng.mpss_spawn(tw.n)
    eng.mpss_write(slots, pos=tw.pos, mat=tw.mat)
    dt = 1.0 / 60
    player = np.zeros(3, np.float32)
    walk = np.array([1.4 * dt, 0.0, 0.0], np.float32)  # walking pace
    for _ in range(12):  # allocations, first classification, first mid tick
        player += walk
        eng.tier_tick(player, dt)
    kinds = {}
    t_all = 0.0
    for _ in range(ticks):
        player += walk
        t0 = time.perf_counter()
        st = eng.tier_tick(player, dt)
        el = time.perf_counter() - t0
        t_all += el
        kind = "mid" if st["mid_ran"] else ("reclassify" if st["reclassified"] else "near")
        if st["far_ran"]:
            kind += "+far"
        kinds.setdefault(kind, []).append(el)
    eng.profile(True)
    eng.profile_reset()
    for _ in range(24):
        player += walk
        eng.tier_tick(player, dt)
    prof = {k: eng.profile_get(k) for k in TIER_KERNELS}
    eng.profile(False)
    count = eng.mpss_state()[0]
    near, mid, far = eng.tier_lists()
    eng.close()
    cls_ms = prof["lod_classify"][0]
    cls_bytes = count * (12 + 1 + 1 + 4)  # pos read, active read, class + distance written
    out = {"workload": f"{tw.name}: capacity {tw.capacity}, {tw.n} live, tiers near/mid/far "
                       f"{near.size}/{mid.size}/{far.size} (max_near 10000), player walking 1.4 m/s, "
                       f"dt 1/60, FP32",
           "ticks": ticks,
           "ms_per_tick": round(t_all / ticks * 1e3, 3),
           "ms_by_tick_kind": {k: {"n": len(v), "avg_ms": round(sum(v) / len(v) * 1e3, 3)}
                               for k, v in sorted(kinds.items())},
           "kernel_avg_ms": {k: round(v[0], 4) for k, v in prof.items() if v[1]},
           "classify_roofline": {"bound": "hbm", "achieved": round(cls_bytes / (cls_ms * 1e-3) / 1e9, 1)
                                 if cls_ms > 0 else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "algorithmic_bytes": cls_bytes, "avg_ms": round(cls_ms, 4)}}
    if cls_ms > 0:
        out["classify_roofline"]["frac"] = round(cls_bytes / (cls_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    if with_cpu:
        out["cpu_port"] = tier_cpu_port(tw, aemh, lod, near, mid, far, dt)
    return out


def extra_fixed(W, aemh, n=1 << 20, res=64, steps=8, warmup=2, with_cpu=True):
    """The all-integer FixedPoint MpmSimulation (ae_physics/src/mpm.rs) on a
    fresh cloud: `warmup` steps, then `steps` in one aemh_fixed_step call
    (one substep each), bracketed by the call's own synchronize; beside it
    the oracle (serial C++ restatement of the reference's loops, one
    thread) on the same scene for 2 steps."""
    import numpy as np
    sc = W.fixed_cloud(n, res)
    zero = np.zeros_like(sc.pos)
    sim = aemh.FixedSim(sc.config)
    sim.add_particles(sc.pos, zero, sc.mass)
    sim.step(warmup)
    t0 = time.perf_counter()
    sim.step(steps)
    el = time.perf_counter() - t0
    st = sim.state()
    sim.close()
    out = {"workload": f"{sc.name}: {sc.n} particles uniform in the middle half of a {res}^3 grid, "
                       f"cell 0.5, E 50, density 1, dt 1/60, elastoplastic, 1 substep, Q32.32 "
                       f"integer arithmetic, bit-exact with the reference's serial loops",
           "steps": steps, "warmup": warmup,
           "particle_updates_per_s": sc.n * steps / el,
           "ms_per_step": round(el / steps * 1e3, 3),
           "scatter_bound_frac": round(st["max_scatter_bound"] / float(1 << 39), 8),
           "checked_substeps": st["checked_substeps"], "serial_substeps": st["serial_substeps"]}
    if with_cpu:
        from oracle import fixedsim as X
        o = X.Sim(sc.config)
        o.add_particles(sc.pos, zero, sc.mass)
        t0 = time.perf_counter()
        o.step(2)
        elc = time.perf_counter() - t0
        o.close()
        out["cpu_port"] = {"value": sc.n * 2 / elc, "unit": "particle-updates/s", "cores": 1,
                           "kind": "port",
                           "sample": "oracle/fixedsim.cpp (serial, the reference's own loop "
                                     "structure) on the same scene, its first 2 steps"}
    return out


def tier_cpu_port(tw, aemh, lod, near, mid, far, dt):
    """The oracle's cost of one average tick of the same world on the host:
    classification (vectorized f32, stable sort for the cull) amortized over
    the 8-tick reclassification, the near step (serial, the reference's path
    below 50,000 indices) every tick, the velocity-only mid step (serial P2G)
    every 12 ticks, two lifetime passes per tick.  Timed once each."""
    import numpy as np
    from oracle import refcpu as R
    f32 = np.float32
    p, m, e = aemh.material_arrays(tw.materials)
    M = R.Materials(p, m, e)
    st = R.State(tw.pos, mat=tw.mat)
    player = np.zeros(3, f32)
    t0 = time.perf_counter()
    d = (st.pos - player).astype(f32)
    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    nr = np.flatnonzero(d2 < f32(lod.near_field_distance) ** 2)
    md = np.flatnonzero((d2 >= f32(lod.near_field_distance) ** 2)
                        & (d2 < f32(lod.mid_field_distance) ** 2))
    _fr = np.flatnonzero(d2 >= f32(lod.mid_field_distance) ** 2)
    order = np.argsort(d2[nr], kind="stable")
    _near_c = nr[order[:lod.max_near_particles]]
    _mid_c = np.concatenate([md, nr[order[lod.max_near_particles:]]])
    t_cls = time.perf_counter() - t0
    half = f32(f32(lod.near_n) * f32(lod.near_dx)) * f32(0.5)
    cfg_n = R.make_config(lod.near_n, lod.near_n, lod.near_n, float(lod.near_dx),
                          tuple(float(-half) for _ in range(3)), 1, (0, -9.81, 0), 0.995, 0, 0.0,
                          R.ACCUM_F32, R.CLAMP_AUTO, threads=1)
    t0 = time.perf_counter()
    R.step(cfg_n, M, st, dt, idx=near)
    t_near = time.perf_counter() - t0
    mh = f32(f32(lod.mid_n) * f32(lod.mid_dx)) * f32(0.5)
    cfg_m = R.make_config(lod.mid_n, lod.mid_n, lod.mid_n, float(lod.mid_dx),
                          tuple(float(-mh) for _ in range(3)), 1, (0, -9.81, 0), 0.995, 0, 0.0,
                          R.ACCUM_F32, R.CLAMP_AUTO, threads=1)
    t0 = time.perf_counter()
    R.step_velocity_only(cfg_m, st, dt * lod.mid_interval, idx=mid)
    t_mid = time.perf_counter() - t0
    age = np.zeros(tw.n, f32)
    life = np.full(tw.n, np.finfo(f32).max, f32)
    t0 = time.perf_counter()
    age += f32(dt)
    _exp = np.flatnonzero(age >= life)
    t_life = time.perf_counter() - t0
    per_tick = t_near + t_cls / 8 + t_mid / lod.mid_interval + 2 * t_life
    return {"ms_per_tick": round(per_tick * 1e3, 3), "cores": 1, "kind": "port",
            "parts_ms": {"classify": round(t_cls * 1e3, 3), "near_step": round(t_near * 1e3, 3),
                         "mid_velocity_only_step": round(t_mid * 1e3, 3),
                         "lifetimes": round(t_life * 1e3, 3)},
            "sample": "oracle components timed once each on the same world, composed with "
                      "the tick cadences (classify 1/8, near 1/1, mid 1/12, lifetimes 2/1)"}


def probe(W, aemh, which="c2", shape=None):
    """Counter-collection run.  The copy moves exactly 256 MiB in and 256 MiB
    out, so FETCH_SIZE / WRITE_SIZE of its dispatch give the KiB each counter
    reports per real KiB; the dispatches of the 5 steps after the autotuner
    (the launch shape the benchmark uses) follow, on the 1M-particle snow
    workload (c2) or the 16M-particle three-box workload (c3); c3late runs
    steps 2-96 of the 16M workload instead, so its last dispatches are late
    steps (yielding snow, splashed fluid); or 10 steps of the 1M-particle
    FixedPoint cloud (fixed)."""
    import torch
    x = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    if which == "fixed":
        import numpy as np
        sc = W.fixed_cloud()
        sim = aemh.FixedSim(sc.config)
        sim.add_particles(sc.pos, np.zeros_like(sc.pos), sc.mass)
        sim.step(10)
        sim.close()
        print("probe done: fixed cloud, 10 steps", flush=True)
        return
    wl = W.c3_three_boxes() if which in ("c3", "c3late") else W.c2_snow_dam()
    eng = aemh.Engine()
    W.load_engine(eng, wl)
    eng.step(wl.dt)
    if shape:
        eng.set_launch_shape(*shape)
        shape = eng.launch_shape()
    else:
        shape = eng.autotune()
    for s in range(95 if which == "c3late" else 5):
        eng.step(wl.dt)
        if s % 10 == 9:
            print("probe step", s + 2, flush=True)
    eng.close()
    print("probe done, launch shape", shape, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=195)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--extras", default="q32,c3,c4,c5,tiers,fixed",
                    help="comma list of the extra workloads to run (q32, c3, c4, c5, tiers, "
                         "fixed)")
    ap.add_argument("--probe", nargs="?", const="c2", choices=("c2", "c3", "c3late", "fixed"),
                    help="short run for hardware-counter collection: a 256 MiB device copy "
                         "(calibrates FETCH_SIZE / WRITE_SIZE) then 5 steps of the main workload "
                         "(c2, the default) or of the 16M-particle workload (c3), steps 2-96 of "
                         "the 16M-particle workload (c3late: the last 5 are its late steps), or "
                         "10 steps of the FixedPoint cloud (fixed)")
    ap.add_argument("--shape", default=None,
                    help="with --probe: a fixed launch shape 'p2g_tiles,p2g_partials,g2p_wg,"
                         "upd_wgs[,p2g_pairs[,p2g_layers]]' instead of the autotuner's pick (the autotuner's "
                         "timings move under a profiler)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: the MPM step runs on a single GPU and does not "
              "shard across GPUs; run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    import aemh
    from aemh import workloads as W

    if args.probe:
        probe(W, aemh, args.probe,
              tuple(int(x) for x in args.shape.split(",")) if args.shape else None)
        return

    log("[bench] main workload: C2 1M snow particles, 128^3")
    wl, el, el_sync, el_prof, kern, blocks, npart, shape = run_main(W, aemh, args.steps, args.warmup)
    value = wl.n * args.steps / el
    dominant = max(("p2g", "g2p", "grid_update"), key=lambda k: kern[k])
    step_bytes = npart * BYTES["step"][0] + blocks * 64 * BYTES["step"][1]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded xoshiro256** particles (seed 42+index), v=0, F=I",
        "config": {"workload": "c2_snow_1m_128: 1,048,576 snow particles in a dam-break block, "
                               "128^3 grid, dx 1/128, dt 5e-5, FP32 P2G",
                   "particles": wl.n, "grid": [128, 128, 128], "accum": "f32", "substeps": 1,
                   "launch_shape": {k: shape[k] for k in ("p2g_tiles", "p2g_partials", "p2g_pairs",
                                                            "p2g_layers",
                                                            "g2p_wg", "upd_wgs")}},
        "roofline": roofline_of(kern, dominant, npart, blocks, TRAFFIC_C2, VALU_PER_WAVE_C2,
                                lanes_of(shape)),
        "step_hbm_frac": round(step_bytes / (el / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
        "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
    }
    extra = {"synchronized_steps_particle_updates_per_s": wl.n * args.steps / el_sync,
             "profiled_window_particle_updates_per_s": wl.n * args.steps / el_prof,
             "g2p_roofline": roofline_of(kern, "g2p", npart, blocks, TRAFFIC_C2, VALU_PER_WAVE_C2),
             "active_blocks": blocks}
    want = set() if args.no_extras else {e.strip() for e in args.extras.split(",") if e.strip()}
    if "q32" in want:
        log("[bench] deterministic Q32.32 mode")
        wq, elq, _, _, _, _, _, shq = run_main(W, aemh, args.steps, args.warmup,
                                               accum=aemh.ACCUM_Q32_32)
        extra["q32_particle_updates_per_s"] = wq.n * args.steps / elq
        extra["q32_launch_shape"] = {k: shq[k] for k in ("p2g_tiles", "p2g_partials", "p2g_layers",
                                                         "g2p_wg", "upd_wgs")}
    if "c3" in want:
        log("[bench] C3 16M particles")
        extra["c3"] = extra_c3(W, aemh)
    if "c4" in want:
        log("[bench] C4 batched vs plain")
        extra["c4"] = extra_batched(W, aemh)
    if "c5" in want:
        log("[bench] C5 mixed precision")
        extra["c5"] = extra_c5(W, aemh)
        if not args.no_cpu:
            log("[bench] C5 error against the FP64 oracle, 65,536 particles x 500 steps")
            extra["c5"]["error_vs_fp64"] = c5_error(W, aemh)
    if "tiers" in want:
        log("[bench] LOD tier tick at engine scale")
        extra["tiers"] = extra_tiers(W, aemh, with_cpu=not args.no_cpu)
    if "fixed" in want:
        log("[bench] all-integer FixedPoint MpmSimulation, 1M particles")
        extra["fixed"] = extra_fixed(W, aemh, with_cpu=not args.no_cpu)
        log("[bench] all-integer FixedPoint MpmSimulation, 16M particles (same density, 160^3)")
        extra["fixed_16m"] = extra_fixed(W, aemh, n=1 << 24, res=160, with_cpu=False)
    line["extra"] = extra
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle)")
        line["cpu_baseline"] = cpu_baseline(W, aemh)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
