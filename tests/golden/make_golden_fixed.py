"""Writes the golden fixtures of the all-integer FixedPoint solver
(fixed_*.npz) from the oracle (oracle/fixedsim.cpp, the serial restatement
of ae_physics/src/mpm.rs MpmSimulation).

Each fixture holds the configuration as the raw bytes of the MpmConfig
mirror, the number of steps, and after them every particle record (29 raw
int64, flags, ids, colour), the scalar state and the fracture events.  The
reference is Rust and cannot be run here; these vectors freeze the oracle's
answers, so a later change to it, or a GPU run, is checked against the same
numbers.

    python tests/golden/make_golden_fixed.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import fixedsim as X  # noqa: E402

CASES = {
    # name: (config, steps)
    "fixed_small_elastoplastic_5": (X.small_config(), 5),
    "fixed_snow125x2_3": (X.small_config(model=X.SNOW, model_param=X.from_f32(0.05), particle_count=125,
                                         substeps=2), 3),
    "fixed_fluid125_8": (X.small_config(model=X.FLUID, model_param=X.from_f32(0.01), particle_count=125), 8),
    "fixed_fracture_5": (X.small_config(enable_fracture=1, fracture_strain=X.from_f32(0.01)), 5),
}


def run(cfg, steps):
    s = X.Sim(cfg)
    s.step(steps)
    d = s.read()
    st = s.state()
    ev = s.events()
    out = dict(q=d["q"], flags=d["flags"], ids=d["ids"], color=d["color"],
               step_count=np.uint64(st["step_count"]), time=np.float64(st["time"]),
               next_id=np.uint64(st["next_id"]), ev_time=ev["time"], ev_pos=ev["pos"],
               ev_strain=ev["strain"], ev_frags=ev["frags"])
    s.close()
    return out


def main():
    for name, (cfg, steps) in CASES.items():
        out = run(cfg, steps)
        np.savez_compressed(os.path.join(HERE, name + ".npz"),
                            config=np.frombuffer(bytes(cfg), np.uint8), steps=np.uint32(steps), **out)
        print(name, "written:", out["q"].shape[0], "particles,", len(out["ev_frags"]), "events")


if __name__ == "__main__":
    main()
