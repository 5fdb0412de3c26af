"""Write the committed golden fixtures: FP64 oracle outputs on small seeded
scenes (inputs regenerate from the seeds in icempm_2d_sp_amd.scenes).

    python tests/golden/make_fixtures.py

The reference has no tests or expected outputs for this path, so these are
regression vectors of the oracle (which tests/test_oracle.py pins against
closed forms); the GPU tests compare the HIP path with them too.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from icempm_2d_sp_amd import scenes  # noqa: E402

CASES = {
    # name: (scene kwargs, substeps)
    "dyn_small_1": (dict(seed=101, gx=20, gy=16, fill=0.9), 1),
    "dyn_small_25": (dict(seed=102, gx=20, gy=16, fill=0.9, speed=400.0), 25),
    "dyn_small_coh_10": (dict(seed=103, gx=18, gy=18, fill=0.9, with_cohesion=True), 10),
}


def scene_for(name):
    kw, steps = CASES[name]
    return scenes.dynamic_scene(**kw), steps


def compute(name):
    sc, steps = scene_for(name)
    o = oracle.Oracle.from_scene(sc, 64)
    for k in range(steps):
        o.substep(1e-4 * k, k == steps - 1)
    pts = o.points_f64()
    mpp, vel, frc = o.grid()
    return dict(points=pts, forces=frc, error=np.uint32(o.error), disabled=np.int32(o.disabled),
                steps=np.int32(steps))


def main():
    for name in CASES:
        out = compute(name)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(path, os.path.getsize(path))


if __name__ == "__main__":
    main()
