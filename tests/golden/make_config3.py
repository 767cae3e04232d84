"""Oracle summary of the 256^3 random-obstacle pressure solve (seed 0: 1-4
solid spheres or boxes, smoothed random MAC divergence, Neumann walls), the
workload of the mixed-precision and learned-preconditioner solvers: the
oracle's FP64 single-reduction CG (oracle/xf_oracle.c xo_cg_cg, the same
recurrence as the GPU's plain CG; its rows are bit-exact against the compiled
reference's) from x0 = 0 to a 1e-12 relative residual with the mean over the
unknowns projected out.  A 256^3 field is 137 MB, so the fixture keeps the
mean-free field's 2-norm over the fluid and its mean-free values at 512 fixed
fluid cells, with the iteration count.

    python tests/golden/make_config3.py      (about 10 minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import xf_oracle as xo  # noqa: E402
from xflow_amd import cases  # noqa: E402

SAMPLES = 512
RTOL = 1e-12


def main():
    xo.set_threads(min(16, os.cpu_count() or 1))
    t0 = time.time()
    case = cases.obstacle_case(256, seed=0)
    fluid = np.flatnonzero(case["codes"].ravel() == 1)
    idx = np.sort(np.random.default_rng(3).choice(fluid, SAMPLES, replace=False)).astype(np.int64)
    og = xo.Grid(case["hx"], case["hy"], case["hz"], case["codes"])
    x, it, res, st = xo.cg_cg(og, case["b"], np.zeros_like(case["b"]), RTOL, 200000, project_mean=True)
    assert st == 0, st
    xf = x.ravel()
    mean = xf[fluid].mean()
    out = {"idx": idx, "iters": np.array([it], dtype=np.int64), "rtol": np.array([RTOL]),
           "norm": np.array([np.linalg.norm(xf[fluid] - mean)]), "x": xf[idx] - mean,
           "fluid_cells": np.array([fluid.size], dtype=np.int64)}
    np.savez_compressed(os.path.join(HERE, "config3_obstacle256.npz"), **out)
    print("oracle CG-CG 256^3 seed 0: %d iterations to %.0e, ||x - mean|| %.12e (%.0f s)" % (
        it, RTOL, out["norm"][0], time.time() - t0), flush=True)


if __name__ == "__main__":
    main()
