"""Summaries of the compiled reference's own CG and BiCG on the 128^3 sphere
workload (solid sphere r = 0.2 at the cube centre, smoothed random MAC
divergence from seed 0): TSLAE::SolveByConjugateGradients (XSLAE.cpp:56-157)
and TSLAE::SolveByBiConjugateGradients (XSLAE.cpp:159-348, with the harness's
operator~ the identity on this symmetric mask) at the reference's absolute
tolerances, the right-hand side scaled as in make_config2.py (CG ||b|| = 1e3,
BiCG ||b|| = 1) so that each tolerance is 1e-8 relative, x0 = 0.  Stored as in
config2_sphere128.npz: iteration count (from the harness's operator-sweep
count), the field's 2-norm over the fluid and its values at the same 512 fluid
cells.

    make -C oracle ref && python tests/golden/make_config2_reference.py   (a few minutes, serial)
"""
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import xf_oracle as xo  # noqa: E402
from xfb import read_xfb, write_xfb  # noqa: E402
from xflow_amd import cases  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
NORMS = {"cg": 1e3, "bicg": 1.0}
# iterations from the operator sweeps the harness counts (see tests/test_oracle_golden.py)
SWEEPS_TO_ITERS = {"cg": lambda s: (s - 3) // 2, "bicg": lambda s: (s - 1) // 3}


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    case = cases.sphere_case(128, seed=0)
    codes = np.ascontiguousarray(case["codes"], dtype=np.int32)
    fluid = np.flatnonzero(codes.ravel() == 1)
    summary = np.load(os.path.join(HERE, "config2_sphere128.npz"))
    idx = summary["idx"]
    nodes = lambda h: np.concatenate([[0.0], np.cumsum(h)])
    out = {"idx": idx}
    for name, norm in NORMS.items():
        b = case["b"] * (norm / np.linalg.norm(case["b"]))
        inp = {"x": nodes(case["hx"]), "y": nodes(case["hy"]), "z": nodes(case["hz"]), "codes": codes,
               "space_mask": xo.space_mask_of(codes), "bvec": b, "x0": np.zeros_like(b)}
        t0 = time.time()
        with tempfile.TemporaryDirectory() as td:
            fi, fo = os.path.join(td, "in.xfb"), os.path.join(td, "out.xfb")
            write_xfb(fi, inp)
            p = subprocess.run([HARNESS, name, fi, fo], stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
            if p.returncode != 0:
                raise RuntimeError("harness %s failed: %s" % (name, p.stderr[-400:]))
            res = read_xfb(fo)
        sweeps, rest = int(res["sweeps"][0]), int(res["sweeps"][1])
        assert rest == 0, (name, rest)
        xf = np.asarray(res["x"]).ravel()
        key = "cg3" if name == "cg" else name
        out[key + "_iters"] = np.array([SWEEPS_TO_ITERS[name](sweeps)], dtype=np.int64)
        out[key + "_norm"] = np.array([np.linalg.norm(xf[fluid])])
        out[key + "_x"] = xf[idx]
        print("reference %s: %d sweeps, %d iterations, ||x|| %.12e (%.0f s)" % (
            name, sweeps, out[key + "_iters"][0], out[key + "_norm"][0], time.time() - t0), flush=True)
    np.savez_compressed(os.path.join(HERE, "config2_sphere128_reference.npz"), **out)


if __name__ == "__main__":
    main()
