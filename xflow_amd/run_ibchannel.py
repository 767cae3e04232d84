"""The immersed-boundary channel (256x128x128 cells, icosphere of 2562
vertices, direct forcing, passive scalar) advanced twice from the same state:
once with the plain FP64 CG pressure solve, once with the V-cycle/network
flexible CG in mixed precision; reports ms per step and the relative L2
difference of velocity and scalar between the two runs.

    python -m xflow_amd.run_ibchannel [--ny 128] [--steps 200] [--tol 1e-8]
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

import xflow_amd as xf
from xflow_amd.drivers import IBChannel
from xflow_amd.unet import UNet

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--weights", default=os.path.join(ROOT, "xflow_amd", "weights", "unet_coarse64.npy"),
                    help="coarsest-grid network")
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    ctx = xf.Context(0)
    net = UNet(ctx, seed=0)
    log("coarsest-grid network: %s" % args.weights)
    net.set_params(np.load(args.weights))
    ref = IBChannel(ctx, ny=args.ny, subdivisions=4, tol=args.tol)
    nn = IBChannel(ctx, ny=args.ny, subdivisions=4, tol=args.tol, solver=xf.XF_FCG_MGNN, precond=net.h)
    times = {"fp64": 0.0, "nn": 0.0}
    hist = []
    for s in range(args.steps):
        for name, ch in (("fp64", ref), ("nn", nn)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ch.step()
            torch.cuda.synchronize()
            times[name] += time.perf_counter() - t0
        if (s + 1) % args.every == 0 or s + 1 == args.steps:
            dv = [float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) for a, b in zip(nn.vel, ref.vel)]
            dc = float(torch.linalg.norm(nn.c - ref.c) / torch.linalg.norm(ref.c))
            rec = {"step": s + 1, "vel_rel_l2": dv, "scalar_rel_l2": dc, "fp64": ref.last, "nn": nn.last}
            hist.append(rec)
            log("step %d: velocity rel L2 %s, scalar %.2e; pressure fp64 %d it %.1f ms, nn %d inner %.1f ms, "
                "div max %.2e / %.2e" % (s + 1, ["%.2e" % v for v in dv], dc, ref.last["pressure_iters"],
                                           ref.last["pressure_ms"], nn.last["inner_iters"], nn.last["pressure_ms"],
                                           ref.last["div_max"], nn.last["div_max"]))
    out = {"cells": [2 * args.ny, args.ny, args.ny], "steps": args.steps, "tol": args.tol,
           "ms_per_step": {k: round(v * 1e3 / args.steps, 2) for k, v in times.items()}, "history": hist}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
