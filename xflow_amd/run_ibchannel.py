"""The immersed-boundary channel (256x128x128 cells, icosphere of 2562
vertices, direct forcing, passive scalar) advanced side by side from the same
state with different pressure solvers: the plain FP64 CG ("fp64"), the
flexible CG in mixed precision with the full-depth classical V-cycle
("cycle") and with the cycle cut at 64^3 and the trained network there
("network").  The parts of each step around the solve run as HIP graphs
(--no-graphs: eagerly).  Reports ms per step (after --warmup steps), the pressure solve's ms,
iterations and true residual per step, and the relative L2 difference of velocity and scalar
of each run against the FP64 one.

    python -m xflow_amd.run_ibchannel [--ny 128] [--steps 200] [--tol 1e-8] [--solvers fp64,cycle,network]
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
    ap.add_argument("--warmup", type=int, default=2,
                    help="steps left out of the timing (the first runs eagerly, the second captures the graphs)")
    ap.add_argument("--solvers", default="fp64,cycle,network")
    ap.add_argument("--no-graphs", action="store_true",
                    help="run the parts of the step around the pressure solve eagerly instead of as HIP graphs")
    ap.add_argument("--no-verify", action="store_true",
                    help="skip the per-step true residual (one FP64 operator pass after each solve)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "xflow_amd", "weights", "unet_coarse64.npy"),
                    help="coarsest-grid network")
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    names = [s for s in args.solvers.split(",") if s]
    assert names and names[0] == "fp64" and set(names) <= {"fp64", "cycle", "network"}, \
        "--solvers: fp64 first, then any of cycle, network"
    ctx = xf.Context(0)
    runs = {}
    for name in names:
        if name == "fp64":
            ch = IBChannel(ctx, ny=args.ny, subdivisions=4, tol=args.tol, graphs=not args.no_graphs)
        elif name == "cycle":
            ch = IBChannel(ctx, ny=args.ny, subdivisions=4, tol=args.tol, solver=xf.XF_FCG_MGNN,
                           graphs=not args.no_graphs)
        else:
            net = UNet(ctx, seed=0)
            log("coarsest-grid network: %s" % args.weights)
            net.set_params(np.load(args.weights))
            ch = IBChannel(ctx, ny=args.ny, subdivisions=4, tol=args.tol, solver=xf.XF_FCG_MGNN, precond=net.h,
                           graphs=not args.no_graphs)
            ch.net = net  # keep the network alive with its run
        ch.verify = not args.no_verify
        runs[name] = ch
    times = {k: 0.0 for k in names}
    acc = {k: {"pressure_ms": 0.0, "iters": 0, "inner": 0, "resid_true_max": 0.0} for k in names}
    hist = []
    ref = runs["fp64"]
    for s in range(args.steps):
        for name, ch in runs.items():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ch.step()
            torch.cuda.synchronize()
            if s < args.warmup:
                continue
            times[name] += time.perf_counter() - t0
            a = acc[name]
            a["pressure_ms"] += ch.last["pressure_ms"]
            a["iters"] += ch.last["pressure_iters"]
            a["inner"] += ch.last["inner_iters"]
            a["resid_true_max"] = max(a["resid_true_max"], ch.last["resid_true"])
        if (s + 1) % args.every == 0 or s + 1 == args.steps:
            rec = {"step": s + 1}
            for name, ch in runs.items():
                rec[name] = dict(ch.last)
                if name != "fp64":
                    rec[name]["vel_rel_l2"] = [float(torch.linalg.norm(a - b) / torch.linalg.norm(b))
                                               for a, b in zip(ch.vel, ref.vel)]
                    rec[name]["scalar_rel_l2"] = float(torch.linalg.norm(ch.c - ref.c) / torch.linalg.norm(ref.c))
            hist.append(rec)
            log("step %d: " % (s + 1) + "; ".join(
                "%s %d it (%d inner) %.1f ms, true resid %.1e, div max %.1e%s" % (
                    n, r["pressure_iters"], r["inner_iters"], r["pressure_ms"], r["resid_true"], r["div_max"],
                    "" if n == "fp64" else ", vel diff %.1e, scalar diff %.1e" % (
                        max(r["vel_rel_l2"]), r["scalar_rel_l2"]))
                for n, r in rec.items() if n != "step"))
    per = {}
    nt = max(1, args.steps - args.warmup)
    for name in names:
        a = acc[name]
        per[name] = {"ms_per_step": round(times[name] * 1e3 / nt, 2),
                     "pressure_ms_per_step": round(a["pressure_ms"] / nt, 2),
                     "non_pressure_ms_per_step": round((times[name] * 1e3 - a["pressure_ms"]) / nt, 2),
                     "iters_per_step": round(a["iters"] / nt, 1),
                     "inner_iters_per_step": round(a["inner"] / nt, 1),
                     "resid_true_max": a["resid_true_max"] if ch.verify else None}
    out = {"cells": [2 * args.ny, args.ny, args.ny], "steps": args.steps, "timed_steps": nt, "tol": args.tol,
           "graphs": not args.no_graphs, "solvers": per,
           "history": hist}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
