"""The 3-D lid-driven cavity on the GPU: 64^3 cells, unit cube, Re = 100, lid
u = 1, dt = 1e-3, 100 steps, each step LOD transport of the three velocity
components, the pressure Poisson equation (Neumann walls) to a 1e-8 relative
residual and the velocity correction.  Reports ms per step (and of its
pressure solve), pressure iterations and the largest divergence, every
`--every` steps.

    python -m xflow_amd.run_cavity [--n 64] [--steps 100] [--tol 1e-8] [--solver cg|cg_mp|cg3|cycle] [--graphs]

`u_max` is the largest |u| over the x-velocity unknowns (ghost cells behind
the lid hold 2 - u and are left out).
"""
import argparse
import json
import os
import sys
import time

if __package__ in (None, ""):  # run as a script: make the package importable
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import torch  # noqa: E402

import xflow_amd as xf  # noqa: E402
from xflow_amd import cases  # noqa: E402
from xflow_amd.drivers import Cavity  # noqa: E402

SOLVERS = {"cg": xf.XF_CG, "cg_mp": xf.XF_CG_MP, "cg3": xf.XF_CG3, "cycle": xf.XF_FCG_MGNN}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--every", type=int, default=20)
    ap.add_argument("--solver", choices=sorted(SOLVERS), default="cg")
    ap.add_argument("--graphs", action="store_true",
                    help="run the parts of the step around the pressure solve as HIP graphs")
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    ctx = xf.Context(0)
    cav = Cavity(ctx, n=args.n, tol=args.tol, solver=SOLVERS[args.solver], graphs=args.graphs)
    inner = torch.from_numpy(cav.vcodes[0] == cases.ACTUAL).cuda()
    hist, wall, pms, iters = [], 0.0, 0.0, 0
    for s in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = cav.step()
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        pms += info["pressure_ms"]
        iters += info["pressure_iters"]
        if (s + 1) % args.every == 0 or s + 1 == args.steps:
            u = cav.vel[0][inner]
            rec = {"step": s + 1, "pressure_iters": info["pressure_iters"], "div_max": info["div_max"],
                   "resid_true": info["resid_true"], "u_max": float(u.abs().max())}
            hist.append(rec)
            log("step %d: pressure %d iterations, div max %.2e, true residual %.2e, max |u| %.3f" %
                (s + 1, rec["pressure_iters"], rec["div_max"], rec["resid_true"], rec["u_max"]))
    out = {"cells": [args.n] * 3, "steps": args.steps, "tol": args.tol, "solver": args.solver, "graphs": args.graphs,
           "ms_per_step": round(wall * 1e3 / args.steps, 3), "pressure_ms_per_step": round(pms / args.steps, 3),
           "pressure_iters_mean": iters / args.steps, "history": hist}
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
