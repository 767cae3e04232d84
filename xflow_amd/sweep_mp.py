"""Mixed-precision pressure solves over the 256^3 random-obstacle problems:
for each seed, the FP64 throughput CG to a 1e-12 relative residual is the
reference field; the mixed-precision CG (FP32 sweeps, FP64 refinement), the
flexible CG with the full-depth classical V-cycle and with the cycle cut at
64^3 and the trained network there solve to 1e-10, and their pressure is
compared with the reference (relative L2 of the mean-free field over the
fluid).  Times are the solvers' own HIP events (the multigrid hierarchy is
built before the timed solves).

    python -m xflow_amd.sweep_mp [--n 256] [--seeds 0-15] [--tol 1e-10]
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

import xflow_amd as xf
from xflow_amd import cases
from xflow_amd.unet import UNet

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def seeds_of(text):
    a, _, b = text.partition("-")
    return list(range(int(a), int(b or a) + 1))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seeds", default="0-15")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--weights", default=os.path.join(ROOT, "xflow_amd", "weights", "unet_coarse64.npy"))
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    ctx = xf.Context(0)
    net = UNet(ctx, seed=0)
    net.set_params(np.load(args.weights))
    rows = []
    for seed in seeds_of(args.seeds):
        t0 = time.time()
        case = cases.obstacle_case(args.n, seed=seed)
        g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
        b = torch.from_numpy(case["b"]).cuda()
        z0 = torch.zeros_like(b)
        m = torch.from_numpy(case["codes"] == 1).cuda()

        def field(x):
            v = x[m]
            return v - v.mean()

        xr = torch.zeros_like(b)
        st = g.solve(b, z0, xr, xf.solve_opts(xf.XF_CG, tol=1e-12, tol_mode=xf.XF_TOL_REL, project_mean=1))
        ref = field(xr)
        row = {"seed": seed, "fluid": int(m.sum()), "fp64_cg": {"iters": st.iters, "ms": round(st.ms, 2),
                                                                 "true_residual": st.resid_true}}
        g.mg_levels()  # builds the hierarchy outside the timed solves
        for name, method, p in (("mixed_precision_cg", xf.XF_CG_MP, None), ("cycle_mp", xf.XF_FCG_MGNN, None),
                                ("nn_cycle_mp", xf.XF_FCG_MGNN, net)):
            x = torch.zeros_like(b)
            o = xf.solve_opts(method, tol=args.tol, tol_mode=xf.XF_TOL_REL, project_mean=1, inner_tol=1e-4)
            o.precond = p.h if p is not None else None
            st = g.solve(b, z0, x, o)
            err = float(torch.linalg.norm(field(x) - ref) / torch.linalg.norm(ref))
            row[name] = {"outer": st.iters, "inner": st.inner_iters, "ms": round(st.ms, 2), "status": st.status,
                         "fallback": st.fallback, "true_residual": st.resid_true, "rel_l2_vs_fp64": err}
        g.close()
        rows.append(row)
        log("seed %2d: fp64 %d it %.0f ms | " % (seed, row["fp64_cg"]["iters"], row["fp64_cg"]["ms"]) + " | ".join(
            "%s %d/%d %.1f ms err %.1e" % (k, row[k]["outer"], row[k]["inner"], row[k]["ms"], row[k]["rel_l2_vs_fp64"])
            for k in ("mixed_precision_cg", "cycle_mp", "nn_cycle_mp")) + " (%.1f s)" % (time.time() - t0))
    summary = {"fp64_cg_ms_mean": float(np.mean([r["fp64_cg"]["ms"] for r in rows])),
               "fp64_cg_iters_mean": float(np.mean([r["fp64_cg"]["iters"] for r in rows]))}
    for k in ("mixed_precision_cg", "cycle_mp", "nn_cycle_mp"):
        summary[k] = {"ms_mean": float(np.mean([r[k]["ms"] for r in rows])),
                      "inner_mean": float(np.mean([r[k]["inner"] for r in rows])),
                      "rel_l2_vs_fp64_max": float(max(r[k]["rel_l2_vs_fp64"] for r in rows)),
                      "true_residual_max": float(max(r[k]["true_residual"] for r in rows)),
                      "fallbacks": int(sum(r[k]["fallback"] for r in rows))}
    print(json.dumps({"n": args.n, "tol": args.tol, "summary": summary, "rows": rows}))


if __name__ == "__main__":
    main()
