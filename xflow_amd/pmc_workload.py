"""A short, fixed workload for hardware-counter runs (rocprofv3 --pmc).

It launches a plain device copy of a known size, after evicting its source
from the Infinity Cache (the calibration kernel: its FETCH_SIZE and
WRITE_SIZE should each read the copied bytes), then a few
iterations of the FP64 and FP32 throughput CG on the 256^3 random-obstacle
pressure problem, so every counter pass sees the same dispatches:

    rocprofv3 --pmc FETCH_SIZE -d OUT -o pmc -- python3 <repo>/xflow_amd/pmc_workload.py
    rocprofv3 --pmc WRITE_SIZE -d OUT -o pmc -- python3 <repo>/xflow_amd/pmc_workload.py

    python3 xflow_amd/pmc_workload.py [--n 256] [--iters 8] [--copy-mib 512]

(run from any directory: the profiler is started from /tmp).
"""
import argparse
import os
import sys
import time

if __package__ in (None, ""):  # run as a script: make the package importable
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--iters", type=int, default=8)
    ap.add_argument("--copy-mib", type=int, default=512)
    args = ap.parse_args()

    import torch
    import xflow_amd as xf
    from xflow_amd import cases

    t0 = time.time()
    nbytes = args.copy_mib << 20
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    # evict src from the 256 MB Infinity Cache first (fill 1 GiB elsewhere), so
    # the copy's reads come from HBM
    flush = torch.empty((1 << 30) // 8, dtype=torch.float64, device="cuda")
    flush.fill_(2.0)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del flush
    print("copy: %d bytes read, %d bytes written" % (nbytes, nbytes), flush=True)

    case = cases.obstacle_case(args.n, seed=0)
    ctx = xf.Context(0)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = torch.from_numpy(case["b"]).cuda()
    x0 = torch.zeros_like(b)
    x = torch.zeros_like(b)
    for method in (xf.XF_CG, xf.XF_CG_MP):
        o = xf.solve_opts(method, tol=1e-30, tol_mode=xf.XF_TOL_REL, project_mean=1, maxit=args.iters,
                          inner_tol=1e-30)
        st = g.solve(b, x0, x, o)
        print("method %d: %d iterations (%d inner), status %d" % (method, st.iters, st.inner_iters, st.status),
              flush=True)
    torch.cuda.synchronize()
    g.close()
    ctx.close()
    print("done in %.1f s; interior cells %d, array cells %d" % (time.time() - t0, args.n ** 3, (args.n + 2) ** 3),
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
