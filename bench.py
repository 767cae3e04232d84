"""Pressure-solve benchmark: ms per solve to 1e-8 relative residual at 256^3.

A step is one pressure solve of the 256^3 random-obstacle Poisson workload
(seed 0: 1-4 solid spheres/boxes, smoothed random MAC divergence as the
right-hand side, Neumann walls) with inputs resident in HBM.  The JSON line
reports the plain FP64 GPU CG and, where built, the mixed-precision solver.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
    python bench.py --pmc    (the fixed workload of the counter passes only)
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# measured HBM traffic of one k_cgcg_fused<double,double,4,10,1> launch at
# 256^3 (mean of an even and an odd iteration; rocprofv3 counters scaled by a
# copy of known size, as recorded in profiles/cg_kernels_traffic_pmc.md)
KF_TRAFFIC_BYTES = 733.9e6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _size(case):
    return "x".join(str(d - 2) for d in case["codes"].shape)


def host_cpu():
    """The host the CPU baselines ran on: model name and logical CPUs."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "usable_cpus": usable}


def cpu_port(case, gpu_iters, sample_iters):
    """The CPU oracle's FP64 CG (OpenMP stencil, serial dot products), timed
    for a bounded number of iterations on all usable cores (at most 16) and on
    one, and scaled to the GPU's iteration count."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import xf_oracle as xo
    host = host_cpu()
    threads = min(16, host["usable_cpus"])
    og = xo.Grid(case["hx"], case["hy"], case["hz"], case["codes"])
    x0 = np.zeros_like(case["b"])
    per = {}
    for nt in (threads, 1):
        xo.set_threads(nt)
        t0 = time.perf_counter()
        _, it, _, _ = xo.cg_cg(og, case["b"], x0, 1e-30, sample_iters, project_mean=True)
        per[nt] = (time.perf_counter() - t0) * 1e3 / max(it, 1)
    xo.set_threads(threads)
    out = {"value": per[threads] * gpu_iters, "unit": "ms/solve", "cores": threads, "kind": "port",
           "one_thread_value": per[1] * gpu_iters,
           "sample": "%d iterations of the oracle FP64 single-reduction CG (xo_cg_cg, OpenMP stencil, serial "
                     "dot products) at the same %s case: %.1f ms/iteration on %d threads, %.1f ms on one, scaled "
                     "by the GPU solve's %d iterations" % (it, _size(case), per[threads], threads, per[1], gpu_iters)}
    out.update(host)
    return out


REF_HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def cpu_reference(case, gpu_iters, sample_iters):
    """The reference's own CG (TSLAE::SolveByConjugateGradients, XSLAE.cpp:56-157,
    compiled from its sources into oracle/_ref/ref_harness; serial: the
    reference builds without OpenMP and its CG has no parallel region) on the
    same case, stopped after `sample_iters` iterations, its steady-state time
    per iteration scaled to the GPU solve's iteration count.  None when the
    harness was not built."""
    if not os.path.exists(REF_HARNESS):
        return None
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import xf_oracle as xo
    from xfb import read_xfb, write_xfb
    codes = np.ascontiguousarray(case["codes"], dtype=np.int32)
    nodes = lambda h: np.concatenate([[0.0], np.cumsum(h)])
    inp = {"x": nodes(case["hx"]), "y": nodes(case["hy"]), "z": nodes(case["hz"]), "codes": codes,
           "space_mask": xo.space_mask_of(codes), "bvec": case["b"], "x0": np.zeros_like(case["b"]),
           "max_sweeps": np.array([3 + 2 * sample_iters], dtype=np.int32)}
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.xfb"), os.path.join(td, "out.xfb")
        write_xfb(fi, inp)
        del inp
        p = subprocess.run([REF_HARNESS, "cg_timed", fi, fo], stdout=subprocess.DEVNULL, stderr=subprocess.PIPE,
                           text=True, timeout=600)
        if p.returncode != 0:
            raise RuntimeError("reference CG timing failed: %s" % p.stderr[-400:])
        o = read_xfb(fo)
    it = int(o["iters"][0])
    per_iter_ms = float(o["seconds"][0]) * 1e3 / it
    return {**host_cpu(), "value": per_iter_ms * gpu_iters, "unit": "ms/solve", "cores": 1, "kind": "reference",
            "sample": "%d iterations of the reference's TSLAE::SolveByConjugateGradients (XSLAE.cpp:56-157, "
                      "compiled -O2 from its sources; serial, as the reference builds it) at the same %s case "
                      "after its first three operator sweeps, %.1f ms/iteration, scaled by the GPU solve's %d "
                      "iterations (set-up of its masks and vectors, %.1f s, not counted)"
                      % (it, _size(case), per_iter_ms, gpu_iters, float(o["setup_seconds"][0]))}


def load_network(ctx, path):
    """A trained network: saved parameters if present, else None."""
    from xflow_amd.unet import UNet
    if not os.path.exists(path):
        log("no trained network at %s: skipping its solve" % path)
        return None
    net = UNet(ctx, seed=0)
    net.set_params(np.load(path))
    return net


def pmc_workload(n, iters, copy_mib):
    """A short fixed workload for hardware-counter passes (rocprofv3 --pmc):
    a plain device copy of known size after a 1 GiB fill that evicts its
    source from the Infinity Cache (the calibration dispatch: its FETCH_SIZE
    and WRITE_SIZE should each read the copied bytes), then `iters`
    iterations each of the FP64 and FP32 throughput CG on the n^3
    random-obstacle problem, so every counter pass sees the same dispatches."""
    import torch
    import xflow_amd as xf
    from xflow_amd import cases
    t0 = time.time()
    nbytes = copy_mib << 20
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    flush = torch.empty((1 << 30) // 8, dtype=torch.float64, device="cuda")
    flush.fill_(2.0)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del flush
    log("copy: %d bytes read, %d bytes written" % (nbytes, nbytes))
    case = cases.obstacle_case(n, seed=0)
    ctx = xf.Context(0)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = torch.from_numpy(case["b"]).cuda()
    x0 = torch.zeros_like(b)
    x = torch.zeros_like(b)
    for method in (xf.XF_CG, xf.XF_CG_MP):
        o = xf.solve_opts(method, tol=1e-30, tol_mode=xf.XF_TOL_REL, project_mean=1, maxit=iters, inner_tol=1e-30,
                          verify=0)
        st = g.solve(b, x0, x, o)
        log("method %d: %d iterations (%d inner), status %d" % (method, st.iters, st.inner_iters, st.status))
    torch.cuda.synchronize()
    g.close()
    ctx.close()
    log("done in %.1f s; interior cells %d, array cells %d" % (time.time() - t0, n ** 3, (n + 2) ** 3))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--cpu-ref-iters", type=int, default=12)
    ap.add_argument("--plain-only", action="store_true")
    ap.add_argument("--skip-fullres", action="store_true")
    ap.add_argument("--skip-bicgstab", action="store_true",
                    help="leave out the two BiCGStab solves (reference statements and fused)")
    ap.add_argument("--io", action="store_true", help="also time the output files (flow VTK, .npy) of the solution")
    ap.add_argument("--inner-tol", type=float, default=1e-4)
    ap.add_argument("--nn-inner-tol", type=float, default=1e-4)
    ap.add_argument("--weights", default=os.path.join(ROOT, "xflow_amd", "weights", "unet_poisson.npy"))
    ap.add_argument("--coarse-weights", default=os.path.join(ROOT, "xflow_amd", "weights", "unet_coarse64.npy"))
    ap.add_argument("--pmc", action="store_true",
                    help="run only the fixed counter workload (calibration copy + a few CG iterations) and exit")
    ap.add_argument("--pmc-iters", type=int, default=8)
    ap.add_argument("--pmc-copy-mib", type=int, default=512)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every box in this setup has a single MI355X; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.pmc:
        pmc_workload(args.n, args.pmc_iters, args.pmc_copy_mib)
        return

    import torch
    import xflow_amd as xf
    from xflow_amd import cases, profile

    t0 = time.perf_counter()
    case = cases.obstacle_case(args.n, seed=args.seed)
    log("case generated in %.1f s: %d fluid cells" % (time.perf_counter() - t0, int((case["codes"] == 1).sum())))
    ctx = xf.Context(0)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = torch.from_numpy(case["b"]).cuda()
    x0 = torch.zeros_like(b)
    x = torch.zeros_like(b)
    fluid = int((case["codes"] == 1).sum())

    # Timed solves skip the library's after-solve check of the true residual
    # with the reference rows (opts.verify = 0: two operator passes and a
    # temporary field, not part of the solver); an untimed solve of the same
    # problem with the check on supplies the reported true residual.
    def solve(method, verify=0, **kw):
        o = xf.solve_opts(method, tol=1e-8, tol_mode=xf.XF_TOL_REL, project_mean=1, verify=verify, **kw)
        return g.solve(b, x0, x, o)

    resid_true = None
    for _ in range(args.warmup):
        st = solve(xf.XF_CG, verify=1)
        resid_true = st.resid_true
        log("warmup: %d iters, %.2f ms, true resid %.2e" % (st.iters, st.ms, st.resid_true))
    # timed region: K plain-FP64 solves (no per-kernel events inside it)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    iters = []
    for _ in range(args.steps):
        st = solve(xf.XF_CG)
        iters.append(st.iters)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    assert st.status == 0, st.status
    # kernel durations for the roofline: one more solve with HIP events on the
    # library's stream around a sample of the launches (two of every eight
    # iterations), outside the timed region
    profile.enable(ctx, True)
    solve(xf.XF_CG)
    kstats = profile.read(ctx)
    profile.enable(ctx, False)
    # streaming ceilings of this GPU for the kernels' access patterns (1 GiB buffers)
    ceiling = {name: round(profile.stream_ceiling(ctx, kind), 1)
               for kind, name in ((0, "read"), (1, "copy"), (2, "read2_write2"))}
    log("streaming ceilings (GB/s): %s" % ceiling)
    if resid_true is None:
        resid_true = solve(xf.XF_CG, verify=1).resid_true
    plain = {"ms": round(wall_ms, 3), "iterations": float(np.mean(iters)), "true_residual": resid_true}
    log("plain FP64 CG: %.2f ms/solve wall, %s iters, true resid %.2e" % (wall_ms, iters, resid_true))

    solvers = {"plain_fp64_cg": plain}
    # mixed precision: FP32 CG sweeps + FP64 refinement
    if not args.plain_only:
        st = solve(xf.XF_CG_MP, verify=1, inner_tol=args.inner_tol)
        mp_resid_true = st.resid_true
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = solve(xf.XF_CG_MP, inner_tol=args.inner_tol)
        torch.cuda.synchronize()
        mp_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        solvers["mixed_precision_cg"] = {"ms": round(mp_ms, 3), "outer": st.iters, "inner_iterations": st.inner_iters,
                                         "true_residual": mp_resid_true, "status": st.status}
        log("mixed precision: %.2f ms/solve, %d outer / %d inner, true resid %.2e" %
            (mp_ms, st.iters, st.inner_iters, mp_resid_true))

        def timed(o, label):
            o.verify = 1
            st = g.solve(b, x0, x, o)
            resid_true = st.resid_true
            o.verify = 0
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                st = g.solve(b, x0, x, o)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / args.steps
            rec = {"ms": round(ms, 3), "outer": st.iters, "inner_iterations": st.inner_iters,
                   "fallback": st.fallback, "true_residual": resid_true, "status": st.status,
                   "speedup_vs_plain_fp64": round(wall_ms / ms, 2)}
            log("%s: %.2f ms/solve, %d outer / %d inner, fallback %d, true resid %.2e" %
                (label, ms, st.iters, st.inner_iters, st.fallback, resid_true))
            return rec

        def opts(method, net):
            o = xf.solve_opts(method, tol=1e-8, tol_mode=xf.XF_TOL_REL, project_mean=1, inner_tol=args.nn_inner_tol)
            o.precond = net.h if net is not None else None
            return o

        coarse = load_network(ctx, args.coarse_weights)
        solvers["mp_fcg_cycle_no_network"] = timed(opts(xf.XF_FCG_MGNN, None),
                                                   "MP flexible CG, full-depth V-cycle, exact solve on the 4^3 grid")
        if coarse is not None:
            solvers["nn_preconditioned_mp"] = timed(opts(xf.XF_FCG_MGNN, coarse),
                                                    "MP flexible CG, V-cycle to 64^3, network solve there")
        full = load_network(ctx, args.weights)
        if full is not None and not args.skip_fullres:
            solvers["nn_fullres_mp"] = timed(opts(xf.XF_FCG_NN, full), "MP flexible CG, full-resolution network")
        if not args.skip_bicgstab:
            # the reference's BiCGStab (max |r| <= 1e-5 absolute) on b scaled to
            # max |b| = 1e3, one timed solve each: statement by statement and fused
            bs = b * (1e3 / float(b.abs().max()))
            for name, method in (("bicgstab_reference_statements", xf.XF_BICGSTAB),
                                 ("bicgstab_fused", xf.XF_BICGSTAB_T)):
                o = xf.solve_opts(method)
                o.verify = 0
                # a two-iteration solve first: workspaces and the per-grid
                # tables are built on first use, outside the timed solve
                w = xf.solve_opts(method, maxit=2)
                w.verify = 0
                g.solve(bs, x0, x, w)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                st = g.solve(bs, x0, x, o)
                torch.cuda.synchronize()
                ms = (time.perf_counter() - t0) * 1e3
                solvers[name] = {"ms": round(ms, 2), "iterations": st.iters, "status": st.status,
                                 "max_abs_residual": st.resid, "ms_per_iteration": round(ms / max(st.iters, 1), 4)}
                log("%s: %.1f ms, %d iterations, max |r| %.2e" % (name, ms, st.iters, st.resid))

    # host-buffer boundary: the drop-in classes upload b and download x once per
    # solve; time that PCIe round trip (pinned host memory) to quote the
    # PCIe-inclusive rate beside the resident-input value
    hb = torch.empty(b.shape, dtype=torch.float64, pin_memory=True)
    hb.copy_(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        b.copy_(hb, non_blocking=True)
        hb.copy_(x, non_blocking=True)
    torch.cuda.synchronize()
    pcie_ms = (time.perf_counter() - t0) * 1e3 / 3
    log("PCIe round trip of b and x (%.0f MB each): %.2f ms" % (b.numel() * 8 / 1e6, pcie_ms))

    io = None
    if args.io:
        # the reference's output path on the solved pressure: node-centred
        # flow VTK (GPU node averages + threaded printf) and a .npy of p
        import tempfile
        with tempfile.TemporaryDirectory() as td:
            coords = [np.concatenate([[0.0], np.cumsum(h)]) for h in (case["hx"], case["hy"], case["hz"])]
            zero = torch.zeros_like(x)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xf.write_vtk_flow(ctx, os.path.join(td, "flow.vtk"), g, *coords, zero, zero, zero, p=x)
            vtk_s = time.perf_counter() - t0
            vtk_mb = os.path.getsize(os.path.join(td, "flow.vtk")) / 1e6
            t0 = time.perf_counter()
            xf.npy_save(ctx, os.path.join(td, "p.npy"), x)
            npy_s = time.perf_counter() - t0
            npy_mb = os.path.getsize(os.path.join(td, "p.npy")) / 1e6
        io = {"flow_vtk_s": round(vtk_s, 3), "flow_vtk_mb": round(vtk_mb, 1),
              "npy_s": round(npy_s, 3), "npy_mb": round(npy_mb, 1)}
        log("output files: flow VTK %.1f MB in %.2f s, .npy %.1f MB in %.3f s" % (vtk_mb, vtk_s, npy_mb, npy_s))

    interior = (args.n) ** 3
    # algorithmic bytes per interior cell of one iteration's kernel: r, p_old
    # and the flag byte read, p and r written (33 B), plus x read and written
    # on odd iterations (49 B; the sampled launches alternate, so the mean is
    # 41 B).  The fused kernel forms the next iteration's dot products from
    # the same bytes; where the rows are too long for it, a separate dots
    # kernel reads r and the flags again (9 B).
    kf = kstats.get("cg_fused_fp64", {"ms": float("nan"), "count": 0})
    ka = kstats.get("cg_dots_fp64", {"ms": float("nan"), "count": 0})
    kb = kstats.get("cg_step_fp64", {"ms": float("nan"), "count": 0})
    fused = kf["count"] > 0
    kd = kf if fused else kb
    kd_bytes = 41.0 * interior
    kd_gbs = kd_bytes / (kd["ms"] * 1e-3) / 1e9
    it_bytes = kd_bytes if fused else 50.0 * interior
    it_us = kd["ms"] * 1e3 if fused else (ka["ms"] + kb["ms"]) * 1e3
    it_mean = plain["iterations"]
    mcups = fluid * it_mean / (wall_ms * 1e-3) / 1e6

    cpu = port = None
    if not args.no_cpu:
        log("CPU baseline: the reference CG for %d iterations" % args.cpu_ref_iters)
        cpu = cpu_reference(case, int(it_mean), args.cpu_ref_iters)
        log("CPU baseline: the oracle CG for %d iterations" % args.cpu_iters)
        port = cpu_port(case, int(it_mean), args.cpu_iters)
        if cpu is None:
            cpu, port = port, None

    out = {
        "metric": "ms per pressure solve to 1e-8 rel. residual at 256³ (plain vs NN-precond MP)",
        "value": round(wall_ms, 3),
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "pressure Poisson %d^3, random obstacles seed %d, solve to 1e-8 relative residual" %
                               (args.n, args.seed),
                   "value_solver": "plain FP64 CG (single-reduction, A p recomputed, %s)" % (
                       "one kernel per iteration" if fused else "two kernels per iteration"), "iterations": it_mean,
                   "fluid_cells": fluid, "mcell_updates_per_s": round(mcups, 1),
                   "solvers": solvers},
        "roofline": {"bound": "hbm", "achieved": round(kd_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kd_gbs / HBM_PEAK_GBS, 4),
                     "traffic": KF_TRAFFIC_BYTES if (args.n == 256 and fused) else None,
                     "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE, scaled by "
                                     "a 512 MiB copy (profiles/cg_kernels_traffic_pmc.md); 688 MB algorithmic, "
                                     "the excess is the halo rows and planes read again (mean of an even and "
                                     "an odd update)",
                     "kernel": ("k_cgcg_fused<double,double,4,10,1>" if fused else "k_cgcg_step<double,double,2>") +
                               " (41 B per interior cell, mean of 33 and 49)",
                     "kernel_us": round(kd["ms"] * 1e3, 2), "launches": kd["count"],
                     "iteration_bytes_per_cell": it_bytes / interior,
                     "iteration_us": round(it_us, 2),
                     "iteration_gbs": round(it_bytes / (it_us * 1e-6) / 1e9, 1),
                     "stream_ceiling_gbs": ceiling,
                     "stream_ceiling_note": "the fastest of a grid-stride loop and flat launches (plain and "
                                            "non-temporal) of 16-byte streams over 1 GiB buffers on the same GPU "
                                            "in the same run: read-only reduction, copy, and two reads + two "
                                            "writes in place (the iteration kernel's pattern)",
                     "kernel_frac_of_read2_write2_ceiling": round(kd_gbs / ceiling["read2_write2"], 3),
                     # the textbook Hestenes-Stiefel iteration (three fused kernels, 89 B per cell in
                     # FP64: A p, two dots, three axpys) would take this long at 100 % of peak
                     "textbook_hs_cg_us_at_peak": round(89.0 * interior / (HBM_PEAK_GBS * 1e9) * 1e6, 1)},
        "cpu_baseline": cpu,
    }
    if port is not None:
        out["cpu_oracle_port"] = port
    out["host_boundary"] = {"pcie_round_trip_ms": round(pcie_ms, 3),
                            "plain_fp64_cg_ms_with_pcie": round(wall_ms + pcie_ms, 3)}
    nn = solvers.get("nn_preconditioned_mp")
    if nn is not None:
        out["nn_precond_mp"] = {"ms_per_solve": nn["ms"], "speedup_vs_plain_fp64": nn["speedup_vs_plain_fp64"],
                                "solver": "FP32 flexible CG in FP64 refinement, V-cycle preconditioner with the "
                                          "trained network as its coarsest-grid (64^3) solve",
                                "true_residual": nn["true_residual"],
                                "ms_with_pcie": round(nn["ms"] + pcie_ms, 3)}
    if io is not None:
        out["output_files"] = io
    print(json.dumps(out), flush=True)
    g.close()
    ctx.close()


if __name__ == "__main__":
    main()
