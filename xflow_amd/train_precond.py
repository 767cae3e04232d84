"""Training of the learned preconditioner on residual/error pairs harvested
from the solver itself.

Harvest: for each Poisson problem (random solid spheres/boxes, RHS from the
smoothed-random-velocity generator, seeds 100+), the FP64 GPU CG solves to a
1e-12 relative residual for x*, then reruns with iteration caps k to collect
the CG iterates x_k; each pair is (r_k = b - A x_k, e_k = x* - x_k) with the
mean over the fluid removed from e_k, both divided by s = rms(r_k) over the
fluid (the scaling the preconditioner applies).  Training draws random crops
of the pairs and minimises the masked mean-squared error with Adam.

    python -m xflow_amd.train_precond --out weights.npy [--problems 40 --steps 1500]

With --pairs N the harvest keeps about N pairs from ncell^3 problems, each as
one random crop^3 window in BF16 (the mask as bytes), with the problems
generated by --workers CPU processes while the GPU solves:

    python -m xflow_amd.train_precond --out unet_poisson.npy --ncell 128 --pairs 20000 --steps 10000
"""
import argparse
import multiprocessing
import os
import sys
import time

import numpy as np
import torch

import xflow_amd as xf
from xflow_amd import cases
from xflow_amd.unet import LOSS_COSINE, LOSS_ENERGY, LOSS_MSE, UNet

ITERATES = (0, 2, 5, 10, 20, 40, 80, 160, 320)
# the iteration caps of the cropped harvest (18 pairs per problem)
ITERATES_DENSE = (0, 1, 2, 3, 5, 7, 10, 14, 20, 28, 40, 56, 80, 113, 160, 226, 320, 452)


def coarse_case(ncell, fine, seed):
    """A problem on the grid the multigrid cycle reaches from a fine^3 problem:
    the fine random-obstacle mask coarsened (cases.coarsen_codes) down to
    ncell^3, the right-hand side generated on that grid."""
    codes = cases.neumann_codes(fine, cases.random_obstacles(seed))
    while codes.shape[0] - 2 > ncell:
        codes = cases.coarsen_codes(codes)
    b = cases.divergence_rhs(codes, seed)
    h = cases.axis_steps(ncell)
    return {"codes": codes, "hx": h, "hy": h.copy(), "hz": h.copy(), "b": b}


def box_case(fine_shape, seed):
    """The coarsest grid of a uniform-step box (fine_shape cells, e.g. the IB
    channel's 256x128x128 pressure grid): all fluid (the immersed body is not
    in the pressure mask) for even seeds, random obstacles for odd ones; the
    right-hand side generated on the coarsest grid."""
    obs = cases.random_obstacles(seed) if seed % 2 else []
    length = tuple(s / fine_shape[1] for s in fine_shape)
    codes = cases.neumann_codes(None, obs, shape=fine_shape, length=length)
    while (codes.shape[2] - 2) // 2 >= 64 and all((s - 2) % 2 == 0 for s in codes.shape):
        codes = cases.coarsen_codes(codes)
    shape = tuple(s - 2 for s in codes.shape)
    hc = length[0] / shape[0]
    h = [np.full(n + 1, hc) for n in shape]
    b = cases.divergence_rhs(codes, seed, h=[hc, hc, hc])
    return {"codes": codes, "hx": h[0], "hy": h[1], "hz": h[2], "b": b}


def harvest(ctx, ncell, problems, seed0=100, coarsen_from=0, box=None, log=print):
    """Returns device tensors R, E, M of shape [P*len(ITERATES), D, H, W] (interior only)."""
    Rs, Es, Ms = [], [], []
    for pidx in range(problems):
        seed = seed0 + pidx
        if box:
            case = box_case(box, seed)
        elif coarsen_from:
            case = coarse_case(ncell, coarsen_from, seed)
        else:
            case = cases.poisson_case(ncell, cases.random_obstacles(seed), seed)
        g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
        b = torch.from_numpy(case["b"]).cuda()
        x0 = torch.zeros_like(b)
        xs = torch.zeros_like(b)
        st = g.solve(b, x0, xs, xf.solve_opts(xf.XF_CG, tol=1e-12, tol_mode=xf.XF_TOL_REL, project_mean=1))
        fluid = torch.from_numpy((case["codes"] == 1).astype(np.float32)).cuda()
        nf = fluid.sum()
        for k in ITERATES:
            xk = torch.zeros_like(b)
            if k > 0:
                g.solve(b, x0, xk, xf.solve_opts(xf.XF_CG, tol=1e-30, tol_mode=xf.XF_TOL_REL, project_mean=1,
                                                 maxit=k))
            rk = torch.zeros_like(b)
            g.apply(xk, rk, b, -1.0)  # A x - b
            rk = -rk * fluid.double()
            ek = (xs - xk) * fluid.double()
            ek = (ek - (ek.sum() / nf) * fluid.double())
            s = torch.sqrt((rk * rk).sum() / nf)
            if s <= 0:
                continue
            sl = (slice(1, -1),) * 3
            Rs.append((rk / s).float()[sl].contiguous())
            Es.append((ek / s).float()[sl].contiguous())
            Ms.append(fluid[sl].contiguous())
        g.close()
        log("harvest problem %d: CG %d iterations" % (pidx, st.iters))
    return torch.stack(Rs), torch.stack(Es), torch.stack(Ms)


def _obstacle_problem(arg):
    """CPU worker: the random-obstacle problem of one seed."""
    ncell, seed = arg
    return seed, cases.poisson_case(ncell, cases.random_obstacles(seed), seed)


def harvest_crops(ctx, ncell, pairs, crop, pool, seed0=100, log=print):
    """About `pairs` pairs (r_k, e_k) from ncell^3 random-obstacle problems
    (seeds seed0, seed0 + 1, ...; len(ITERATES_DENSE) CG iterates each), one
    random crop^3 window of each kept: R, E in BF16, M as bytes, on the
    device.  `pool` generates the problems on the CPU while the GPU solves."""
    per = len(ITERATES_DENSE)
    problems = (pairs + per - 1) // per
    cap = problems * per
    R = torch.empty((cap, crop, crop, crop), dtype=torch.bfloat16, device="cuda")
    E = torch.empty_like(R)
    M = torch.empty((cap, crop, crop, crop), dtype=torch.uint8, device="cuda")
    gen = torch.Generator().manual_seed(seed0)
    n, t0 = 0, time.time()
    jobs = [(ncell, seed0 + q) for q in range(problems)]
    for pidx, (seed, case) in enumerate(pool.imap(_obstacle_problem, jobs)):
        g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
        b = torch.from_numpy(case["b"]).cuda()
        x0 = torch.zeros_like(b)
        xs = torch.zeros_like(b)
        st = g.solve(b, x0, xs, xf.solve_opts(xf.XF_CG, tol=1e-12, tol_mode=xf.XF_TOL_REL, project_mean=1))
        if st.status != 0:  # no reference solution: no pairs from this problem
            g.close()
            log("harvest problem %d/%d (seed %d): CG status %d after %d iterations, skipped" %
                (pidx + 1, problems, seed, st.status, st.iters))
            continue
        fluid = torch.from_numpy((case["codes"] == 1).astype(np.float64)).cuda()
        nf = fluid.sum()
        for k in ITERATES_DENSE:
            xk = torch.zeros_like(b)
            if k > 0:
                g.solve(b, x0, xk, xf.solve_opts(xf.XF_CG, tol=1e-30, tol_mode=xf.XF_TOL_REL, project_mean=1,
                                                 maxit=k))
            rk = torch.zeros_like(b)
            g.apply(xk, rk, b, -1.0)  # A x - b
            rk = -rk * fluid
            ek = (xs - xk) * fluid
            ek = ek - (ek.sum() / nf) * fluid
            sc = torch.sqrt((rk * rk).sum() / nf)
            if sc <= 0:
                continue
            o = [1 + int(torch.randint(0, ncell - crop + 1, (1,), generator=gen)) for _ in range(3)]
            sl = tuple(slice(v, v + crop) for v in o)
            R[n] = (rk[sl] / sc).to(torch.bfloat16)
            E[n] = (ek[sl] / sc).to(torch.bfloat16)
            M[n] = fluid[sl].to(torch.uint8)
            n += 1
        g.close()
        if pidx % 25 == 0 or pidx == problems - 1:
            log("harvest problem %d/%d (seed %d): CG %d iterations, %d pairs, %.1f s" %
                (pidx + 1, problems, seed, st.iters, n, time.time() - t0))
    return R[:n], E[:n], M[:n]


def crops(R, E, M, B, S, gen):
    n, D, H, W = R.shape
    if S is None or S <= 0 or S >= D:  # whole fields (or whole stored windows)
        idx = torch.randint(0, n, (B,), generator=gen).to(R.device)
        return R[idx].float().contiguous(), M[idx].float().contiguous(), E[idx].float().contiguous()
    idx = torch.randint(0, n, (B,), generator=gen)
    zs = torch.randint(0, D - S + 1, (B,), generator=gen)
    ys = torch.randint(0, H - S + 1, (B,), generator=gen)
    xs = torch.randint(0, W - S + 1, (B,), generator=gen)
    rr, ee, mm = [], [], []
    for q in range(B):
        sl = (idx[q], slice(zs[q], zs[q] + S), slice(ys[q], ys[q] + S), slice(xs[q], xs[q] + S))
        rr.append(R[sl])
        ee.append(E[sl])
        mm.append(M[sl])
    return torch.stack(rr).contiguous(), torch.stack(mm).contiguous(), torch.stack(ee).contiguous()


def train(ctx, R, E, M, steps=1500, batch=4, crop=64, lr=1e-3, seed=0, loss="cosine", log=print):
    net = UNet(ctx, seed=seed)
    net.set_loss({"energy": LOSS_ENERGY, "cosine": LOSS_COSINE, "mse": LOSS_MSE}[loss])
    gen = torch.Generator().manual_seed(seed)
    t0 = time.time()
    hist = []
    for it in range(steps):
        lr_t = lr * (0.1 ** (it / max(steps, 1)))  # exponential decay to lr/10
        r, m, e = crops(R, E, M, batch, crop, gen)
        loss = net.train_step(r, m, e, lr_t)
        hist.append(loss)
        if it % 100 == 0 or it == steps - 1:
            log("step %d loss %.4e (%.1f s)" % (it, loss, time.time() - t0))
    return net, hist


def evaluate_mg(ctx, net, ncell, seed, log=print):
    """Plain CG vs the V-cycle-preconditioned flexible CG without and with the
    network on one held-out problem (iterations to 1e-8)."""
    case = cases.poisson_case(ncell, cases.random_obstacles(seed), seed)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = torch.from_numpy(case["b"]).cuda()
    x = torch.zeros_like(b)
    st0 = g.solve(b, torch.zeros_like(b), x, xf.solve_opts(xf.XF_CG, tol=1e-8, tol_mode=xf.XF_TOL_REL, project_mean=1))
    res = {"plain": st0.iters}
    for name, p in (("cycle", None), ("cycle+net", net)):
        o = xf.solve_opts(xf.XF_FCG_MGNN, tol=1e-8, tol_mode=xf.XF_TOL_REL, project_mean=1, inner_tol=1e-4)
        o.precond = p.h if p is not None else None
        x.zero_()
        st = g.solve(b, torch.zeros_like(b), x, o)
        res[name] = (st.inner_iters, st.fallback, st.ms, st.resid_true)
    g.close()
    log("evaluate %d^3 seed %d: plain CG %d; cycle without network %s; cycle with network %s "
        "(inner iterations, fallback, ms, true residual)" % (ncell, seed, res["plain"], res["cycle"], res["cycle+net"]))
    return res


def evaluate(ctx, net, ncell, seed, log=print):
    """Plain FP32-refined CG vs the network-preconditioned flexible CG on one
    held-out problem: (plain iterations, network iterations, fallback)."""
    case = cases.poisson_case(ncell, cases.random_obstacles(seed), seed)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = torch.from_numpy(case["b"]).cuda()
    x = torch.zeros_like(b)
    st0 = g.solve(b, torch.zeros_like(b), x, xf.solve_opts(xf.XF_CG, tol=1e-8, tol_mode=xf.XF_TOL_REL, project_mean=1))
    o = xf.solve_opts(xf.XF_FCG_NN, tol=1e-8, tol_mode=xf.XF_TOL_REL, project_mean=1, inner_tol=1e-4)
    o.precond = net.h
    x.zero_()
    st1 = g.solve(b, torch.zeros_like(b), x, o)
    g.close()
    log("evaluate %d^3 seed %d: plain CG %d iterations; network FCG %d outer / %d inner, fallback %d, "
        "true residual %.2e" % (ncell, seed, st0.iters, st1.iters, st1.inner_iters, st1.fallback, st1.resid_true))
    return st0.iters, st1.inner_iters, st1.fallback


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--ncell", type=int, default=128)
    ap.add_argument("--problems", type=int, default=24)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--crop", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--loss", choices=("energy", "cosine", "mse"), default="energy")
    ap.add_argument("--eval-ncell", type=int, default=128)
    ap.add_argument("--coarsen-from", type=int, default=0,
                    help="train the coarsest-grid network of the V-cycle: fine^3 masks coarsened to ncell^3")
    ap.add_argument("--pairs", type=int, default=0,
                    help="cropped harvest of about this many pairs from ncell^3 problems (see the module notes)")
    ap.add_argument("--workers", type=int, default=16, help="CPU processes generating problems (--pairs)")
    ap.add_argument("--box", default="",
                    help="train on the coarsest grid of a DxHxW-cell box (e.g. 256x128x128), whole fields")
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    # the problem generators fork before this process touches the GPU
    pool = multiprocessing.get_context("fork").Pool(args.workers) if args.pairs else None
    ctx = xf.Context(0)
    t0 = time.time()
    box = tuple(int(v) for v in args.box.split("x")) if args.box else None
    if args.pairs:
        R, E, M = harvest_crops(ctx, args.ncell, args.pairs, args.crop, pool, log=log)
        pool.close()
        pool.join()
    else:
        R, E, M = harvest(ctx, args.ncell, args.problems, coarsen_from=args.coarsen_from, box=box, log=log)
    if box:
        args.crop = 0
    log("harvested %d pairs in %.1f s" % (R.shape[0], time.time() - t0))
    net, hist = train(ctx, R, E, M, args.steps, args.batch, args.crop, args.lr, loss=args.loss, log=log)
    np.save(args.out, net.get_params())
    log("saved %s; loss %.4e -> %.4e" % (args.out, hist[0], np.mean(hist[-50:])))
    if box:
        pass
    elif args.eval_ncell and args.coarsen_from:
        evaluate_mg(ctx, net, args.eval_ncell, 999, log=log)
    elif args.eval_ncell:
        evaluate(ctx, net, args.eval_ncell, 999, log=log)


if __name__ == "__main__":
    main()
