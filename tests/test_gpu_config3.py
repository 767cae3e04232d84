"""The 256^3 random-obstacle pressure problem (seed 0: 16,444,373 fluid cells,
Neumann walls) solved on the GPU by every throughput solver to a 1e-10
relative residual, against the oracle's FP64 CG solution to 1e-12
(tests/golden/make_config3.py: 2781 iterations), stored as the mean-free
field's 2-norm over the fluid and its values at 512 fixed fluid cells.

Solvers: the plain FP64 single-reduction CG, FP32 CG inside FP64 iterative
refinement, and the flexible CG in mixed precision preconditioned by the
full-depth classical V-cycle and by the cycle with the trained coarse-grid
network (xflow_amd/weights/unet_coarse64.npy).  Bars: the solve's own true
residual (FP64, after the solve) at most 1.5e-10; the norm and the samples
within 1e-6 relative of the oracle's.
"""
import os

import numpy as np
import pytest
import torch

import xflow_amd as xf
from xflow_amd import cases

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def problem():
    ctx = xf.Context(0)
    case = cases.obstacle_case(256, seed=0)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    fx = np.load(os.path.join(HERE, "golden", "config3_obstacle256.npz"))
    fluid = torch.from_numpy(case["codes"].ravel() == 1).cuda()
    yield ctx, g, torch.from_numpy(case["b"]).cuda(), fluid, fx
    g.close()
    ctx.close()


@pytest.mark.parametrize("solver", ["plain FP64 CG", "mixed precision", "cycle", "cycle + network"])
def test_obstacle256_matches_oracle(problem, solver):
    ctx, g, b, fluid, fx = problem
    from xflow_amd.unet import UNet
    net = None
    if solver == "plain FP64 CG":
        method = xf.XF_CG
    elif solver == "mixed precision":
        method = xf.XF_CG_MP
    else:
        method = xf.XF_FCG_MGNN
        if solver == "cycle + network":
            net = UNet(ctx, seed=0)
            net.set_params(np.load(os.path.join(os.path.dirname(HERE), "xflow_amd", "weights", "unet_coarse64.npy")))
    o = xf.solve_opts(method, tol=1e-10, tol_mode=xf.XF_TOL_REL, project_mean=1, verify=1, inner_tol=1e-4)
    o.precond = net.h if net is not None else None
    x = torch.zeros_like(b)
    st = g.solve(b, torch.zeros_like(b), x, o)
    assert st.status == 0 and st.fallback == 0, (st.status, st.fallback)
    xf_ = x.reshape(-1)[fluid]
    xf_ = xf_ - xf_.mean()
    norm = float(torch.linalg.norm(xf_))
    full = x.reshape(-1) - x.reshape(-1)[fluid].mean()
    samples = full[torch.from_numpy(fx["idx"]).cuda()].cpu().numpy()
    dn = abs(norm - fx["norm"][0]) / fx["norm"][0]
    dx = np.linalg.norm(samples - fx["x"]) / np.linalg.norm(fx["x"])
    print("%s: %d outer / %d inner, true resid %.2e, %.1f ms; norm %.2e, samples %.2e" % (
        solver, st.iters, st.inner_iters, st.resid_true, st.ms, dn, dx))
    assert 0 < st.resid_true <= 1.5e-10, st.resid_true
    assert dn <= 1e-6 and dx <= 1e-6, (dn, dx)
