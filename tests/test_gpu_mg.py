"""The V-cycle preconditioner of XF_FCG_MGNN on the GPU against its numpy
restatement (oracle/mg_ref.py) and the numpy coarsening of point types.

Bars: coarse-grid masks bit-exact; one cycle without the network within
1e-4 relative L2 of the FP32 restatement (same operations, different
grouping of the FP32 sums); solves reach the requested true residual.
"""
import numpy as np
import pytest
import torch

import mg_ref
import xflow_amd as xf
from xflow_amd import cases
from xflow_amd.unet import UNet

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = xf.Context(0)
    yield c
    c.close()


def case_of(shape, seed, dirichlet):
    codes = cases.neumann_codes(None, cases.random_obstacles(seed), shape=shape).astype(np.int32)
    if dirichlet:
        codes[0, 1:-1, 1:-1] = 3
    h = [cases.axis_steps(s) for s in shape]
    return codes, h


@pytest.mark.parametrize("shape,dirichlet", [((8, 8, 256), False), ((16, 12, 128), True)])
def test_hierarchy_masks_match_numpy(ctx, shape, dirichlet):
    codes, h = case_of(shape, 11, dirichlet)
    g = xf.Grid(ctx, h[0], h[1], h[2], codes)
    nlev = g.mg_levels()
    want = codes
    for k in range(1, nlev):
        want = cases.coarsen_codes(want)
        sm, _, _ = g.mg_level(k).export(activity=False)
        assert np.array_equal(sm, (want == 1).astype(np.int32)), k
    # halved to the first grid of at most 512 cells (ring included) or odd extents:
    # (8, 8, 256) -> (2, 2, 64); (16, 12, 128) -> (8, 6, 64) -> (4, 3, 32)
    assert nlev == 3 == len(mg_ref.hierarchy(codes, tuple(1.0 / s for s in shape)))


# (8, 10, 300): the residual restricted in the sweep pass with a shifted last
# k-tile and a block whose last two rows lie past the grid; (16, 16, 64) and
# (16, 16, 32): halved down to a grid of at most 512 cells, solved with its
# pseudo-inverse (pure Neumann: singular; with a boundary-value plane: regular),
# the levels of at most 2048 cells in one workgroup (k_mg_bottom: from level 2,
# level 1, and (8, 8, 16) from level 0)
@pytest.mark.parametrize("shape,dirichlet", [((8, 8, 256), False), ((16, 12, 128), True), ((20, 16, 64), True),
                                             ((8, 10, 300), True), ((16, 16, 64), False), ((16, 16, 32), True),
                                             ((8, 8, 16), True)])
def test_cycle_matches_restatement(ctx, shape, dirichlet):
    codes, h = case_of(shape, 12, dirichlet)
    g = xf.Grid(ctx, h[0], h[1], h[2], codes)
    rng = np.random.default_rng(7)
    r = np.where(codes == 1, rng.standard_normal(codes.shape), 0).astype(np.float32)
    z = torch.zeros(codes.shape, dtype=torch.float32, device="cuda")
    g.mg_apply(None, torch.from_numpy(r).cuda(), z)
    torch.cuda.synchronize()
    want, nlev = mg_ref.apply(codes, tuple(1.0 / s for s in shape), r)
    assert nlev == g.mg_levels()
    got = z.cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-4 * np.linalg.norm(want)


def test_flexible_cg_with_cycle_converges(ctx):
    case = cases.obstacle_case(128, seed=4)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = torch.from_numpy(case["b"]).cuda()
    x = torch.zeros_like(b)
    st0 = g.solve(b, torch.zeros_like(b), x, xf.solve_opts(xf.XF_CG, tol=1e-8, tol_mode=xf.XF_TOL_REL,
                                                             project_mean=1))
    for net in (None, UNet(ctx, seed=9)):
        o = xf.solve_opts(xf.XF_FCG_MGNN, tol=1e-8, tol_mode=xf.XF_TOL_REL, project_mean=1, inner_tol=1e-4)
        o.precond = net.h if net is not None else None
        x.zero_()
        st = g.solve(b, torch.zeros_like(b), x, o)
        assert st.status == 0, st.status
        assert st.resid_true <= 1.5e-8
        if net is None:
            # the classical cycle alone is a strong preconditioner
            assert st.inner_iters * 3 < st0.iters, (st.inner_iters, st0.iters)


def test_trained_network_gain(ctx):
    """What the trained network does against the method without it, on a
    held-out problem (training used seeds 100+, this is seed 7): the flexible
    CG with the full-depth classical V-cycle (no network: down to 4^3 and an
    exact pseudo-inverse solve there), with the cycle cut at 64^3 and an
    untrained network there, and with the trained one (xflow_amd/weights),
    and plain CG for reference.  Every solve must reach the true residual.
    The trained network is far better than an untrained one; against the
    full-depth classical cycle it needs more iterations (the times are
    printed, not asserted)."""
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "xflow_amd", "weights",
                        "unet_coarse64.npy")
    trained = UNet(ctx, seed=0)
    trained.set_params(np.load(path))
    untrained = UNet(ctx, seed=1)
    case = cases.obstacle_case(128, seed=7)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = torch.from_numpy(case["b"]).cuda()
    x = torch.zeros_like(b)
    plain = g.solve(b, torch.zeros_like(b), x, xf.solve_opts(xf.XF_CG, tol=1e-8, tol_mode=xf.XF_TOL_REL,
                                                              project_mean=1))
    res = {}
    for name, net in (("no network", None), ("untrained", untrained), ("trained", trained)):
        o = xf.solve_opts(xf.XF_FCG_MGNN, tol=1e-8, tol_mode=xf.XF_TOL_REL, project_mean=1, inner_tol=1e-4)
        o.precond = net.h if net is not None else None
        x.zero_()
        st = g.solve(b, torch.zeros_like(b), x, o)
        assert st.status == 0, (name, st.status)
        assert st.resid_true <= 1.5e-8, (name, st.resid_true)
        res[name] = st
    print("plain CG %d iterations %.1f ms; cycle inner iterations / ms: %s" % (
        plain.iters, plain.ms, {k: (v.inner_iters, round(v.ms, 1)) for k, v in res.items()}))
    t = res["trained"].inner_iters
    assert t * 3 < res["untrained"].inner_iters
    assert t * 15 < plain.iters
    assert res["no network"].inner_iters * 15 < plain.iters


# (16, 14, 200): a block whose last two rows lie past the grid
@pytest.mark.parametrize("shape", [(16, 16, 300), (16, 12, 200), (16, 14, 200)])
def test_flexible_cg_long_rows(ctx, shape):
    """Rows of 300 cells take the flexible-CG step's 4-floats-per-lane kernels
    with the last k-tile shifted back over the one before it, rows of 200 the
    2-floats-per-lane ones; the solve matches the FP64 CG's field."""
    length = tuple(s / shape[2] for s in shape)  # equal steps on all axes (the cycle needs an isotropic operator)
    codes = cases.neumann_codes(None, [cases.box_obstacle((0.5 * length[0], 0.5 * length[1], 0.4), 0.01)],
                                shape=shape, length=length)
    hc = 1.0 / shape[2]
    b = cases.divergence_rhs(codes, seed=5, h=[hc, hc, hc])
    h = [cases.axis_steps(s, s * hc) for s in shape]
    g = xf.Grid(ctx, h[0], h[1], h[2], codes)
    bd = torch.from_numpy(b).cuda()
    xr = torch.zeros_like(bd)
    st0 = g.solve(bd, torch.zeros_like(bd), xr, xf.solve_opts(xf.XF_CG, tol=1e-11, tol_mode=xf.XF_TOL_REL,
                                                               project_mean=1))
    assert st0.status == 0
    o = xf.solve_opts(xf.XF_FCG_MGNN, tol=1e-9, tol_mode=xf.XF_TOL_REL, project_mean=1, inner_tol=1e-4)
    o.precond = None
    x = torch.zeros_like(bd)
    st = g.solve(bd, torch.zeros_like(bd), x, o)
    assert st.status == 0 and st.fallback == 0, (st.status, st.fallback)
    assert st.resid_true <= 2e-9
    m = torch.from_numpy(codes == 1).cuda()
    a, r = x[m] - x[m].mean(), xr[m] - xr[m].mean()
    assert float(torch.linalg.norm(a - r) / torch.linalg.norm(r)) <= 1e-6
    g.close()


def test_nonzero_mean_rhs_all_methods(ctx):
    """A Neumann right-hand side whose mean over the unknowns is not zero:
    with project_mean every method stops on the norm of the mean-free
    residual (the plain CG's rule), so the mixed-precision and flexible-CG
    solves converge like the plain CG instead of refining until their cap,
    and all three fields agree up to a constant."""
    case = cases.obstacle_case(64, seed=2)
    fluid = case["codes"] == 1
    bh = case["b"].copy()
    bh[fluid] += 0.05 * np.abs(bh).max()
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = torch.from_numpy(bh).cuda()
    m = torch.from_numpy(fluid).cuda()
    fields = {}
    for method in (xf.XF_CG, xf.XF_CG_MP, xf.XF_FCG_MGNN):
        o = xf.solve_opts(method, tol=1e-10, tol_mode=xf.XF_TOL_REL, project_mean=1, inner_tol=1e-4, maxit=20000)
        o.precond = None
        x = torch.zeros_like(b)
        st = g.solve(b, torch.zeros_like(b), x, o)
        print("method %d: status %d, outer %d, inner %d, resid %.2e" % (method, st.status, st.iters,
                                                                      st.inner_iters, st.resid))
        assert st.status == 0, (method, st.status)
        assert st.resid <= 1e-10
        assert method == xf.XF_CG or st.iters <= 8, (method, st.iters)
        fields[method] = x[m] - x[m].mean()
    ref = fields[xf.XF_CG]
    for method in (xf.XF_CG_MP, xf.XF_FCG_MGNN):
        assert float(torch.linalg.norm(fields[method] - ref) / torch.linalg.norm(ref)) <= 1e-6, method
    g.close()
