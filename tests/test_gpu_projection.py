"""GPU projection glue (pressure right-hand side, velocity correction,
divergence monitor) and the mixed-precision solver, against the oracle.

The glue is +-*/ in the reference's order: bit-exact.  The mixed-precision
solve is checked against the FP64 oracle solution: the pressure field within
1e-6 relative L2 (mean-free, pure-Neumann system) at a 1e-10 relative
residual, which the true FP64 residual recomputed by the library confirms.
"""
import numpy as np
import pytest
import torch

import xflow_amd as xf
import xf_oracle as xo
from xflow_amd import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = xf.Context(0)
    yield c
    c.close()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def mixed_codes(nc):
    codes = cases.neumann_codes(nc, [cases.sphere_obstacle((0.5, 0.5, 0.5), 0.2)]).astype(np.int32)
    return codes


def test_pressure_rhs_and_divergence_bit_exact(ctx):
    rng = np.random.default_rng(5)
    nc = 30
    codes = mixed_codes(nc)
    codes[0, 5:9, 5:9] = 4  # PreDefined pressure points take their boundary values
    n = nc + 2
    hs = [rng.uniform(0.02, 0.04, n - 1) for _ in range(3)]
    u, v, w, rho, pbc = (rng.standard_normal(codes.shape) for _ in range(5))
    rho = 1 + 0.5 * np.abs(rho)
    g = xf.Grid(ctx, hs[0], hs[1], hs[2], codes)
    rhs = dev(np.zeros(codes.shape))
    g.pressure_rhs(dev(u), dev(v), dev(w), hs[0], hs[1], hs[2], 0.01, rhs, rho=dev(rho), pbc=dev(pbc))
    want = xo.pressure_rhs(codes, u, v, w, hs[0], hs[1], hs[2], 0.01, rho=rho, pbc=pbc)
    assert np.array_equal(host(rhs), want)
    val, cell = g.divergence_max(dev(u), dev(v), dev(w), hs[0], hs[1], hs[2])
    wv, wc = xo.divergence_max(codes, u, v, w, hs[0], hs[1], hs[2])
    assert val == wv and cell == wc


def test_velocity_correction_bit_exact(ctx):
    rng = np.random.default_rng(6)
    nc = 24
    n = nc + 2
    codes = np.full((n, n, n), 1, dtype=np.int32)
    codes[:, :, 0] = codes[:, :, -1] = 0
    codes[:, 0, :] = codes[:, -1, :] = 0
    codes[-1] = 0
    codes[0] = 3     # Defined inflow plane
    codes[-2] = 5    # Normal outflow plane (its gradient reads the plane beyond)
    codes[:, 0, :] = codes[:, -1, :] = 0
    codes[:, :, 0] = codes[:, :, -1] = 0
    codes[10:14, 10:14, 10:14] = 0
    hP = rng.uniform(0.03, 0.05, n - 1)
    us, p, rho, bc = (rng.standard_normal(codes.shape) for _ in range(4))
    rho = 1 + np.abs(rho)
    us[codes == 3] = bc[codes == 3]  # the reference checks Defined values
    g = xf.Grid(ctx, np.full(n - 1, 0.04), np.full(n - 1, 0.04), np.full(n - 1, 0.04), codes)
    for axis in (1, 2, 3):
        out = dev(np.zeros(codes.shape))
        g.velocity_correct(axis, dev(us), dev(p), hP, 0.02, out, rho=dev(rho), bc=dev(bc))
        want = xo.velocity_correct(codes, axis, us, p, hP, 0.02, rho=rho, bc=bc)
        got = host(out)
        assert np.array_equal(got, want)
    bad = us.copy()
    bad[0, 5, 5] += 1.0  # a Defined value that disagrees with its boundary condition
    with pytest.raises(xf.XfError):
        g.velocity_correct(1, dev(bad), dev(p), hP, 0.02, dev(np.zeros(codes.shape)), bc=dev(bc))


@pytest.mark.parametrize("ncell", [32, 64])
def test_mixed_precision_solver(ctx, ncell):
    case = cases.obstacle_case(ncell, seed=1)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = dev(case["b"])
    x = torch.zeros_like(b)
    st = g.solve(b, torch.zeros_like(b), x, xf.solve_opts(xf.XF_CG_MP, tol=1e-10, tol_mode=xf.XF_TOL_REL,
                                                            project_mean=1, inner_tol=1e-4))
    assert st.status == 0, st.status
    assert st.resid_true <= 2e-10
    og = xo.Grid(case["hx"], case["hy"], case["hz"], case["codes"])
    xr, it, _, s2 = xo.cg_cg(og, case["b"], np.zeros_like(case["b"]), 1e-12, 100000, project_mean=True)
    m = case["codes"] == 1
    xg = host(x)
    a, r = xg[m] - xg[m].mean(), xr[m] - xr[m].mean()
    assert np.linalg.norm(a - r) / np.linalg.norm(r) <= 1e-6
    assert st.iters >= 2 and st.inner_iters > it * 0.8


@pytest.mark.parametrize("shape", [(14, 12, 300), (12, 10, 257)])
def test_mixed_precision_long_rows(ctx, shape):
    """Rows of at least 256 cells take the 4-floats-per-lane stencil (the last
    k-tile of a row is shifted back and overlaps the one before it)."""
    codes = cases.neumann_codes(None, [cases.box_obstacle((0.5, 0.5, 0.4), 0.2)], shape=shape)
    bb = cases.divergence_rhs(codes, seed=2, sigma=1.5)
    h = [cases.axis_steps(s_) for s_ in shape]
    g = xf.Grid(ctx, h[0], h[1], h[2], codes)
    b = dev(bb)
    x = torch.zeros_like(b)
    st = g.solve(b, torch.zeros_like(b), x, xf.solve_opts(xf.XF_CG_MP, tol=1e-10, tol_mode=xf.XF_TOL_REL,
                                                            project_mean=1, inner_tol=1e-4))
    assert st.status == 0, st.status
    assert st.resid_true <= 2e-10
    og = xo.Grid(h[0], h[1], h[2], codes)
    xr, it, _, s2 = xo.cg_cg(og, bb, np.zeros_like(bb), 1e-12, 100000, project_mean=True)
    m = codes == 1
    xg = host(x)
    a, r = xg[m] - xg[m].mean(), xr[m] - xr[m].mean()
    assert np.linalg.norm(a - r) / np.linalg.norm(r) <= 1e-6


@pytest.mark.parametrize("axis", [0, 1, 2, 3])
def test_transport_terms_bit_exact(ctx, axis):
    """xf_transport_velocity against the oracle's restatement of the
    reference's inline construction (XProblem.cpp:5090-5303, 5357-5417 for
    the velocity components, 5889-5944 for the concentration): transport
    velocities, viscosity mean and right-hand side bit for bit, on random
    fields, non-uniform steps and every point type the reference accepts for
    the equation (Normal points with an x normal for the first component)."""
    from xflow_amd import drivers
    rng = np.random.default_rng(40 + axis)
    normals = None
    if axis == 0:
        codes = drivers.channel_scalar_codes(12, 10, 8)
    else:
        codes, _ = drivers.velocity_codes(12, axis)
        if axis == 1:
            codes[12, 4:7, 3:9] = 5  # Normal points on the outflow wall of u (normal along x)
            normals = np.zeros(codes.shape + (3,))
            normals[..., 0] = 1.0
    shape = codes.shape
    hs = [rng.uniform(0.05, 0.1, s - 1) for s in shape]
    u, v, w, vis, bc = (rng.standard_normal(shape) for _ in range(5))
    vis = 0.01 + np.abs(vis)
    g = xf.Grid(ctx, hs[0], hs[1], hs[2], codes, normals=normals)
    out = [dev(np.full(shape, 9.0)) for _ in range(5)]
    if axis == 0:
        g.transport_terms(0, [dev(u), dev(v), dev(w)], out[:3], bc=dev(bc), rp=out[4])
        want = xo.transport_velocity(0, codes, u, v, w, bc=bc, h=hs)
        names = (0, 1, 2, 4)
    else:
        g.transport_terms(axis, [dev(u), dev(v), dev(w)], out[:3], vis=dev(vis), bc=dev(bc), tvis=out[3], rp=out[4])
        want = xo.transport_velocity(axis, codes, u, v, w, vis=vis, bc=bc, h=hs, normals=normals)
        names = range(5)
    for q in names:
        got = host(out[q])
        assert np.array_equal(got, want[q]), (axis, q, np.abs(got - want[q]).max())
    # point types the reference rejects for this equation
    bad = codes.copy()
    bad[5, 5, 5] = 3 if axis == 0 else 7  # Defined concentration point / Equation velocity point
    gb = xf.Grid(ctx, hs[0], hs[1], hs[2], bad)
    with pytest.raises(xf.XfError):
        gb.transport_terms(axis, [dev(u), dev(v), dev(w)], out[:3])
    with pytest.raises(ValueError):
        xo.transport_velocity(axis, bad, u, v, w)
    g.close()
    gb.close()
