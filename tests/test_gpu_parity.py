"""GPU parity: libxflow_hip.so through its C ABI against the reference's golden
vectors (tests/golden, written by the compiled reference) and the CPU oracle.

Bars: integer/index arrays bit-exact; operator rows bit-exact (same arithmetic
order, no FMA contraction, correctly rounded division); reductions within
1e-13 relative (different summation order); solver fields within 1e-10
relative L2 of the reference and iteration counts within +-2.
"""
import numpy as np
import pytest
import torch

import xflow_amd as xf
import xf_oracle as xo
from conftest import golden

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


def grid_of(ctx, ins):
    return xf.Grid(ctx, np.diff(ins["x"]), np.diff(ins["y"]), np.diff(ins["z"]), ins["codes"],
                   ins.get("normals"), ins.get("has_normal"))


def rel_l2(a, b):
    d = np.linalg.norm(a - b)
    n = np.linalg.norm(b)
    return d / n if n > 0 else d


@pytest.mark.parametrize("name", ["op_mixed", "op_sphere16"])
def test_grid_export_bit_exact(ctx, name):
    ins, outs = golden(name)
    g = grid_of(ctx, ins)
    sm, rn, act = g.export()
    assert np.array_equal(sm, ins["space_mask"])
    assert np.array_equal(rn, outs["rn"])
    assert g.n_unknowns == outs["act"].shape[0]
    assert np.array_equal(act, outs["act"])


@pytest.mark.parametrize("g", ["U", "V", "W", "P"])
def test_reference_made_masks(ctx, g):
    """Grids from the reference's own cylinder-channel generator and
    LocalPrepareVariables (XStatement.cpp:7717-7977, XProblem.cpp:4116-4206):
    space mask and Rn numbering bit-exact on all four grids; on the pressure
    grid (with its normals) the activity table and the operator rows too."""
    ins, outs = golden("vp_cylinder")
    grid = xf.Grid(ctx, np.diff(outs[g + "_x"]), np.diff(outs[g + "_y"]), np.diff(outs[g + "_z"]),
                   outs[g + "_codes"], outs[g + "_normals"], outs[g + "_has_normal"])
    sm, rn, act = grid.export(activity=g == "P")
    assert np.array_equal(sm, (outs[g + "_rn"] > 0).astype(np.int32))
    assert np.array_equal(rn, outs[g + "_rn"])
    if g == "P":
        assert np.array_equal(act, outs["P_act"])
        y = dev(np.zeros(outs["P_codes"].shape))
        grid.apply(dev(ins["xvec"]), y)
        assert np.array_equal(host(y), outs["P_Ax"])
    grid.close()


@pytest.mark.parametrize("name,rho", [("op_mixed", False), ("op_sphere16", False), ("divop_mixed", True)])
def test_operator_rows_bit_exact(ctx, name, rho):
    ins, outs = golden(name)
    g = grid_of(ctx, ins)
    if rho:
        g.set_density(dev(ins["rho"]))
    x, b = dev(ins["xvec"]), dev(ins["bvec"])
    y = dev(ins["yinit"])
    g.apply(x, y)
    assert np.array_equal(host(y), outs["Ax"])
    y2 = dev(ins["yinit"])
    g.apply(x, y2, b, -1.0)
    assert np.array_equal(host(y2), outs["Ax_minus_b"])
    assert g.apply_dot(x) == pytest.approx(outs["dots"][0], rel=1e-13)
    assert g.dot(x, b) == pytest.approx(outs["dots"][1], rel=1e-13, abs=1e-13)
    if "absmax" in outs:
        v, idx = g.absmax(x)
        assert v == outs["absmax"][0] and idx == outs["absmax_rn"][0]


def test_vector_kernels_match_oracle(ctx):
    ins, _ = golden("op_mixed")
    g = grid_of(ctx, ins)
    og = xo.Grid(np.diff(ins["x"]), np.diff(ins["y"]), np.diff(ins["z"]), ins["codes"], ins["space_mask"],
                 ins["normals"], ins["has_normal"])
    rng = np.random.default_rng(11)
    s1, s2, r0 = (rng.standard_normal(g.shape) for _ in range(3))
    for p1, p2 in [(0.7, -1.3), (1.0, 0.0), (0.0, 2.5), (0.0, 0.0), (-1.0, 1.0)]:
        res = dev(r0)
        g.lincomb(p1, dev(s1), p2, dev(s2), res)
        want = r0.copy()
        xo.lib().xo_lincomb(og.ref, p1, s1.ctypes.data, p2, s2.ctypes.data, want.ctypes.data)
        assert np.array_equal(host(res), want)
    res = dev(r0)
    g.copy(dev(s1), res)
    want = r0.copy()
    want[ins["space_mask"] > 0] = s1[ins["space_mask"] > 0]
    assert np.array_equal(host(res), want)


def test_operator_rejects_rows_the_reference_throws_on(ctx):
    codes = np.zeros((6, 6, 6), dtype=np.int32)
    codes[1:-1, 1:-1, 1:-1] = 1
    codes[0, 2, 2] = 1  # an Actual row on the array face reads outside the grid
    g = xf.Grid(ctx, np.full(5, 0.2), np.full(5, 0.2), np.full(5, 0.2), codes)
    x = dev(np.ones(codes.shape))
    y = dev(np.zeros(codes.shape))
    with pytest.raises(xf.XfError):
        g.apply(x, y)


SOLVERS = {"cg": xf.XF_CG3, "bicg": xf.XF_BICG, "mr": xf.XF_MR, "bicgstab": xf.XF_BICGSTAB}
SWEEPS_TO_ITERS = {"cg": lambda s: (s - 3) // 2, "bicg": lambda s: (s - 1) // 3, "mr": lambda s: (s - 1) // 2,
                   "bicgstab": lambda s: (s - 1) // 2}


@pytest.mark.parametrize("name,solver", [("cg_dirichlet16", "cg"), ("cg_neumann20", "cg"),
                                         ("bicg_dirichlet12", "bicg"), ("mr_dirichlet8", "mr"),
                                         ("bicgstab_dirichlet12", "bicgstab")])
def test_reference_solvers(ctx, name, solver):
    ins, outs = golden(name)
    g = grid_of(ctx, ins)
    b, x0 = dev(ins["bvec"]), dev(ins["x0"])
    x = torch.zeros_like(x0)
    st = g.solve(b, x0, x, xf.solve_opts(SOLVERS[solver]))
    assert st.status == 0
    ref_iters = SWEEPS_TO_ITERS[solver](int(outs["sweeps"][0]))
    assert abs(st.iters - ref_iters) <= 2, (st.iters, ref_iters)
    assert rel_l2(host(x), outs["x"]) <= 1e-10


def sphere_case(ncell, seed=0):
    from xflow_amd import cases
    return cases.sphere_case(ncell, seed)


# (24, 24, 24), (40, 40, 40), (30, 45, 37): rows shorter than a wave (lanes past
# the row end clamped); (16, 18, 100): one double per lane, last k-tile shifted
# back; (17, 19, 131), (20, 24, 300), (20, 23, 256), (33, 10, 200): rows long
# enough for the two-doubles-per-lane stencil, with the last k-tile shifted back
# over the one before it where the row is not a whole number of tiles, and a
# last block of rows that ends past the grid (23 - 2 rows); (9, 7, 130): tiles
# of two planes.  Dirichlet variants: Defined points on the i = 0 and j = l - 1
# faces (open faces towards the ring) and a Defined box inside the grid.
@pytest.mark.parametrize("shape,dirichlet", [((24, 24, 24), False), ((40, 40, 40), False), ((30, 45, 37), False),
                                             ((16, 18, 100), False), ((17, 19, 131), False), ((20, 24, 300), False),
                                             ((20, 23, 256), False), ((33, 10, 200), False), ((9, 7, 130), False),
                                             ((24, 24, 24), True), ((20, 23, 256), True), ((17, 19, 131), True)])
def test_throughput_cg_matches_oracle(ctx, shape, dirichlet):
    from xflow_amd import cases
    codes = cases.neumann_codes(None, [cases.sphere_obstacle((0.5, 0.5, 0.5), 0.2)], shape=shape)
    if dirichlet:
        codes[0, 1:-1, 1:-1] = 3
        codes[1:-1, -1, 1:-1] = 3
        n, l, m = codes.shape
        codes[n // 4:n // 4 + 2, l // 4:l // 4 + 2, m // 4:m // 4 + 3] = 3
    b = cases.divergence_rhs(codes, seed=0)
    case = {"codes": codes, "hx": cases.axis_steps(shape[0]), "hy": cases.axis_steps(shape[1]),
            "hz": cases.axis_steps(shape[2]), "b": b}
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = dev(case["b"])
    x0 = torch.zeros_like(b)
    x = torch.zeros_like(b)
    pm = 0 if dirichlet else 1
    st = g.solve(b, x0, x, xf.solve_opts(xf.XF_CG, tol=1e-12, tol_mode=xf.XF_TOL_REL, project_mean=pm))
    assert st.status == 0
    og = xo.Grid(case["hx"], case["hy"], case["hz"], case["codes"])
    xr, it, res, s2 = xo.cg_cg(og, case["b"], np.zeros_like(case["b"]), 1e-12, 100000, project_mean=bool(pm))
    assert s2 == 0
    print("%s dirichlet=%s: gpu %d iterations, oracle %d" % (shape, dirichlet, st.iters, it))
    assert abs(st.iters - it) <= 2, (st.iters, it)
    m = case["codes"] == 1
    xg = host(x)
    if dirichlet:
        a, r = xg[m], xr[m]
    else:
        # the pure-Neumann solution is defined up to a constant: compare mean-free fields
        a = xg[m] - xg[m].mean()
        r = xr[m] - xr[m].mean()
    assert rel_l2(a, r) <= 1e-9
    assert st.resid_true <= 2e-12


def test_throughput_cg_128_sphere(ctx):
    """The 128^3 sphere workload: plain FP64 GPU CG to 1e-8 relative."""
    case = sphere_case(128)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = dev(case["b"])
    x = torch.zeros_like(b)
    st = g.solve(b, torch.zeros_like(b), x, xf.solve_opts(xf.XF_CG, tol=1e-8, tol_mode=xf.XF_TOL_REL,
                                                            project_mean=1))
    assert st.status == 0
    assert st.resid_true <= 1.5e-8
    assert 50 < st.iters < 2000


@pytest.mark.parametrize("method", [xf.XF_CG, xf.XF_CG_MP])
def test_verify_flag_skips_only_the_check(ctx, method):
    """opts.verify = 0 skips the after-solve residual check with the reference
    rows (stats.resid_true = -1) and changes nothing else: same iterations,
    bit-identical x."""
    case = sphere_case(40, seed=3)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = dev(case["b"])
    out = []
    for verify in (1, 0):
        x = torch.zeros_like(b)
        st = g.solve(b, torch.zeros_like(b), x, xf.solve_opts(method, tol=1e-9, tol_mode=xf.XF_TOL_REL,
                                                                project_mean=1, verify=verify))
        assert st.status == 0
        out.append((st, host(x)))
    (s1, x1), (s0, x0) = out
    assert s1.resid_true > 0 and s1.resid_true <= 2e-9
    assert s0.resid_true == -1.0
    assert (s1.iters, s1.inner_iters) == (s0.iters, s0.inner_iters)
    assert np.array_equal(x1, x0)
    g.close()


CONFIG2_NORMS = {"cg3": (xf.XF_CG3, 1e3), "bicg": (xf.XF_BICG, 1.0), "mr": (xf.XF_MR, 10.0)}


@pytest.mark.parametrize("solver", ["cg3", "bicg", "mr"])
def test_config2_sphere128_reference_solvers(ctx, solver):
    """The 128^3 sphere workload solved by the reference solvers (their own
    absolute tolerances; b scaled so that each means 1e-8 relative) against
    the compiled reference's own results at the same size for CG and BiCG
    (tests/golden/make_config2_reference.py: 809 and 810 iterations) and the
    oracle's for minimal residuals (make_config2.py: 151234 iterations, too
    long to run the reference for): iterations within 2, the field's norm
    over the fluid and 512 sampled values within 1e-10 (measured: the same
    iteration counts, differences 1e-14 to 1e-13)."""
    import os
    name = "config2_sphere128_reference.npz" if solver in ("cg3", "bicg") else "config2_sphere128.npz"
    fx = np.load(os.path.join(os.path.dirname(__file__), "golden", name))
    method, norm = CONFIG2_NORMS[solver]
    case = sphere_case(128)
    b = case["b"] * (norm / np.linalg.norm(case["b"]))
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    bd = dev(b)
    x = torch.zeros_like(bd)
    st = g.solve(bd, torch.zeros_like(bd), x, xf.solve_opts(method))
    assert st.status == 0, st.status
    it_ref = int(fx[solver + "_iters"][0])
    assert abs(st.iters - it_ref) <= 2, (st.iters, it_ref)
    tol = 1e-10
    xs = host(x).ravel()
    fluid = case["codes"].ravel() == 1
    dn = abs(np.linalg.norm(xs[fluid]) - fx[solver + "_norm"][0]) / fx[solver + "_norm"][0]
    dx = rel_l2(xs[fx["idx"]], fx[solver + "_x"])
    print("%s: %d iterations (oracle %d), norm %.2e, samples %.2e" % (solver, st.iters, it_ref, dn, dx))
    assert dn <= tol and dx <= tol, (dn, dx)
    g.close()
