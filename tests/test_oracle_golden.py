"""The CPU oracle against the compiled reference's golden vectors.

Fixtures in tests/golden were written by tests/golden/make_golden.py from the
reference compiled by oracle/Makefile.  Integer arrays must match exactly; the
oracle repeats the reference's arithmetic in the same order, so floating-point
outputs must match bit for bit too (np.array_equal).
"""
import numpy as np
import pytest

import xf_oracle as xo
from conftest import golden


def grid_from(ins, rho=False):
    hx, hy, hz = np.diff(ins["x"]), np.diff(ins["y"]), np.diff(ins["z"])
    return xo.Grid(hx, hy, hz, ins["codes"], ins.get("space_mask"), ins.get("normals"),
                   ins.get("has_normal"), ins.get("rho") if rho else None)


@pytest.mark.parametrize("name", ["op_mixed", "op_sphere16"])
def test_masks_numbering_activity(name):
    ins, outs = golden(name)
    sm = xo.space_mask_of(ins["codes"])
    assert np.array_equal(sm, ins["space_mask"])
    rn, cnt = xo.rn_number(sm)
    assert np.array_equal(rn, outs["rn"])
    act = xo.activity(sm, rn, cnt)
    assert np.array_equal(act, outs["act"])


@pytest.mark.parametrize("name,rho", [("op_mixed", False), ("op_sphere16", False), ("divop_mixed", True)])
def test_operator_rows_and_dots(name, rho):
    ins, outs = golden(name)
    g = grid_from(ins, rho)
    y = xo.apply(g, ins["xvec"], y=ins["yinit"])
    assert np.array_equal(y, outs["Ax"])
    y2 = xo.apply(g, ins["xvec"], ins["bvec"], -1.0, y=ins["yinit"])
    assert np.array_equal(y2, outs["Ax_minus_b"])
    assert xo.apply_dot(g, ins["xvec"]) == outs["dots"][0]
    assert xo.dot(g, ins["xvec"], ins["bvec"]) == outs["dots"][1]
    if "absmax" in outs:
        rn, _ = xo.rn_number(g.space_mask)
        v, idx = xo.absmax(g, ins["xvec"], rn)
        assert v == outs["absmax"][0] and idx == outs["absmax_rn"][0]


# iterations from operator sweeps counted by the harness
SWEEPS = {
    "cg": lambda it: 3 + 2 * it,
    "bicg": lambda it: 1 + 3 * it,
    "mr": lambda it: 1 + 2 * it,
    "bicgstab": lambda it: 1 + 2 * it,
}


@pytest.mark.parametrize("name,solver,eps", [
    ("cg_dirichlet16", "cg", 1e-5), ("cg_neumann20", "cg", 1e-5), ("bicg_dirichlet12", "bicg", 1e-8),
    ("mr_dirichlet8", "mr", 1e-7), ("bicgstab_dirichlet12", "bicgstab", 1e-5)])
def test_solvers_bit_exact(name, solver, eps):
    ins, outs = golden(name)
    g = grid_from(ins)
    fn = {"cg": xo.cg3, "bicg": xo.bicg, "mr": xo.mr, "bicgstab": xo.bicgstab}[solver]
    x, it, res, st = fn(g, ins["bvec"], ins["x0"], eps)
    assert st == xo.OK
    assert outs["sweeps"][1] == 0
    assert SWEEPS[solver](it) == outs["sweeps"][0]
    assert np.array_equal(x, outs["x"])


def test_thomas():
    ins, outs = golden("thomas24")
    u = xo.thomas(ins["u"], ins["a"], ins["b"], ins["c"], ins["f"])
    assert np.array_equal(u, outs["u"])


@pytest.mark.parametrize("name", ["lod_scheme0", "lod_scheme1"])
def test_lod_transport(name):
    ins, outs = golden(name)
    g = grid_from(ins)
    f = xo.lod(g, ins["f"], ins["f_prev"], ins["u"], ins["v"], ins["w"], ins["rp"], ins["kappa"],
               float(ins["tau"][0]), int(ins["scheme"][0]))
    assert outs["tau_used"][0] == ins["tau"][0]
    assert np.array_equal(f, outs["f"])


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_ib_interp_spread(axis):
    ins, outs = golden("ib_axis%d" % axis)
    g = grid_from(ins)
    nodes = ins["nodes"]
    h = np.diff(ins["x"])[0]
    E = len(ins["x"]) - 1
    base = np.array([[xo.ib_index(nodes[q, t], t + 1, axis, h, E) for t in range(3)] for q in range(len(nodes))])
    assert np.array_equal(base, outs["base"])
    U = xo.ib_interp(g, axis, nodes, ins["vel"])
    assert np.array_equal(U, outs["U"])
    F = xo.ib_spread(g, axis, nodes, ins["force_n"])
    assert np.array_equal(F, outs["force"])


@pytest.mark.parametrize("name", ["pack_mixed", "pack_sphere10"])
def test_pack_operator_rows_and_products(name):
    """T1dPackOperator: packed rows of A and A^T (counts, values, 0-based
    columns, order) and the products A x, A^T x on the Rn-ordered x."""
    ins, outs = golden(name)
    g = grid_from(ins)
    rn, _ = xo.rn_number(ins["space_mask"])
    x_rn = np.zeros(int(rn.max()))
    x_rn[rn[rn > 0] - 1] = ins["xvec"][rn > 0]
    for sfx, tr in (("", False), ("T", True)):
        counts, vals, cols = xo.pack(g, transpose=tr)
        assert np.array_equal(counts, outs["counts" + sfx])
        assert np.array_equal(cols, outs["cols" + sfx])
        assert np.array_equal(vals, outs["vals" + sfx])
        y = xo.pack_apply(counts, vals, cols, x_rn)
        assert np.array_equal(y, outs["Ax" if not tr else "ATx"])


@pytest.mark.parametrize("name,solver", [("pack_bicg_mixed", "bicg"), ("pack_bicgstab_mixed", "bicgstab"),
                                         ("pack_mr_mixed", "mr"), ("pack_bicgstab_dirichlet12", "bicgstab"),
                                         ("pack_mr_dirichlet8", "mr")])
def test_pack_solvers_bit_exact(name, solver):
    """The reference solvers run on the packed operator (BiCG's shadow sequence
    on the transposed rows of the non-symmetric mixed-border mask):
    solution bit-exact, iteration count from the harness's row evaluations."""
    ins, outs = golden(name)
    g = grid_from(ins)
    rn, _ = xo.rn_number(ins["space_mask"])
    m = rn > 0
    b, x0 = np.zeros(int(rn.max())), np.zeros(int(rn.max()))
    b[rn[m] - 1], x0[rn[m] - 1] = ins["bvec"][m], ins["x0"][m]
    x, it, res, st = xo.pack_solve(solver, xo.pack(g), xo.pack(g, True), b, x0)
    assert st == xo.OK
    assert outs["sweeps"][1] == 0
    # BiCG's transposed copy (A.Copy(), XSLAE.cpp:205) is a plain T1dPackOperator
    # whose row evaluations the harness does not count: two counted sweeps per iteration
    assert (1 + 2 * it if solver == "bicg" else SWEEPS[solver](it)) == outs["sweeps"][0]
    assert np.array_equal(x, outs["x"])


@pytest.mark.parametrize("g", ["U", "V", "W", "P"])
def test_reference_made_masks(g):
    """Masks the reference generates itself: the cylinder-channel statement of
    its live problem and LocalPrepareVariables (XStatement.cpp:7717-7977,
    XProblem.cpp:4116-4206), run by the harness.  The oracle's space-mask rule
    and Rn numbering reproduce the reference's registered space models on all
    four grids; on the pressure grid (with its normals) the activity table and
    A x too."""
    ins, outs = golden("vp_cylinder")
    codes = outs[g + "_codes"]
    sm = xo.space_mask_of(codes)
    assert np.array_equal(sm, (outs[g + "_rn"] > 0).astype(np.int32))
    rn, cnt = xo.rn_number(sm)
    assert np.array_equal(rn, outs[g + "_rn"])
    if g != "P":
        return
    act = xo.activity(sm, rn, cnt)
    assert np.array_equal(act, outs["P_act"])
    grid = xo.Grid(np.diff(outs["P_x"]), np.diff(outs["P_y"]), np.diff(outs["P_z"]), codes, None,
                   outs["P_normals"], outs["P_has_normal"])
    y = xo.apply(grid, ins["xvec"], y=np.zeros(codes.shape))
    assert np.array_equal(y, outs["P_Ax"])
    # the generator's point types: Actual interior, PreDefined inflow (inside
    # the cylinder opening) and outflow planes, Fictive walls and edges
    assert set(np.unique(codes)) == {0, 1, 4}
    assert (codes[0] == 4).sum() > 0 and (codes[-1, 1:-1, 1:-1] == 4).all()


def test_config2_sphere128_oracle_equals_reference_run():
    """At 128^3 the oracle's summaries of CG and BiCG (make_config2.py) are
    the compiled reference's own results (make_config2_reference.py runs
    TSLAE::SolveByConjugateGradients and SolveByBiConjugateGradients on the
    same case): same iteration counts, norms and sampled values, bit for bit."""
    import os
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    o = np.load(os.path.join(here, "config2_sphere128.npz"))
    r = np.load(os.path.join(here, "config2_sphere128_reference.npz"))
    assert np.array_equal(o["idx"], r["idx"])
    for s in ("cg3", "bicg"):
        for q in ("_iters", "_norm", "_x"):
            assert np.array_equal(o[s + q], r[s + q]), s + q
    assert int(r["cg3_iters"][0]) == 809 and int(r["bicg_iters"][0]) == 810
