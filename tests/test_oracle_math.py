"""Independent pins of the oracle: closed forms, invariants and brute force.

These complement the golden vectors (tests/golden) and cover what they leave
open, chosen so that a dropped term, wrong sign or transposed index fails:
  * operator: manufactured solution u = sin x sin y sin z (-Lap u = 3u,
    XProblem.cpp:554-577) with second-order truncation error; symmetry of the
    dense matrix probed with unit vectors (as ConstructMatrix does,
    XOperator.cpp:383-488); constants in the null space of a pure-Neumann mask;
  * solvers: dense numpy solves on small systems; MR residual monotone;
  * Thomas: scipy.linalg.solve_banded;
  * IB kernel: partition of unity on a uniform grid (the cosine kernel has
    no vanishing first moment except at symmetric offsets), interpolation of
    constants exact, of linear fields exact at symmetric offsets,
    spread/interp adjointness;
  * projection glue: exact divergence of linear face fields and exact
    gradients of linear pressures.
"""
import numpy as np
import pytest
import scipy.linalg

import xf_oracle as xo
from xflow_amd import cases


def grid_from_codes(codes, h):
    n, l, m = codes.shape
    return xo.Grid(np.full(n - 1, h), np.full(l - 1, h), np.full(m - 1, h), codes)


def dense_matrix(g):
    """Columns A e_j over the unknowns (unit-vector probing)."""
    sm = g.space_mask
    idx = np.flatnonzero(sm.ravel())
    A = np.zeros((idx.size, idx.size))
    for col, c in enumerate(idx):
        e = np.zeros(sm.size)
        e[c] = 1.0
        y = xo.apply(g, e.reshape(sm.shape)).ravel()
        A[:, col] = y[idx]
    return A, idx


def test_manufactured_solution_second_order():
    errs = []
    for nc in (8, 16, 32):
        h = np.pi / nc
        x = np.arange(nc + 1) * h
        codes = np.full((nc + 1,) * 3, 3, dtype=np.int32)
        codes[1:-1, 1:-1, 1:-1] = 1
        g = grid_from_codes(codes, h)
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        u = np.sin(X) * np.sin(Y) * np.sin(Z)
        r = xo.apply(g, u, 3 * u, -1.0)
        errs.append(np.abs(r[codes == 1]).max())
    assert errs[1] < errs[0] / 3.5 and errs[2] < errs[1] / 3.5


def test_operator_symmetric_and_neumann_null_space():
    codes = cases.neumann_codes(6, [cases.sphere_obstacle((0.5, 0.5, 0.5), 0.25)]).astype(np.int32)
    g = grid_from_codes(codes, 1 / 6)
    A, idx = dense_matrix(g)
    assert np.allclose(A, A.T, atol=1e-12 * np.abs(A).max())
    ones = np.ones(idx.size)
    assert np.abs(A @ ones).max() < 1e-9 * np.abs(A).max()
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > -1e-8 * ev.max()  # positive semi-definite
    # the matrix is the 7-point Laplacian: diagonal = (#open faces)/h^2
    assert np.isclose(A[0, 0], np.count_nonzero(A[0]) - 1, rtol=0) or True
    rows = (np.abs(A) > 0).sum(axis=1)
    assert np.all(np.isclose(np.diag(A) * (1 / 6) ** 2, rows - 1))


def spd_case(nc=6):
    codes = cases.neumann_codes(nc, []).astype(np.int32)
    codes[0, 1:-1, 1:-1] = 3  # Dirichlet face -> SPD
    return codes, grid_from_codes(codes, 1.0 / nc)


@pytest.mark.parametrize("solver,eps", [("cg3", 1e-11), ("bicg", 1e-11), ("mr", 1e-11), ("bicgstab", 1e-12),
                                        ("cg_hs", 1e-13), ("cg_cg", 1e-13)])
def test_solvers_match_dense_solve(solver, eps):
    codes, g = spd_case()
    A, idx = dense_matrix(g)
    rng = np.random.default_rng(0)
    b = np.zeros(codes.shape)
    b.ravel()[idx] = rng.standard_normal(idx.size)
    want = np.linalg.solve(A, b.ravel()[idx])
    x, it, res, st = getattr(xo, solver)(g, b, np.zeros_like(b), eps)
    assert st == xo.OK
    assert np.allclose(x.ravel()[idx], want, rtol=1e-8, atol=1e-9 * np.abs(want).max())
    if solver in ("cg3", "bicg", "cg_hs", "cg_cg"):
        assert it <= idx.size + 5  # CG-type: at most n steps in exact arithmetic (+ rounding slack)


def test_cg3_and_hs_cg_iterates_agree():
    codes, g = spd_case(8)
    rng = np.random.default_rng(1)
    b = np.zeros(codes.shape)
    b[g.space_mask > 0] = rng.standard_normal(int(g.space_mask.sum()))
    b *= 1e3 / np.linalg.norm(b)
    x3, it3, _, _ = xo.cg3(g, b, np.zeros_like(b), 1e-5)
    xh, ith, _, _ = xo.cg_hs(g, b, np.zeros_like(b), 1e-8)
    assert abs(it3 - ith) <= 2
    assert np.linalg.norm(x3 - xh) <= 1e-6 * np.linalg.norm(xh)


@pytest.mark.parametrize("neumann", [False, True])
def test_single_reduction_cg_agrees_with_hs_cg(neumann):
    """Chronopoulos-Gear and Hestenes-Stiefel CG produce the same iterates in
    exact arithmetic: same iteration count (+-2 for rounding) and the same
    solution; a wrong alpha or beta recurrence changes the count or diverges."""
    from xflow_amd import cases as cs
    if neumann:
        codes = cs.neumann_codes(14, [cs.sphere_obstacle((0.5, 0.5, 0.5), 0.25)]).astype(np.int32)
        b = cs.divergence_rhs(codes, seed=4)
    else:
        codes, _ = spd_case(12)
        b = cs.divergence_rhs(codes, seed=4)
        b[codes == 1] += 0.2
    g = grid_from_codes(codes, 1.0 / (codes.shape[0] - 2))
    xh, ith, rh, sh = xo.cg_hs(g, b, np.zeros_like(b), 1e-11, project_mean=neumann)
    xc, itc, rc, sc = xo.cg_cg(g, b, np.zeros_like(b), 1e-11, project_mean=neumann)
    assert sh == xo.OK and sc == xo.OK
    assert abs(ith - itc) <= 2, (ith, itc)
    m = g.space_mask > 0
    a, r = xh[m], xc[m]
    if neumann:
        a, r = a - a.mean(), r - r.mean()
    assert np.linalg.norm(a - r) <= 1e-9 * np.linalg.norm(a)


def test_mr_residual_monotone():
    codes, g = spd_case(5)
    rng = np.random.default_rng(2)
    b = np.zeros(codes.shape)
    b[g.space_mask > 0] = rng.standard_normal(int(g.space_mask.sum()))
    prev = np.inf
    for it in range(1, 30):
        x, k, res, st = xo.mr(g, b, np.zeros_like(b), 0.0, maxit=it)
        assert res <= prev * (1 + 1e-12)
        prev = res


def test_thomas_matches_banded_solve():
    rng = np.random.default_rng(3)
    M = 40
    a, c = rng.uniform(-1, 1, M + 1), rng.uniform(-1, 1, M + 1)
    b = 3 + rng.uniform(0, 1, M + 1)
    f = rng.standard_normal(M + 1)
    u = np.zeros(M + 1)
    u[0], u[M] = 0.5, -0.25
    got = xo.thomas(u, a, b, c, f)
    rhs = f[1:M].copy()
    rhs[0] -= a[1] * u[0]
    rhs[-1] -= c[M - 1] * u[M]
    ab = np.zeros((3, M - 1))
    ab[0, 1:] = c[1:M - 1]
    ab[1] = b[1:M]
    ab[2, :-1] = a[2:M]
    want = scipy.linalg.solve_banded((1, 1), ab, rhs)
    assert np.allclose(got[1:M], want, rtol=1e-12, atol=1e-12)
    assert got[0] == u[0] and got[M] == u[M]


def test_ib_delta_moments():
    h = 0.05
    for x in np.linspace(0, h, 7):
        pts = np.arange(-5, 6) * h
        w = np.array([xo.ib_delta(abs(x - p), h) for p in pts])
        assert np.isclose((w * h).sum(), 1.0, atol=1e-14)  # partition of unity
    for x in (0.0, 0.5 * h):  # symmetric offsets: the first moment vanishes
        w = np.array([xo.ib_delta(abs(x - p), h) for p in pts])
        assert abs((w * h * (pts - x)).sum()) < 1e-15
    assert xo.ib_delta(2 * h, h) == 0.0 and xo.ib_delta(0.0, h) == pytest.approx(0.5 / h)


def test_ib_interp_linear_exact_and_adjoint():
    nc = 20
    h = 1.0 / nc
    codes = np.ones((nc + 2,) * 3, dtype=np.int32)
    g = grid_from_codes(codes, h)
    rng = np.random.default_rng(4)
    # node coordinates on cell-centre / face offsets (symmetric kernel positions)
    xyz = (rng.integers(6, 14, (15, 3)) + 0.5 * rng.integers(0, 2, (15, 3))) * h
    idx = np.arange(nc + 2) * h
    for axis in (1, 2, 3):
        # staggered coordinates of this component's nodes (XProblem.cpp:9005)
        cx = [idx - (0.5 * h if t + 1 != axis else 0) for t in range(3)]
        X, Y, Z = np.meshgrid(cx[0], cx[1], cx[2], indexing="ij")
        lin = 0.3 + 1.5 * X - 0.7 * Y + 2.0 * Z
        U = xo.ib_interp(g, axis, xyz, lin)
        want = 0.3 + 1.5 * xyz[:, 0] - 0.7 * xyz[:, 1] + 2.0 * xyz[:, 2]
        assert np.allclose(U, want, atol=1e-12)
        Ur = xo.ib_interp(g, axis, rng.uniform(0.3, 0.7, (15, 3)), np.full(codes.shape, 2.5))
        assert np.allclose(Ur, 2.5, atol=1e-13)
        F = rng.standard_normal(len(xyz))
        u = rng.standard_normal(codes.shape)
        lhs = (xo.ib_spread(g, axis, xyz, F) * u).sum()
        rhs = (F * xo.ib_interp(g, axis, xyz, u)).sum() / h ** 3
        assert np.isclose(lhs, rhs, rtol=1e-12)


def test_projection_glue_closed_forms():
    nc = 10
    h = 1.0 / nc
    codes = cases.neumann_codes(nc, []).astype(np.int32)
    n = nc + 2
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    # faces at x = i*h (face between cells i and i+1): u = 2x, v = -3y, w = 5z
    u, v, w = 2.0 * I * h, -3.0 * J * h, 5.0 * K * h
    hs = np.full(n - 1, h)
    rhs = xo.pressure_rhs(codes, u, v, w, hs, hs, hs, tau=0.5)
    assert np.allclose(rhs[codes == 1], -(2.0 - 3.0 + 5.0) / 0.5)
    assert np.all(rhs[codes == 0] == 0)
    val, cell = xo.divergence_max(codes, u, v, w, hs, hs, hs)
    assert np.isclose(val, 4.0) and codes.ravel()[cell] == 1
    p = 7.0 * I * h  # linear pressure: grad p = 7 along x
    ucodes = codes.copy()
    out = xo.velocity_correct(ucodes, 1, u, p, hs, tau=0.1)
    assert np.allclose(out[ucodes == 1], (u - 0.1 * 7.0)[ucodes == 1])
    with pytest.raises(ValueError):
        bad = codes.copy()
        bad[0, 3, 3] = 6  # PreNormal pressure points: the reference throws
        xo.pressure_rhs(bad, u, v, w, hs, hs, hs, tau=0.5)


def test_oracle_cavity_projection_is_divergence_free():
    """Two projection steps of the cavity on the oracle: after the correction
    the discrete divergence is at the pressure solve's tolerance."""
    import sys
    import os
    sys.path.insert(0, os.path.dirname(__file__))
    from test_gpu_cavity import oracle_cavity
    n = 8
    vel, p = oracle_cavity(n, 2)
    st = np.full(n + 1, 1.0 / n)
    pcodes = cases.neumann_codes(n, []).astype(np.int32)
    div, _ = xo.divergence_max(pcodes, vel[0], vel[1], vel[2], st, st, st)
    assert div < 1e-8
    assert np.abs(vel[0]).max() > 0.1  # the lid drives the flow


# ------------------------------------------------------------ V-cycle restatement
def _mg_case(shape, seed, dirichlet=True):
    from xflow_amd import cases as cs
    codes = cs.neumann_codes(None, cs.random_obstacles(seed), shape=shape).astype(np.int32)
    if dirichlet:
        codes[0, 1:-1, 1:-1] = 3
    h = tuple(1.0 / s for s in shape)
    return codes, h


def test_mg_operator_matches_oracle_rows():
    """The restatement's A equals the oracle's reference rows on Actual cells
    (power-of-two steps: the coefficient form is exact)."""
    import mg_ref
    codes, h = _mg_case((16, 8, 32), 1)
    lev = mg_ref.Level(codes, h)
    g = xo.Grid(np.full(17, h[0]), np.full(9, h[1]), np.full(33, h[2]), codes)
    rng = np.random.default_rng(3)
    z = rng.standard_normal(codes.shape).astype(np.float32)
    want = np.zeros(codes.shape)
    z64 = z.astype(np.float64)  # keep the buffer alive across the call
    xo.lib().xo_apply(g.ref, z64.ctypes.data, None, 0.0, want.ctypes.data)
    got = lev.apply(z)
    m = lev.act_in
    assert np.allclose(got[m], want[m], rtol=1e-5, atol=1e-5 * np.abs(want[m]).max())


def test_mg_transfer_keeps_smooth_fields():
    """Restriction then prolongation returns a constant unchanged on the unknowns."""
    import mg_ref
    codes, h = _mg_case((8, 8, 128), 2, dirichlet=False)
    levels = mg_ref.hierarchy(codes, h)
    assert len(levels) == 2
    t = np.where(levels[0].unknown, np.float32(2.5), np.float32(0))
    rc = mg_ref.restrict(levels[0], t, levels[1])
    assert np.allclose(rc[levels[1].unknown], 2.5)
    back = mg_ref.prolong(levels[0], rc)
    assert np.allclose(back[levels[0].unknown & levels[0].inner], 2.5)


def test_mg_jacobi_reduces_energy_error():
    import mg_ref
    codes, h = _mg_case((12, 12, 16), 3)
    lev = mg_ref.Level(codes, h)
    rng = np.random.default_rng(4)
    e = np.where(lev.act_in, rng.standard_normal(codes.shape), 0).astype(np.float32)
    r = lev.apply(e)
    z = lev.jacobi(r)
    errs = []
    for _ in range(6):
        d = (e - z).astype(np.float64)
        errs.append(float(np.sum(d * lev.apply(d.astype(np.float32)))))
        z = lev.jacobi(r, z)
    assert all(b < a for a, b in zip(errs, errs[1:]))


def test_mg_cycle_preconditions_cg():
    """PCG with the cycle as preconditioner needs far fewer iterations than CG
    on a Dirichlet problem with obstacles (a wrong restriction, prolongation
    or sign in the cycle breaks this)."""
    import mg_ref
    codes, h = _mg_case((16, 16, 128), 5)
    levels = mg_ref.hierarchy(codes, h)
    assert len(levels) == 2
    L = levels[0]
    rng = np.random.default_rng(6)
    b = np.where(L.act_in, rng.standard_normal(codes.shape), 0).astype(np.float64)

    def A(v):
        return L.apply(v.astype(np.float32)).astype(np.float64)

    def pcg(prec, tol=1e-5, maxit=2000):
        x = np.zeros_like(b)
        r = b.copy()
        z = prec(r)
        p = z.copy()
        rz = np.sum(r * z)
        r0 = np.linalg.norm(r)
        for it in range(1, maxit + 1):
            q = A(p)
            a = rz / np.sum(p * q)
            x += a * p
            r -= a * q
            if np.linalg.norm(r) <= tol * r0:
                return it
            z = prec(r)
            rz_new = np.sum(r * z)
            p = z + (rz_new / rz) * p
            rz = rz_new
        return maxit

    plain = pcg(lambda r: r.copy())
    mg = pcg(lambda r: mg_ref.vcycle(levels, 0, r.astype(np.float32)).astype(np.float64))
    assert mg * 4 < plain, (mg, plain)
