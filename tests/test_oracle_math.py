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
    gradients of linear pressures, the density's square in the divergence
    and its division in the correction, the Defined check and PreDefined
    mirroring of the correction;
  * transport terms: four-face averages and viscosity means exact for linear
    fields, the viscosity-gradient term against its closed form, boundary
    values by point type, and the operand order through a sum whose rounding
    depends on it.
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


def test_pressure_rhs_density_closed_form():
    """The density enters the divergence as rho^2 du/dx + dv/dy + dw/dz (the
    reference's precedence, XProblem.cpp:5495): with rho = 2 and linear face
    fields rhs = -(4 a + b + c)/tau; with rho varying per cell, -(rho^2 a + b +
    c)/tau; PreDefined points take their boundary value."""
    nc = 8
    h = 1.0 / nc
    codes = cases.neumann_codes(nc, []).astype(np.int32)
    codes[0, 3:5, 3:5] = 4
    n = nc + 2
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u, v, w = 2.0 * I * h, -3.0 * J * h, 5.0 * K * h
    hs = np.full(n - 1, h)
    pbc = np.full(codes.shape, 0.25)
    act = codes == 1
    rhs = xo.pressure_rhs(codes, u, v, w, hs, hs, hs, tau=0.5, rho=np.full(codes.shape, 2.0), pbc=pbc)
    assert np.allclose(rhs[act], -(4 * 2.0 - 3.0 + 5.0) / 0.5, rtol=1e-12)
    assert np.all(rhs[codes == 4] == 0.25) and np.all(rhs[codes == 0] == 0)
    rho = 1.0 + 0.1 * I
    rhs = xo.pressure_rhs(codes, u, v, w, hs, hs, hs, tau=0.5, rho=rho)
    assert np.allclose(rhs[act], (-(rho ** 2 * 2.0 - 3.0 + 5.0) / 0.5)[act], rtol=1e-12)


def test_velocity_correction_density_and_border_branches():
    """u = u* - tau grad p / rho at Actual points (XProblem.cpp:5660-5675),
    Defined points keep their boundary value and are checked against it
    (5676-5691), PreDefined points mirror their one Actual neighbour about the
    boundary value after the other points (5692-5756)."""
    from xflow_amd.drivers import velocity_codes
    nc = 6
    h = 1.0 / nc
    n = nc + 2
    codes, _ = velocity_codes(nc, 2)  # v: Defined walls j = 0, n; PreDefined ghosts along i and k
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    hs = np.full(n - 1, h)
    tau, rho = 0.1, 2.0
    p = 7.0 * J * h  # grad p = 7 along y
    bc = np.where(codes == 3, 0.5, np.where(codes == 4, -0.25, 0.0))
    us = 1.0 + 0.3 * I + 0.2 * K
    us[codes == 3] = 0.5
    # PreDefined values consistent with the mirror check: us = 2 bc - us(inner)
    pre = np.argwhere(codes == 4)
    nb = {}
    for (i, j, k) in pre:
        inner = [(i + d * (a == 0), j + d * (a == 1), k + d * (a == 2)) for a in range(3) for d in (-1, 1)]
        inner = [q for q in inner if all(0 <= q[x] < n for x in range(3)) and codes[q] == 1]
        assert len(inner) == 1
        nb[(i, j, k)] = inner[0]
        us[i, j, k] = 2 * bc[i, j, k] - us[inner[0]]
    out = xo.velocity_correct(codes, 2, us, p, hs, tau, rho=np.full(codes.shape, rho), bc=bc)
    act = codes == 1
    assert np.allclose(out[act], (us - tau * 7.0 / rho)[act], rtol=1e-13)
    assert np.all(out[codes == 3] == 0.5)
    for q, a in nb.items():
        assert np.isclose(out[q], 2 * bc[q] - out[a], rtol=1e-13)
    assert np.all(out[codes == 0] == 0)
    bad = us.copy()
    bad[codes == 3] = 0.75  # a Defined value that disagrees with its boundary condition
    with pytest.raises(ValueError):
        xo.velocity_correct(codes, 2, bad, p, hs, tau, bc=bc)
    bad = us.copy()
    q = tuple(pre[0])
    bad[q] += 1e-6  # a PreDefined value off the mirror of its inner neighbour
    with pytest.raises(ValueError):
        xo.velocity_correct(codes, 2, bad, p, hs, tau, bc=bc)


def _face_positions(shape, h, comp):
    """Coordinates of the nodes of component comp (0..2; 3 = cell centres) on
    the pressure array shape: x_comp = i h, other axes (i - 1/2) h."""
    idx = np.indices(shape).astype(float)
    return [idx[d] * h if d == comp else (idx[d] - 0.5) * h for d in range(3)]


def test_transport_velocity_closed_forms():
    """The transport terms of XProblem.cpp:5090-5303, 5357-5417 and 5889-5944:
    four-face averages of linear velocity fields are the fields at the node;
    the viscosity mean is the viscosity at the node; the viscosity-gradient
    term is grad(vis) . (grad f_a + d_a f) for linear fields; point types give
    0 and boundary values; the reference's operand order (c, +a, -e, +a-e)
    shows in a sum whose rounding depends on it."""
    from xflow_amd.drivers import channel_scalar_codes, velocity_codes
    nc, h = 8, 1.0 / 8
    G = np.array([[0.5, -1.0, 2.0], [1.5, 0.25, -0.5], [-2.0, 1.0, 0.75]])  # G[e][d] = d f_e / d x_d
    c0 = np.array([0.1, -0.2, 0.3])
    gv = np.array([0.02, -0.03, 0.05])  # grad vis
    for axis in (1, 2, 3):
        a = axis - 1
        codes, bc = velocity_codes(nc, axis)
        bc = np.where((codes == 3) | (codes == 4), 0.125 + 0.0 * bc, 0.0)
        shape = codes.shape
        f = []
        for e in range(3):
            x = _face_positions(shape, h, e)
            f.append(c0[e] + sum(G[e][d] * x[d] for d in range(3)))
        xp = _face_positions(shape, h, 3)
        vis = 0.01 + sum(gv[d] * xp[d] for d in range(3))
        hs = [np.full(s - 1, h) for s in shape]
        tu, tv, tw, tvis, rp = xo.transport_velocity(axis, codes, *f, vis=vis, bc=bc, h=hs)
        xn = _face_positions(shape, h, a)
        act = codes == 1
        for e, t in enumerate((tu, tv, tw)):
            want = c0[e] + sum(G[e][d] * xn[d] for d in range(3))
            assert np.allclose(t[act], want[act], rtol=0, atol=1e-13), (axis, e)
            assert np.all(t[~act] == 0)
        wvis = 0.01 + sum(gv[d] * xn[d] for d in range(3))
        assert np.allclose(tvis[act], wvis[act], rtol=0, atol=1e-15)
        assert np.all(tvis[~act] == 0)
        # where every neighbour is Actual the central differences are exact
        inner = act.copy()
        for d in range(3):
            inner &= np.roll(act, 1, d) & np.roll(act, -1, d)
        want_rp = sum(gv[d] * (G[a][d] + G[d][a]) for d in range(3))
        assert inner.sum() > 0
        assert np.allclose(rp[inner], want_rp, rtol=1e-9, atol=1e-12), axis
        border = (codes == 3) | (codes == 4)
        assert np.all(rp[border] == 0.125) and np.all(rp[codes == 0] == 0)
        # operand order: at one Actual node put 1e16, 1, -1e16, 3 on the four
        # faces of another component in the order c, +a, -e, +a-e:
        # ((1e16 + 1) - 1e16) + 3 = 3 in doubles (1e16 + 1 rounds to 1e16)
        e = (a + 1) % 3
        S = [shape[1] * shape[2], shape[2], 1]
        q = np.ravel_multi_index((4, 4, 4), shape)
        fe = np.zeros(shape).ravel()
        for off, val in ((0, 1e16), (S[a], 1.0), (-S[e], -1e16), (S[a] - S[e], 3.0)):
            fe[q + off] = val
        ff = [np.zeros(shape) for _ in range(3)]
        ff[e] = fe.reshape(shape)
        out = xo.transport_velocity(axis, codes, *ff)
        assert out[e].ravel()[q] == 0.75, (axis, out[e].ravel()[q])
        bad = codes.copy()
        bad[4, 4, 4] = 7  # Equation points: the reference throws
        with pytest.raises(ValueError):
            xo.transport_velocity(axis, bad, *f)
    # the concentration: cell-centre averages, inlet value on i = 0 only
    codes = channel_scalar_codes(8, 6, 6)
    shape = codes.shape
    f = []
    for e in range(3):
        x = _face_positions(shape, h, e)
        f.append(c0[e] + sum(G[e][d] * x[d] for d in range(3)))
    xc = _face_positions(shape, h, 3)
    bc = np.full(shape, 0.5)
    tu, tv, tw, _, rp = xo.transport_velocity(0, codes, *f, bc=bc)
    act = codes == 1
    for e, t in enumerate((tu, tv, tw)):
        want = c0[e] + sum(G[e][d] * xc[d] for d in range(3))
        assert np.allclose(t[act], want[act], rtol=0, atol=1e-13)
        assert np.all(t[~act] == 0)
    I = np.indices(shape)[0]
    assert np.all(rp[(codes == 4) & (I == 0)] == 0.5)
    assert np.all(rp[(codes != 4) | (I != 0)] == 0)
    bad = codes.copy()
    bad[3, 3, 3] = 3  # Defined concentration points: the reference throws
    with pytest.raises(ValueError):
        xo.transport_velocity(0, bad, *f)


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
    assert len(levels) == 3  # (8, 8, 128) -> (4, 4, 64) -> (2, 2, 32)
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
    """Flexible CG (Notay's FCG(1), the solver's own use of the cycle) with
    the cycle as preconditioner needs far fewer iterations than with none
    (which is plain CG) on a Dirichlet problem with obstacles (a wrong
    restriction, prolongation or sign in the cycle breaks this).  The cycle
    is not symmetric positive definite with four sweeps of weight 6/7 (the
    rediscretised coarse operators overcorrect), so plain PCG is not the
    measure."""
    import mg_ref
    codes, h = _mg_case((16, 16, 128), 5)
    levels = mg_ref.hierarchy(codes, h)
    assert len(levels) == 4 and levels[-1].codes.size <= mg_ref.COARSE_CELLS  # dense solve on (2, 2, 16)
    L = levels[0]
    rng = np.random.default_rng(6)
    b = np.where(L.act_in, rng.standard_normal(codes.shape), 0).astype(np.float64)

    def A(v):
        return L.apply(v.astype(np.float32)).astype(np.float64)

    def fcg(prec, tol=1e-5, maxit=2000):
        x = np.zeros_like(b)
        r = b.copy()
        r0 = np.linalg.norm(r)
        p_old = q_old = None
        for it in range(1, maxit + 1):
            z = prec(r)
            p = z.copy() if p_old is None else z - (np.sum(z * q_old) / np.sum(p_old * q_old)) * p_old
            q = A(p)
            a = np.sum(p * r) / np.sum(p * q)
            x += a * p
            r -= a * q
            if np.linalg.norm(r) <= tol * r0:
                return it
            p_old, q_old = p, q
        return maxit

    plain = fcg(lambda r: r.copy())
    mg = fcg(lambda r: mg_ref.vcycle(levels, 0, r.astype(np.float32)).astype(np.float64))
    assert mg * 4 < plain, (mg, plain)


@pytest.mark.parametrize("dirichlet", [True, False])
def test_mg_coarse_solve_solves_the_level(dirichlet):
    """The deepest level's solve: with a boundary-value plane A z = r
    exactly; on a pure-Neumann grid A z = r - mean(r) with z mean-free
    (the minimum-norm least-squares solution)."""
    import mg_ref
    codes, h = _mg_case((4, 6, 8), 8, dirichlet=dirichlet)
    lev = mg_ref.Level(codes, h)
    assert lev.codes.size <= mg_ref.COARSE_CELLS
    rng = np.random.default_rng(9)
    r = np.where(lev.act_in, rng.standard_normal(codes.shape), 0).astype(np.float32)
    z = mg_ref.coarse_solve(lev, r)
    m = lev.act_in
    Az = lev.apply(z).astype(np.float64)[m]
    rhs = r.astype(np.float64)[m]
    if not dirichlet:
        rhs = rhs - rhs.mean()
        assert abs(z[m].astype(np.float64).mean()) <= 1e-6 * np.abs(z[m]).max()
    assert np.linalg.norm(Az - rhs) <= 1e-5 * np.linalg.norm(rhs)


def test_lincomb_and_copy_special_cases():
    """xo_lincomb and xo_copy as the reference's masked vector statements
    (XSpace.cpp:2293-2363, 2489-3010): only unknowns are written; a zero
    coefficient drops its vector entirely (an Inf there does not reach the
    result), both zero give 0; x |= s copies at unknowns only."""
    codes = np.zeros((5, 6, 7), dtype=np.int32)
    codes[1:-1, 1:-1, 1:-1] = 1
    h = np.full(4, 0.25), np.full(5, 0.2), np.full(6, 0.1)
    g = xo.Grid(*h, codes)
    rng = np.random.default_rng(2)
    s1, s2 = rng.standard_normal(codes.shape), rng.standard_normal(codes.shape)
    s2_inf = s2.copy()
    s2_inf[2, 2, 2] = np.inf
    unk = codes == 1
    for p1, p2, b, want in ((1.5, -0.5, s2, 1.5 * s1 + -0.5 * s2), (2.0, 0.0, s2_inf, 2.0 * s1),
                            (0.0, 3.0, s2, 3.0 * s2), (0.0, 0.0, s2_inf, np.zeros_like(s1))):
        res = np.full(codes.shape, 7.0)
        xo.lib().xo_lincomb(g.ref, p1, s1.ctypes.data, p2, b.ctypes.data, res.ctypes.data)
        assert np.array_equal(res[unk], want[unk]) and np.all(res[~unk] == 7.0), (p1, p2)
    res = np.full(codes.shape, 7.0)
    xo.lib().xo_copy(g.ref, s1.ctypes.data, res.ctypes.data)
    assert np.array_equal(res[unk], s1[unk]) and np.all(res[~unk] == 7.0)
