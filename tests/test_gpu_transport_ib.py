"""GPU parity of the LOD transport step, batched Thomas and IB interpolate /
spread against the reference's golden vectors (tests/golden) and the oracle.

Transport and Thomas use only +-*/ in the reference's order: bit-exact.  IB
weights go through cos(), whose device and libm implementations may differ
in the last bit: 1e-14 relative (a cos difference of 1 ulp moves a weight
by ~1e-16 relative).  Node window indices are bit-exact.
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


@pytest.mark.parametrize("name", ["lod_scheme0", "lod_scheme1"])
def test_lod_transport_bit_exact(ctx, name):
    ins, outs = golden(name)
    g = xf.Grid(ctx, np.diff(ins["x"]), np.diff(ins["y"]), np.diff(ins["z"]), ins["codes"])
    f = dev(ins["f"])
    g.lod_transport(f, dev(ins["f_prev"]), dev(ins["u"]), dev(ins["v"]), dev(ins["w"]), dev(ins["rp"]),
                    dev(ins["kappa"]), float(ins["tau"][0]), int(ins["scheme"][0]))
    assert np.array_equal(host(f), outs["f"])


def test_lod_transport_large_matches_oracle(ctx):
    """A 48x40x36 grid with obstacles and every border type: GPU == oracle bit for bit."""
    rng = np.random.default_rng(21)
    n, l, m = 48, 40, 36
    codes = np.full((n, l, m), 3, dtype=np.int32)
    codes[1:-1, 1:-1, 1:-1] = 1
    codes[0], codes[-1] = 4, 5
    codes[:, 0], codes[:, -1] = 3, 6
    codes[:, :, 0], codes[:, :, -1] = 7, 4
    codes[10:14, 8:20, 5:9] = 3
    codes[30:33, 25:28, 20:30] = 4
    hx, hy, hz = rng.uniform(0.02, 0.03, n - 1), rng.uniform(0.02, 0.03, l - 1), rng.uniform(0.02, 0.03, m - 1)
    fields = {k: rng.standard_normal((n, l, m)) for k in ("f", "f_prev", "u", "v", "w", "rp")}
    kappa = rng.uniform(0.001, 0.01, (n, l, m))
    g = xf.Grid(ctx, hx, hy, hz, codes)
    og = xo.Grid(hx, hy, hz, codes)
    for scheme in (0, 1):
        f = dev(fields["f"])
        g.lod_transport(f, dev(fields["f_prev"]), dev(fields["u"]), dev(fields["v"]), dev(fields["w"]),
                        dev(fields["rp"]), dev(kappa), 0.005, scheme)
        want = xo.lod(og, fields["f"], fields["f_prev"], fields["u"], fields["v"], fields["w"], fields["rp"],
                      kappa, 0.005, scheme)
        assert np.array_equal(host(f), want)


def test_lod_transport_batch_equals_single_calls(ctx):
    """Three problems on grids of different shapes (the golden grid, a
    48x40x36 grid with obstacles, and its transpose-like 36x40x48 variant)
    advanced in one batched call: each equals its own single call bit for bit."""
    rng = np.random.default_rng(5)
    probs = []
    ins, _ = golden("lod_scheme0")
    probs.append((np.diff(ins["x"]), np.diff(ins["y"]), np.diff(ins["z"]), ins["codes"]))
    for (n, l, m) in ((48, 40, 36), (36, 40, 48)):
        codes = np.full((n, l, m), 3, dtype=np.int32)
        codes[1:-1, 1:-1, 1:-1] = 1
        codes[0], codes[-1] = 4, 5
        codes[:, :, 0], codes[:, :, -1] = 7, 4
        codes[10:14, 8:20, 5:9] = 3
        probs.append((rng.uniform(0.02, 0.03, n - 1), rng.uniform(0.02, 0.03, l - 1), rng.uniform(0.02, 0.03, m - 1),
                      codes))
    grids, fields = [], []
    for hx, hy, hz, codes in probs:
        grids.append(xf.Grid(ctx, hx, hy, hz, codes))
        fields.append({k: dev(rng.standard_normal(codes.shape)) for k in ("f", "f_prev", "u", "v", "w", "rp")})
        fields[-1]["kappa"] = dev(rng.uniform(0.001, 0.01, codes.shape))
    for scheme in (0, 1):
        single = []
        for g, fl in zip(grids, fields):
            f = fl["f"].clone()
            g.lod_transport(f, fl["f_prev"], fl["u"], fl["v"], fl["w"], fl["rp"], fl["kappa"], 0.004, scheme)
            single.append(host(f))
        outs = [fl["f"].clone() for fl in fields]
        xf.lod_transport_batch(ctx, grids, outs, *[[fl[k] for fl in fields] for k in ("f_prev", "u", "v", "w", "rp",
                                                                                       "kappa")], 0.004, scheme)
        for got, want in zip(outs, single):
            assert np.array_equal(host(got), want)


def test_lod_rejects_bad_sector(ctx):
    codes = np.full((6, 6, 6), 3, dtype=np.int32)
    codes[1:-1, 1:-1, 1:-1] = 1
    codes[0, 2, 2] = 0  # a Fictive sector border: the reference throws
    g = xf.Grid(ctx, np.full(5, 0.2), np.full(5, 0.2), np.full(5, 0.2), codes)
    z = dev(np.zeros(codes.shape))
    with pytest.raises(xf.XfError):
        g.lod_transport(z.clone(), z, z, z, z, z, z + 0.1, 0.01, 0)


def test_thomas_batched_bit_exact(ctx):
    ins, outs = golden("thomas24")
    u = dev(np.stack([ins["u"]] * 3))
    a, b, c, f = (dev(np.stack([ins[k]] * 3)) for k in "abcf")
    xf.thomas_batched(ctx, u, a, b, c, f)
    got = host(u)
    for s in range(3):
        assert np.array_equal(got[s], outs["u"])


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_ib_interp_spread(ctx, axis):
    ins, outs = golden("ib_axis%d" % axis)
    g = xf.Grid(ctx, np.diff(ins["x"]), np.diff(ins["y"]), np.diff(ins["z"]), ins["codes"])
    xyz = dev(ins["nodes"])
    U = torch.zeros(xyz.shape[0], dtype=torch.float64, device="cuda")
    g.ib_interp(axis, xyz, dev(ins["vel"]), U)
    np.testing.assert_allclose(host(U), outs["U"], rtol=1e-14, atol=1e-14 * np.abs(outs["U"]).max())
    force = torch.zeros(g.ncell, dtype=torch.float64, device="cuda").reshape(g.shape)
    g.ib_spread(axis, xyz, dev(ins["force_n"]), force)
    ref = outs["force"]
    np.testing.assert_allclose(host(force), ref, rtol=1e-14, atol=1e-14 * np.abs(ref).max())


def test_ib_adjoint_and_partition_of_unity(ctx):
    """<spread F, u> = h^-3 <F, interp u> (exact adjoints) and interp(1) = 1."""
    rng = np.random.default_rng(3)
    nc = 32
    h = 1.0 / nc
    codes = np.ones((nc + 2,) * 3, dtype=np.int32)
    g = xf.Grid(ctx, np.full(nc + 1, h), np.full(nc + 1, h), np.full(nc + 1, h), codes)
    xyz = dev(rng.uniform(0.2, 0.8, (200, 3)))
    F = dev(rng.standard_normal(200))
    u = dev(rng.standard_normal(codes.shape))
    for axis in (1, 2, 3):
        U = torch.zeros(200, dtype=torch.float64, device="cuda")
        g.ib_interp(axis, xyz, u, U)
        force = torch.zeros_like(u)
        g.ib_spread(axis, xyz, F, force)
        lhs = float((force * u).sum())
        rhs = float((F * U).sum()) / h ** 3
        assert lhs == pytest.approx(rhs, rel=1e-12)
        one = torch.ones_like(u)
        g.ib_interp(axis, xyz, one, U)
        np.testing.assert_allclose(host(U), 1.0, rtol=1e-12)


def test_ib_many_nodes_match_oracle(ctx):
    """20 000 nodes (79 chunks of 256, more than the gather keeps in LDS), a
    cluster in one corner where the windows are clamped to the grid, an
    anisotropic grid: spreading and interpolation equal the oracle's serial
    scatter and sums to 1e-14 relative (device and libm cos)."""
    rng = np.random.default_rng(17)
    n, l, m = 42, 38, 34
    hx, hy, hz = 1.0 / 40, 1.0 / 36, 1.0 / 32
    codes = np.ones((n, l, m), dtype=np.int32)
    g = xf.Grid(ctx, np.full(n - 1, hx), np.full(l - 1, hy), np.full(m - 1, hz), codes)
    og = xo.Grid(np.full(n - 1, hx), np.full(l - 1, hy), np.full(m - 1, hz), codes)
    pts = np.concatenate([rng.uniform(0.05, 0.95, (18000, 3)), rng.uniform(0.0, 0.08, (2000, 3))])
    F = rng.standard_normal(pts.shape[0])
    vel = rng.standard_normal((n, l, m))
    for axis in (1, 2, 3):
        force = torch.full((n, l, m), np.nan, dtype=torch.float64, device="cuda")  # every cell is written
        g.ib_spread(axis, dev(pts), dev(F), force)
        want = xo.ib_spread(og, axis, pts, F)
        np.testing.assert_allclose(host(force), want, rtol=1e-14, atol=1e-14 * np.abs(want).max())
        U = torch.zeros(pts.shape[0], dtype=torch.float64, device="cuda")
        g.ib_interp(axis, dev(pts), dev(vel), U)
        wantU = xo.ib_interp(og, axis, pts, vel)
        np.testing.assert_allclose(host(U), wantU, rtol=1e-14, atol=1e-14 * np.abs(wantU).max())
    force = torch.full((n, l, m), np.nan, dtype=torch.float64, device="cuda")
    g.ib_spread(1, dev(np.zeros((0, 3))), dev(np.zeros(0)), force)  # no nodes: zeros everywhere
    assert not host(force).any()


def test_ib_update_position(ctx):
    xyz = dev(np.arange(12, dtype=np.float64).reshape(4, 3))
    prev = torch.zeros_like(xyz)
    U = dev(np.array([1.0, 2.0, 3.0, 4.0]))
    xf.ib_update_position(ctx, 2, xyz, U, 0.5, prev)
    got = host(xyz)
    assert np.array_equal(got[:, 1], np.array([1, 4, 7, 10]) + 0.5 * np.array([1, 2, 3, 4]))
    assert np.array_equal(host(prev)[:, 1], np.array([1, 4, 7, 10], dtype=np.float64))
