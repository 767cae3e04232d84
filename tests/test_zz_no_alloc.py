"""xf_solve_noalloc (plain FP64 CG on the library's fixed per-device work
array) against xf_solve with XF_CG on the same grids, through the C ABI: two
calls in a row, the second on a smaller grid (the array still holds the first
call's vectors), bit-identical pressure, the same iteration count and
residuals; no device memory taken by the call; grids beyond the array's
capacity rejected before any work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve(lib, ctx, g, b, x0, x, opts, noalloc):
    from xflow_amd._lib import SolveStats
    stats = SolveStats()
    f = lib.xf_solve_noalloc if noalloc else lib.xf_solve
    rc = f(ctx.h, g.h, C.c_void_p(b.data_ptr()), C.c_void_p(x0.data_ptr()), C.c_void_p(x.data_ptr()),
           C.byref(opts), C.byref(stats))
    return rc, stats


def test_noalloc_matches_xf_solve_two_calls():
    import torch
    import xflow_amd as xf
    from xflow_amd import cases
    from xflow_amd._lib import lib, XF_EINVAL

    L = lib()
    ctx = xf.Context(0)
    # load every kernel module once (a tiny solve) before memory is compared
    small = cases.obstacle_case(16, seed=1)
    gs = xf.Grid(ctx, small["hx"], small["hy"], small["hz"], small["codes"])
    bs = torch.from_numpy(small["b"]).cuda()
    opts = xf.solve_opts(xf.XF_CG, tol=1e-10, tol_mode=xf.XF_TOL_REL, project_mean=1, verify=1)
    rc, _ = _solve(L, ctx, gs, bs, torch.zeros_like(bs), torch.zeros_like(bs), opts, True)
    assert rc == 0, L.xf_last_error(ctx.h)
    gs.close()

    for call, (ncell, seed, x0_scale) in enumerate([(100, 3, 0.0), (40, 5, 1e-3)]):
        case = cases.obstacle_case(ncell, seed=seed)
        g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
        assert g.ncell <= 2358272
        b = torch.from_numpy(case["b"]).cuda()
        rng = np.random.default_rng(seed)
        x0 = torch.from_numpy(x0_scale * rng.standard_normal(case["b"].shape)).cuda()
        xa = torch.full_like(b, 7.0)
        xb = torch.full_like(b, -3.0)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rc_a, sa = _solve(L, ctx, g, b, x0, xa, opts, True)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        print("call %d: %d^3 cells, noalloc rc %d iters %d resid %.3e true %.3e, device memory %+d B"
              % (call, ncell, rc_a, sa.iters, sa.resid, sa.resid_true, free1 - free0), flush=True)
        assert rc_a == 0 and sa.status == 0
        # the solve's four work vectors would be 32 B per cell on the device
        assert free0 - free1 < min(4 << 20, 16 * g.ncell), (free0, free1)
        rc_b, sb = _solve(L, ctx, g, b, x0, xb, opts, False)
        assert rc_b == 0 and sb.status == 0
        A, B = xa.cpu().numpy(), xb.cpu().numpy()
        assert sa.iters == sb.iters
        assert sa.resid == sb.resid and sa.resid_true == sb.resid_true
        assert np.array_equal(A, B), np.abs(A - B).max()
        assert sa.resid_true < 1e-9
        g.close()

    # beyond the array's capacity: rejected before any work, x untouched
    big = np.zeros((134, 134, 134), dtype=np.int32)
    big[1:-1, 1:-1, 1:-1] = 1
    h = np.full(133, 1.0 / 132)
    gb = xf.Grid(ctx, h, h, h, big)
    assert gb.ncell > 2358272
    bb = torch.zeros(gb.ncell, dtype=torch.float64, device="cuda")
    xx = torch.full_like(bb, 5.0)
    rc, _ = _solve(L, ctx, gb, bb, torch.zeros_like(bb), xx, opts, True)
    assert rc == XF_EINVAL
    assert "exceeds the fixed work array" in L.xf_last_error(ctx.h).decode()
    assert bool((xx == 5.0).all())
    gb.close()
    ctx.close()
