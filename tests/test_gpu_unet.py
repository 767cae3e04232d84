"""The learned preconditioner's kernels against the CPU reference network
(oracle/unet_ref.py: torch FP32 with the same BF16 rounding points).

Tolerances: BF16 activations (8-bit mantissa) make single roundings differ
when the summation order differs, so the forward output is compared at 2e-2
relative L2 and each layer's weight gradient at 5e-2 relative L2 (the device
input-gradient path also carries BF16 gradients).
"""
import numpy as np
import pytest
import torch

import unet_ref
import xflow_amd as xf
from xflow_amd import cases
from xflow_amd.unet import LOSS_COSINE, LOSS_ENERGY, UNet, layer_slices

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = xf.Context(0)
    yield c
    c.close()


def crop_batch(seed, B=1, D=16, H=16, W=64):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((B, D, H, W)).astype(np.float32)
    mask = (rng.random((B, D, H, W)) > 0.15).astype(np.float32)
    r *= mask
    target = (0.05 * rng.standard_normal((B, D, H, W))).astype(np.float32) * mask
    return r, mask, target


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_reference(ctx):
    net = UNet(ctx, seed=3)
    r, mask, _ = crop_batch(1)
    y = torch.zeros(r.shape, dtype=torch.float32, device="cuda")
    net.forward(torch.from_numpy(r).cuda(), torch.from_numpy(mask).cuda(), y)
    torch.cuda.synchronize()
    layers = unet_ref.unflatten(net.get_params())
    want = unet_ref.forward(layers, torch.from_numpy(r), torch.from_numpy(mask)).detach().numpy()
    assert rel(y.cpu().numpy(), want) < 2e-2


def test_gradients_match_reference(ctx):
    net = UNet(ctx, seed=4)
    r, mask, target = crop_batch(2, B=2)
    loss = net.loss_grad(torch.from_numpy(r).cuda(), torch.from_numpy(mask).cuda(), torch.from_numpy(target).cuda())
    grads = net.get_grads()
    wl, wg, _ = unet_ref.loss_and_grads(net.get_params(), r, mask, target)
    assert loss == pytest.approx(wl, rel=2e-2)
    for l, (ws, bs, _) in enumerate(layer_slices()):
        assert rel(grads[ws], wg[ws]) < 5e-2, ("weights", l, rel(grads[ws], wg[ws]))
        assert rel(grads[bs], wg[bs]) < 5e-2, ("bias", l, rel(grads[bs], wg[bs]))


def test_cosine_loss_gradients_match_reference(ctx):
    net = UNet(ctx, seed=8)
    net.set_loss(LOSS_COSINE)
    r, mask, target = crop_batch(9, B=2)
    loss = net.loss_grad(torch.from_numpy(r).cuda(), torch.from_numpy(mask).cuda(), torch.from_numpy(target).cuda())
    grads = net.get_grads()
    wl, wg, _ = unet_ref.loss_and_grads(net.get_params(), r, mask, target, cosine=True)
    # the cosine gradient g (t - c y) cancels between its two terms, so BF16
    # roundings weigh more than in the squared-error test: 8e-2 relative L2
    assert loss == pytest.approx(wl, rel=2e-2, abs=2e-3)
    for l, (ws, bs, _) in enumerate(layer_slices()):
        assert rel(grads[ws], wg[ws]) < 8e-2, ("weights", l, rel(grads[ws], wg[ws]))
        assert rel(grads[bs], wg[bs]) < 8e-2, ("bias", l, rel(grads[bs], wg[bs]))


def test_energy_loss_gradients_match_reference(ctx):
    net = UNet(ctx, seed=10)
    net.set_loss(LOSS_ENERGY)
    r, mask, _ = crop_batch(11, B=2)
    # a target with (t, r) > 0: t solves a smoothed problem, r is its residual
    t = torch.from_numpy(r).unsqueeze(1)
    target = torch.nn.functional.avg_pool3d(t, 5, stride=1, padding=2)[:, 0].numpy() * mask
    rin = unet_ref.crop_laplacian(torch.from_numpy(target), torch.from_numpy(mask)).numpy().astype(np.float32)
    loss = net.loss_grad(torch.from_numpy(rin).cuda(), torch.from_numpy(mask).cuda(),
                         torch.from_numpy(np.ascontiguousarray(target, dtype=np.float32)).cuda())
    grads = net.get_grads()
    wl, wg, _ = unet_ref.loss_and_grads(net.get_params(), rin, mask, target, energy=True)
    assert 0 <= loss <= 1
    assert loss == pytest.approx(wl, rel=2e-2, abs=2e-3)
    for l, (ws, bs, _) in enumerate(layer_slices()):
        assert rel(grads[ws], wg[ws]) < 8e-2, ("weights", l, rel(grads[ws], wg[ws]))
        # the last bias moves every output alike, along the null space of the
        # Neumann crop operator: its gradient is ~0 and compared absolutely
        tol = 8e-2 * max(np.linalg.norm(wg[bs]), 1e-2 * np.linalg.norm(wg[ws]))
        assert np.linalg.norm(grads[bs] - wg[bs]) <= tol, ("bias", l, rel(grads[bs], wg[bs]))


def test_training_reduces_loss(ctx):
    net = UNet(ctx, seed=5)
    r, mask, target = crop_batch(3, B=2)
    # a learnable target: a smoothed copy of the input
    t = torch.from_numpy(r).unsqueeze(1)
    target = torch.nn.functional.avg_pool3d(t, 3, stride=1, padding=1)[:, 0].numpy() * mask
    dr, dm, dt = (torch.from_numpy(a).cuda() for a in (r, mask, target))
    first = net.train_step(dr, dm, dt, 1e-3)
    for _ in range(60):
        last = net.train_step(dr, dm, dt, 1e-3)
    assert last < 0.5 * first


def test_grid_apply_is_scaled_crop_forward(ctx):
    net = UNet(ctx, seed=6)
    case = cases.obstacle_case(64, seed=2)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    r = torch.from_numpy(case["b"].astype(np.float32)).cuda()
    z = torch.zeros_like(r)
    net.apply(g, r, z)
    m = (case["codes"] == 1).astype(np.float32)
    rr = case["b"].astype(np.float32)
    # the kernels' own arithmetic: r * float(1/sqrt(ms)) in, y * float(sqrt(ms)) out, ms the
    # mean square over the unknowns; a BF16 network amplifies any 1-ulp input change (a
    # flipped rounding cascades), so the crop is fed bit-identical inputs
    ms = float((rr[m > 0].astype(np.float64) ** 2).sum()) / float((m > 0).sum())
    inv = np.float32(1.0 / np.sqrt(ms))
    ri = ((rr * inv) * m)[1:-1, 1:-1, 1:-1][None].astype(np.float32)
    mi = m[1:-1, 1:-1, 1:-1][None].astype(np.float32)
    y = torch.zeros(ri.shape, dtype=torch.float32, device="cuda")
    net.forward(torch.from_numpy(np.ascontiguousarray(ri)).cuda(), torch.from_numpy(np.ascontiguousarray(mi)).cuda(), y)
    torch.cuda.synchronize()
    want = np.zeros_like(rr)
    want[1:-1, 1:-1, 1:-1] = (y.cpu().numpy()[0] * np.float32(np.sqrt(ms))) * mi[0]
    assert rel(z.cpu().numpy(), want) < 1e-5


def test_fcg_with_network_converges(ctx):
    net = UNet(ctx, seed=7)
    case = cases.obstacle_case(64, seed=3)
    g = xf.Grid(ctx, case["hx"], case["hy"], case["hz"], case["codes"])
    b = torch.from_numpy(case["b"]).cuda()
    x = torch.zeros_like(b)
    o = xf.solve_opts(xf.XF_FCG_NN, tol=1e-8, tol_mode=xf.XF_TOL_REL, project_mean=1, inner_tol=1e-4)
    o.precond = net.h
    st = g.solve(b, torch.zeros_like(b), x, o)
    assert st.status == 0, st.status
    assert st.resid_true <= 1.5e-8
