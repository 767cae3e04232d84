"""CPU reference of the learned preconditioner network — TEST INFRASTRUCTURE ONLY.

torch (CPU, FP32) restatement of the U-Net in xflow_amd/csrc/xf_unet.hip with
BF16 rounding at the points where the device kernels round (activations after
each layer and pool, packed weights of c1..c8); the rounding is a
straight-through operation so autograd gives the FP32 gradient of the
rounded network.  Used by tests/test_gpu_unet.py as the checker of the
forward pass and of every parameter gradient.
"""
import numpy as np
import torch
import torch.nn.functional as F

LCIN = (2, 16, 16, 32, 32, 64, 96, 32, 48, 16)
LCOUT = (16, 16, 32, 32, 64, 64, 32, 32, 16, 1)


def rb(t):
    """Round to bfloat16, identity gradient."""
    return t + (t.to(torch.bfloat16).float() - t).detach()


def unflatten(params):
    """Flat FP32 parameter vector -> [(W [Cout, Cin, 3, 3, 3], b [Cout])]."""
    p = torch.as_tensor(np.asarray(params, dtype=np.float32))
    out, off = [], 0
    for ci, co in zip(LCIN, LCOUT):
        nw = co * 27 * ci
        w = p[off:off + nw].reshape(co, 27, ci).permute(0, 2, 1).reshape(co, ci, 3, 3, 3)
        b = p[off + nw:off + nw + co]
        out.append((w.clone().requires_grad_(True), b.clone().requires_grad_(True)))
        off += nw + co
    return out


def flatten_grads(layers):
    parts = []
    for w, b in layers:
        co, ci = w.shape[0], w.shape[1]
        parts.append(w.grad.reshape(co, ci, 27).permute(0, 2, 1).reshape(-1))
        parts.append(b.grad.reshape(-1))
    return torch.cat(parts).numpy()


def forward(layers, r, mask):
    """r, mask: [B, D, H, W] float32 tensors -> y [B, D, H, W]."""
    x = torch.stack([r, mask], dim=1)  # [B, 2, D, H, W]

    def conv(t, l, relu=True, round_w=True):
        w, b = layers[l]
        y = F.conv3d(t, rb(w) if round_w else w, b, padding=1)
        return rb(F.relu(y)) if relu else y

    up = lambda t: t.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
    a0 = conv(x, 0, round_w=False)
    a1 = conv(a0, 1)
    p1 = rb(F.avg_pool3d(a1, 2))
    a2 = conv(p1, 2)
    a3 = conv(a2, 3)
    p2 = rb(F.avg_pool3d(a3, 2))
    a4 = conv(p2, 4)
    a5 = conv(a4, 5)
    a6 = conv(torch.cat([up(a5), a3], 1), 6)
    a7 = conv(a6, 7)
    a8 = conv(torch.cat([up(a7), a1], 1), 8)
    y = conv(a8, 9, relu=False, round_w=False)
    return y[:, 0]


def crop_laplacian(y, mask):
    """L y on [B, D, H, W] crops: sum over faces to masked neighbours inside the
    crop of (y - y_nb), at masked cells (0 elsewhere)."""
    out = torch.zeros_like(y)
    for dim in (1, 2, 3):
        n = y.shape[dim]
        a = y.narrow(dim, 0, n - 1)
        b = y.narrow(dim, 1, n - 1)
        ma = mask.narrow(dim, 0, n - 1)
        mb = mask.narrow(dim, 1, n - 1)
        open_ = (ma > 0) & (mb > 0)
        d = torch.where(open_, a - b, torch.zeros_like(a))
        out.narrow(dim, 0, n - 1).add_(d)
        out.narrow(dim, 1, n - 1).sub_(d)
    return out * (mask > 0)


def loss_and_grads(params, r, mask, target, cosine=False, energy=False):
    """Loss and parameter gradients by autograd: the masked mean squared error,
    or (cosine=True) the mean over samples of 1 - (y,t)^2/((y,y)(t,t)) over the
    mask, or (energy=True) 1 - (y,Lt)^2/((y,Ly)(t,Lt)) with L = crop_laplacian."""
    layers = unflatten(params)
    r, mask, target = (torch.as_tensor(np.asarray(a, dtype=np.float32)) for a in (r, mask, target))
    y = forward(layers, r, mask)
    if energy:
        q = crop_laplacian(y, mask)
        qt = crop_laplacian(target, mask)
        m = (mask > 0)
        a = (torch.where(m, y * qt, 0).double()).flatten(1).sum(1)
        bq = (torch.where(m, y * q, 0).double()).flatten(1).sum(1)
        c = (torch.where(m, target * qt, 0).double()).flatten(1).sum(1)
        loss = (1 - a * a / (bq * c)).mean()
    elif cosine:
        ym = (y * mask).double().flatten(1)
        tm = (target * mask).double().flatten(1)
        yt, yy, tt = (ym * tm).sum(1), (ym * ym).sum(1), (tm * tm).sum(1)
        loss = (1 - yt * yt / (yy * tt)).mean()
    else:
        loss = (((y - target) * mask) ** 2).sum() / mask.sum()
    loss.backward()
    return float(loss.detach()), flatten_grads(layers), y.detach().numpy()
