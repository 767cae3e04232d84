"""The learned preconditioner (xf_unet) from Python: parameters, crop
forward/backward, Adam training and the grid preconditioner."""
import ctypes as C

import numpy as np

from ._lib import lib

LOSS_MSE, LOSS_COSINE, LOSS_ENERGY = 0, 1, 2
LCIN = (2, 16, 16, 32, 32, 64, 96, 32, 48, 16)
LCOUT = (16, 16, 32, 32, 64, 64, 32, 32, 16, 1)


def layer_slices():
    """(weight slice, bias slice, weight shape [Cout, 27, Cin]) per layer."""
    out, off = [], 0
    for ci, co in zip(LCIN, LCOUT):
        nw = co * 27 * ci
        out.append((slice(off, off + nw), slice(off + nw, off + nw + co), (co, 27, ci)))
        off += nw + co
    return out


def _p(t):
    return C.c_void_p(t.data_ptr())


class UNet:
    def __init__(self, ctx, seed=0):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(lib().xf_unet_create(ctx.h, int(seed), C.byref(h)), "xf_unet_create")
        self.h = h
        n = C.c_int64()
        lib().xf_unet_param_count(h, C.byref(n))
        self.nparam = n.value

    def get_params(self):
        a = np.zeros(self.nparam, dtype=np.float32)
        self.ctx.check(lib().xf_unet_get_params(self.ctx.h, self.h, a.ctypes.data), "xf_unet_get_params")
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.size == self.nparam
        self.ctx.check(lib().xf_unet_set_params(self.ctx.h, self.h, a.ctypes.data), "xf_unet_set_params")

    def set_loss(self, mode):
        """LOSS_MSE (mean masked squared error) or LOSS_COSINE (per sample
        1 - cos^2 between output and target over the mask, scale-free)."""
        self.ctx.check(lib().xf_unet_set_loss(self.ctx.h, self.h, int(mode)), "xf_unet_set_loss")

    def get_grads(self):
        a = np.zeros(self.nparam, dtype=np.float32)
        self.ctx.check(lib().xf_unet_get_grads(self.ctx.h, self.h, a.ctypes.data), "xf_unet_get_grads")
        return a

    def forward(self, r, mask, y):
        B, D, H, W = r.shape
        self.ctx.check(lib().xf_unet_forward(self.ctx.h, self.h, B, D, H, W, _p(r), _p(mask), _p(y)),
                       "xf_unet_forward")
        return y

    def loss_grad(self, r, mask, target):
        B, D, H, W = r.shape
        loss = C.c_double()
        self.ctx.check(lib().xf_unet_loss_grad(self.ctx.h, self.h, B, D, H, W, _p(r), _p(mask), _p(target),
                                               C.byref(loss)), "xf_unet_loss_grad")
        return loss.value

    def train_step(self, r, mask, target, lr):
        B, D, H, W = r.shape
        loss = C.c_double()
        self.ctx.check(lib().xf_unet_train_step(self.ctx.h, self.h, B, D, H, W, _p(r), _p(mask), _p(target),
                                                C.c_float(lr), C.byref(loss)), "xf_unet_train_step")
        return loss.value

    def apply(self, grid, r, z):
        self.ctx.check(lib().xf_unet_apply(self.ctx.h, self.h, grid.h, _p(r), _p(z)), "xf_unet_apply")
        return z

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().xf_unet_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
