"""Per-kernel HIP-event timing of the throughput solvers (xf_profile_* in the C ABI)."""
import ctypes as C

from ._lib import lib

KERNELS = ("cg_dots_fp64", "cg_step_fp64", "cg_fused_fp64", "cg_dots_fp32", "cg_step_fp32", "cg_fused_fp32",
           "mp_residual_fp64", "mp_correct")


def enable(ctx, on=True):
    lib().xf_profile_reset(ctx.h)
    ctx.check(lib().xf_profile_enable(ctx.h, 1 if on else 0), "xf_profile_enable")


def read(ctx):
    out = {}
    for k in KERNELS:
        ms, n = C.c_double(), C.c_int64()
        lib().xf_profile_read(ctx.h, k.encode(), C.byref(ms), C.byref(n))
        if n.value:
            out[k] = {"ms": ms.value, "count": n.value}
    return out


def stream_ceiling(ctx, kind, nbytes=1 << 30, reps=10):
    """GB/s of a streaming kernel on this GPU: kind 0 read-only, 1 copy,
    2 two reads and two writes in place (xf_stream_ceiling)."""
    out = C.c_double()
    ctx.check(lib().xf_stream_ceiling(ctx.h, int(kind), int(nbytes), int(reps), C.byref(out)), "xf_stream_ceiling")
    return out.value
