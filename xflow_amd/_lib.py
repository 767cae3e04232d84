"""ctypes binding of libxflow_hip.so (the C ABI declared in include/xflow.h).

The library is built in-tree by __graft_entry__.build() (make -C xflow_amd/csrc).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxflow_hip.so")

XF_OK, XF_EINVAL, XF_EMAXIT, XF_EBREAKDOWN, XF_ENAN, XF_EHIP, XF_ENOMEM, XF_ESTALL = 0, -1, -2, -3, -4, -5, -6, -7
XF_CG3, XF_BICG, XF_MR, XF_BICGSTAB, XF_CG, XF_CG_MP, XF_FCG_NN, XF_FCG_MGNN, XF_BICGSTAB_T = range(9)
XF_TOL_ABS, XF_TOL_REL = 0, 1

STATUS_NAMES = {0: "ok", -1: "invalid", -2: "max-iterations", -3: "breakdown", -4: "nan", -5: "hip-error",
                -6: "out-of-memory", -7: "stalled"}


class SolveOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("tol_mode", C.c_int), ("tol", C.c_double), ("maxit", C.c_int),
                ("check_every", C.c_int), ("project_mean", C.c_int), ("inner_tol", C.c_double),
                ("stall_window", C.c_int), ("precond", C.c_void_p), ("verify", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("inner_iters", C.c_int), ("status", C.c_int), ("fallback", C.c_int),
                ("resid", C.c_double), ("resid_true", C.c_double), ("ms", C.c_double)]


class VtkFlow(C.Structure):
    """xf_vtk_flow"""
    _fields_ = [(n, C.c_void_p) for n in ("mask_grid_u", "mask_grid_p", "mask_grid_d", "mask_grid_c", "x", "y", "z",
                                          "u", "v", "w", "p", "density", "concentration", "viscosity")]


# name -> (restype, argtypes); the export test checks this table against include/xflow.h.
P, I, I64, D, SZ = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_size_t
SIGNATURES = {
    "xf_init": (I, [I, C.POINTER(P)]),
    "xf_shutdown": (I, [P]),
    "xf_last_error": (C.c_char_p, [P]),
    "xf_set_stream": (I, [P, P]),
    "xf_synchronize": (I, [P]),
    "xf_set_deferred_errors": (I, [P, I]),
    "xf_check_errors": (I, [P]),
    "xf_malloc": (I, [P, SZ, C.POINTER(P)]),
    "xf_free": (I, [P, P]),
    "xf_memcpy": (I, [P, P, P, SZ, I]),
    "xf_grid_create": (I, [P, I, I, I, P, P, P, P, P, P, C.POINTER(P)]),
    "xf_grid_destroy": (I, [P, P]),
    "xf_grid_info": (I, [P, C.POINTER(I64), C.POINTER(I64)]),
    "xf_grid_export": (I, [P, P, P, P, P]),
    "xf_grid_set_density": (I, [P, P, P]),
    "xf_apply": (I, [P, P, P, P, D, P]),
    "xf_apply_dot": (I, [P, P, P, C.POINTER(D)]),
    "xf_lincomb": (I, [P, P, D, P, D, P, P]),
    "xf_copy": (I, [P, P, P, P]),
    "xf_dot": (I, [P, P, P, P, C.POINTER(D)]),
    "xf_absmax": (I, [P, P, P, C.POINTER(D), C.POINTER(C.c_int32)]),
    "xf_solve_opts_default": (None, [I, C.POINTER(SolveOpts)]),
    "xf_solve": (I, [P, P, P, P, P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "xf_solve_noalloc": (I, [P, P, P, P, P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "xf_lod_transport": (I, [P, P, P, P, P, P, P, P, P, D, I]),
    "xf_lod_transport_batch": (I, [P, I, P, P, P, P, P, P, P, P, D, I]),
    "xf_thomas_batched": (I, [P, I, I, P, P, P, P, P]),
    "xf_ib_interp": (I, [P, P, I, I, P, P, P]),
    "xf_ib_spread": (I, [P, P, I, I, P, P, P]),
    "xf_ib_update_position": (I, [P, I, I, P, P, P, D]),
    "xf_pressure_rhs": (I, [P, P, P, P, P, P, P, P, P, P, D, P]),
    "xf_velocity_correct": (I, [P, P, I, P, P, P, P, P, D, P]),
    "xf_divergence_max": (I, [P, P, P, P, P, P, P, P, C.POINTER(D), C.POINTER(I64)]),
    "xf_unet_create": (I, [P, I, C.POINTER(P)]),
    "xf_unet_destroy": (I, [P, P]),
    "xf_unet_param_count": (I, [P, C.POINTER(I64)]),
    "xf_unet_get_params": (I, [P, P, P]),
    "xf_unet_set_params": (I, [P, P, P]),
    "xf_unet_get_grads": (I, [P, P, P]),
    "xf_unet_forward": (I, [P, P, I, I, I, I, P, P, P]),
    "xf_unet_loss_grad": (I, [P, P, I, I, I, I, P, P, P, C.POINTER(D)]),
    "xf_unet_train_step": (I, [P, P, I, I, I, I, P, P, P, C.c_float, C.POINTER(D)]),
    "xf_unet_apply": (I, [P, P, P, P, P]),
    "xf_unet_set_loss": (I, [P, P, I]),
    "xf_mg_levels": (I, [P, P, C.POINTER(I)]),
    "xf_mg_apply": (I, [P, P, P, P, P]),
    "xf_mg_level": (I, [P, P, I, C.POINTER(P)]),
    "xf_pack_create": (I, [P, P, C.POINTER(P)]),
    "xf_pack_destroy": (I, [P, P]),
    "xf_pack_info": (I, [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64)]),
    "xf_pack_export": (I, [P, P, I, P, P, P]),
    "xf_pack_apply": (I, [P, P, I, P, P]),
    "xf_pack_gather": (I, [P, P, P, P]),
    "xf_pack_scatter": (I, [P, P, P, P]),
    "xf_pack_solve": (I, [P, P, I, P, P, P, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "xf_rn_lincomb": (I, [P, I64, D, P, D, P, P]),
    "xf_vtk_write_flow": (I, [P, C.c_char_p, C.POINTER(VtkFlow)]),
    "xf_vtk_write_boundary": (I, [P, C.c_char_p, I, I, P, P, P, P, P, D]),
    "xf_npy_save": (I, [P, C.c_char_p, P, C.c_char, I, I, P]),
    "xf_rn_dot": (I, [P, I64, P, P, C.POINTER(D)]),
    "xf_rn_absmax": (I, [P, I64, P, C.POINTER(D), C.POINTER(C.c_int32)]),
    "xf_transport_velocity": (I, [P, P, I, P, P, P, P, P, P, P, P, P, P]),
    "xf_profile_enable": (I, [P, I]),
    "xf_profile_reset": (I, [P]),
    "xf_profile_read": (I, [P, C.c_char_p, C.POINTER(D), C.POINTER(I64)]),
    "xf_stream_ceiling": (I, [P, I, SZ, I, C.POINTER(D)]),
}

_LIB = None


def lib():
    """Loads libxflow_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libxflow_hip.so is not built: run __graft_entry__.build() "
                               "(make -C xflow_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


class XfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%s)" % (msg, STATUS_NAMES.get(code, code)))
        self.code = code
