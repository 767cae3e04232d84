"""xflow_amd — MI355X (gfx950) projection-step hot path of xFlow.

Python face of libxflow_hip.so for tests, benchmarks and drivers.  Device
memory comes from torch (ROCm); every call goes through the C ABI in
include/xflow.h.  The C++ drop-in classes for reference-style host code live
in xflow_amd/host.
"""
import ctypes as C

import numpy as np

from ._lib import (XF_BICG, XF_BICGSTAB, XF_BICGSTAB_T, XF_CG, XF_CG3, XF_CG_MP, XF_FCG_MGNN, XF_FCG_NN, XF_MR, XF_OK, XF_TOL_ABS,
                   XF_TOL_REL, SolveOpts, SolveStats, VtkFlow, XfError, lib)

__all__ = ["Context", "Grid", "Pack", "XfError", "XF_CG3", "XF_BICG", "XF_MR", "XF_BICGSTAB", "XF_CG", "XF_CG_MP",
           "XF_FCG_NN", "XF_FCG_MGNN", "XF_BICGSTAB_T", "XF_TOL_ABS", "XF_TOL_REL", "solve_opts", "write_vtk_flow",
           "write_vtk_boundary", "npy_save"]


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _host(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data


class Context:
    """One xf_ctx (device + stream).  Use one per host thread."""

    def __init__(self, device=0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("xflow_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        torch.cuda.set_device(device)
        self.device = device
        h = C.c_void_p()
        self._check(lib().xf_init(device, C.byref(h)), "xf_init", None)
        self.h = h
        self.torch_device = torch.device("cuda", device)
        # run on torch's current stream so torch allocations and our kernels are ordered
        self.set_stream(torch.cuda.current_stream(device).cuda_stream)

    def _check(self, st, what, h=None):
        if st != XF_OK:
            msg = lib().xf_last_error(h if h is not None else getattr(self, "h", None))
            raise XfError(st, "%s: %s" % (what, msg.decode() if msg else ""))

    def check(self, st, what):
        self._check(st, what, self.h)

    def set_stream(self, stream_handle):
        self.check(lib().xf_set_stream(self.h, C.c_void_p(stream_handle)), "xf_set_stream")

    def synchronize(self):
        self.check(lib().xf_synchronize(self.h), "xf_synchronize")

    def set_deferred_errors(self, on=True):
        """Data checks (the velocity correction's boundary values) recorded on
        the device and raised by check_errors() instead of by each call."""
        self.check(lib().xf_set_deferred_errors(self.h, 1 if on else 0), "xf_set_deferred_errors")

    def check_errors(self):
        self.check(lib().xf_check_errors(self.h), "xf_check_errors")

    def close(self):
        if getattr(self, "h", None):
            lib().xf_shutdown(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Grid:
    """A T3dNormalGrid + space model on the device (xf_grid)."""

    def __init__(self, ctx, hx, hy, hz, codes, normals=None, has_normal=None):
        self.ctx = ctx
        codes = np.asarray(codes)
        self.shape = codes.shape
        n, l, m = self.shape
        hx, phx = _host(hx, np.float64)
        hy, phy = _host(hy, np.float64)
        hz, phz = _host(hz, np.float64)
        assert hx.size == n - 1 and hy.size == l - 1 and hz.size == m - 1, "steps must have n-1, l-1, m-1 entries"
        cd, pcd = _host(codes, np.int32)
        pn = ph = None
        if normals is not None:
            nv, pn = _host(np.asarray(normals).reshape(-1, 3), np.float64)
            assert nv.shape[0] == cd.size
        if has_normal is not None:
            hv, ph = _host(has_normal, np.int32)
            assert hv.size == cd.size
        g = C.c_void_p()
        ctx.check(lib().xf_grid_create(ctx.h, n, l, m, phx, phy, phz, pcd, pn, ph, C.byref(g)), "xf_grid_create")
        self.h = g
        nc, nu = C.c_int64(), C.c_int64()
        lib().xf_grid_info(g, C.byref(nc), C.byref(nu))
        self.ncell, self.n_unknowns = nc.value, nu.value

    def export(self, activity=True):
        sm = np.zeros(self.shape, dtype=np.int32)
        rn = np.zeros(self.shape, dtype=np.int32)
        act = np.zeros((self.n_unknowns, 8), dtype=np.int32) if activity else None
        self.ctx.check(lib().xf_grid_export(self.ctx.h, self.h, sm.ctypes.data, rn.ctypes.data,
                                            act.ctypes.data if activity else None), "xf_grid_export")
        return sm, rn, act

    def set_density(self, rho):
        self.ctx.check(lib().xf_grid_set_density(self.ctx.h, self.h, _ptr(rho)), "xf_grid_set_density")

    def _vec(self, t, dtype=None):
        import torch
        assert t.is_cuda and t.is_contiguous() and t.numel() == self.ncell, "expected a contiguous device field"
        assert t.dtype == (dtype or torch.float64)
        return _ptr(t)

    def apply(self, x, y, b=None, sign_b=0.0):
        self.ctx.check(lib().xf_apply(self.ctx.h, self.h, self._vec(x), self._vec(b) if b is not None else None,
                                      float(sign_b), self._vec(y)), "xf_apply")
        return y

    def apply_dot(self, x):
        out = C.c_double()
        self.ctx.check(lib().xf_apply_dot(self.ctx.h, self.h, self._vec(x), C.byref(out)), "xf_apply_dot")
        return out.value

    def lincomb(self, p1, s1, p2, s2, res):
        self.ctx.check(lib().xf_lincomb(self.ctx.h, self.h, float(p1), self._vec(s1) if s1 is not None else None,
                                        float(p2), self._vec(s2) if s2 is not None else None, self._vec(res)),
                       "xf_lincomb")
        return res

    def copy(self, s, res):
        self.ctx.check(lib().xf_copy(self.ctx.h, self.h, self._vec(s), self._vec(res)), "xf_copy")
        return res

    def dot(self, a, b):
        out = C.c_double()
        self.ctx.check(lib().xf_dot(self.ctx.h, self.h, self._vec(a), self._vec(b), C.byref(out)), "xf_dot")
        return out.value

    def absmax(self, x):
        v, idx = C.c_double(), C.c_int32()
        self.ctx.check(lib().xf_absmax(self.ctx.h, self.h, self._vec(x), C.byref(v), C.byref(idx)), "xf_absmax")
        return v.value, idx.value

    def solve(self, b, x0, x, opts):
        st = SolveStats()
        rc = lib().xf_solve(self.ctx.h, self.h, self._vec(b), self._vec(x0), self._vec(x), C.byref(opts),
                            C.byref(st))
        if rc not in (XF_OK,) and st.status == XF_OK:
            self.ctx.check(rc, "xf_solve")
        return st

    def transport_terms(self, axis, vel, tr, vis=None, bc=None, tvis=None, rp=None):
        """xf_transport_velocity: transport velocities into tr (three fields),
        and optionally the viscosity mean tvis and the boundary right-hand
        side rp, of equation `axis` (0 concentration, 1..3 velocity) on this
        grid's point types."""
        v = lambda t: self._vec(t) if t is not None else None
        self.ctx.check(lib().xf_transport_velocity(self.ctx.h, self.h, int(axis), *[self._vec(x) for x in vel], v(vis),
                                                   v(bc), *[self._vec(t) for t in tr], v(tvis), v(rp)),
                       "xf_transport_velocity")

    def lod_transport(self, f, f_prev, u, v, w, rp, kappa, tau, scheme):
        self.ctx.check(lib().xf_lod_transport(self.ctx.h, self.h, self._vec(f), self._vec(f_prev), self._vec(u),
                                              self._vec(v), self._vec(w), self._vec(rp), self._vec(kappa),
                                              float(tau), int(scheme)), "xf_lod_transport")
        return f

    def ib_interp(self, axis, xyz, vel, U):
        nn = xyz.shape[0]
        self.ctx.check(lib().xf_ib_interp(self.ctx.h, self.h, int(axis), nn, _ptr(xyz), self._vec(vel), _ptr(U)),
                       "xf_ib_interp")
        return U

    def ib_spread(self, axis, xyz, F, force):
        nn = xyz.shape[0]
        self.ctx.check(lib().xf_ib_spread(self.ctx.h, self.h, int(axis), nn, _ptr(xyz), _ptr(F), self._vec(force)),
                       "xf_ib_spread")
        return force

    def pressure_rhs(self, u, v, w, hxU, hyV, hzW, tau, rhs, rho=None, pbc=None):
        hx, phx = _host(hxU, np.float64)
        hy, phy = _host(hyV, np.float64)
        hz, phz = _host(hzW, np.float64)
        self.ctx.check(lib().xf_pressure_rhs(self.ctx.h, self.h, self._vec(u), self._vec(v), self._vec(w),
                                             _ptr(rho), phx, phy, phz, _ptr(pbc), float(tau), self._vec(rhs)),
                       "xf_pressure_rhs")
        return rhs

    def velocity_correct(self, axis, u_star, p, hP, tau, u_out, rho=None, bc=None):
        hp, php = _host(hP, np.float64)
        self.ctx.check(lib().xf_velocity_correct(self.ctx.h, self.h, int(axis), self._vec(u_star), self._vec(p),
                                                 _ptr(rho), php, _ptr(bc), float(tau), self._vec(u_out)),
                       "xf_velocity_correct")
        return u_out

    def divergence_max(self, u, v, w, hxU, hyV, hzW):
        hx, phx = _host(hxU, np.float64)
        hy, phy = _host(hyV, np.float64)
        hz, phz = _host(hzW, np.float64)
        val, idx = C.c_double(), C.c_int64()
        self.ctx.check(lib().xf_divergence_max(self.ctx.h, self.h, self._vec(u), self._vec(v), self._vec(w), phx,
                                               phy, phz, C.byref(val), C.byref(idx)), "xf_divergence_max")
        return val.value, idx.value

    def mg_levels(self):
        """Depth of the multigrid hierarchy of XF_FCG_MGNN (built on first use)."""
        n = C.c_int()
        self.ctx.check(lib().xf_mg_levels(self.ctx.h, self.h, C.byref(n)), "xf_mg_levels")
        return n.value

    def mg_level(self, k):
        """The rediscretised grid of hierarchy level k (0 = this grid), owned by this grid."""
        h = C.c_void_p()
        self.ctx.check(lib().xf_mg_level(self.ctx.h, self.h, int(k), C.byref(h)), "xf_mg_level")
        g = Grid.__new__(Grid)
        g.ctx, g.h, g._owner, g._borrowed = self.ctx, h, self, True
        g.shape = tuple(((s - 2) >> k) + 2 for s in self.shape)
        nc, nu = C.c_int64(), C.c_int64()
        lib().xf_grid_info(h, C.byref(nc), C.byref(nu))
        g.ncell, g.n_unknowns = nc.value, nu.value
        return g

    def mg_apply(self, net, r, z):
        """z = M(r): one V-cycle (FP32 fields): with a network, the cycle down to
        the first grid with rows under 128 cells and the network as its solve;
        with net None, the full-depth cycle and the pseudo-inverse of its deepest grid."""
        import torch
        self.ctx.check(lib().xf_mg_apply(self.ctx.h, self.h, net.h if net is not None else None,
                                         self._vec(r, torch.float32), self._vec(z, torch.float32)), "xf_mg_apply")
        return z

    def close(self):
        if getattr(self, "_borrowed", False):
            self.h = None
            return
        if getattr(self, "h", None) and self.ctx.h:
            lib().xf_grid_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lod_transport_batch(ctx, grids, f, f_prev, u, v, w, rp, kappa, tau, scheme):
    """Grid.lod_transport for several independent problems (lists of the same
    length, 1..4) in one set of launches."""
    n = len(grids)
    arr = lambda xs: (C.c_void_p * n)(*[grids[i]._vec(x) for i, x in enumerate(xs)])
    gh = (C.c_void_p * n)(*[g.h for g in grids])
    ctx.check(lib().xf_lod_transport_batch(ctx.h, n, gh, arr(f), arr(f_prev), arr(u), arr(v), arr(w), arr(rp),
                                           arr(kappa), float(tau), int(scheme)), "xf_lod_transport_batch")
    return f


def thomas_batched(ctx, u, a, b, c, f):
    """Solves u.shape[0] tridiagonal systems of size u.shape[1] in place."""
    nsys, M1 = u.shape
    ctx.check(lib().xf_thomas_batched(ctx.h, nsys, M1 - 1, _ptr(u), _ptr(a), _ptr(b), _ptr(c), _ptr(f)),
              "xf_thomas_batched")
    return u


def ib_update_position(ctx, axis, xyz, U, dt, xyz_prev=None):
    ctx.check(lib().xf_ib_update_position(ctx.h, int(axis), xyz.shape[0], _ptr(xyz), _ptr(xyz_prev), _ptr(U),
                                          float(dt)), "xf_ib_update_position")
    return xyz


def solve_opts(method, **kw):
    o = SolveOpts()
    lib().xf_solve_opts_default(method, C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise AttributeError(k)
        setattr(o, k, v)
    return o


class Pack:
    """T1dPackOperator on the device (xf_pack): packed rows of A and A^T over the
    grid's Rn unknowns and their products with Rn-ordered device vectors."""

    def __init__(self, grid):
        self.ctx, self.grid = grid.ctx, grid
        h = C.c_void_p()
        self.ctx.check(lib().xf_pack_create(self.ctx.h, grid.h, C.byref(h)), "xf_pack_create")
        self.h = h
        n, a, t = C.c_int64(), C.c_int64(), C.c_int64()
        lib().xf_pack_info(h, C.byref(n), C.byref(a), C.byref(t))
        self.n_unknowns, self.nnz = n.value, (a.value, t.value)

    def export(self, transpose=False):
        t = 1 if transpose else 0
        counts = np.zeros(self.n_unknowns, dtype=np.int32)
        vals = np.zeros(max(self.nnz[t], 1))
        cols = np.zeros(max(self.nnz[t], 1), dtype=np.int32)
        self.ctx.check(lib().xf_pack_export(self.ctx.h, self.h, t, counts.ctypes.data, vals.ctypes.data,
                                            cols.ctypes.data), "xf_pack_export")
        return counts, vals[:self.nnz[t]], cols[:self.nnz[t]]

    def apply(self, x, y, transpose=False):
        self.ctx.check(lib().xf_pack_apply(self.ctx.h, self.h, 1 if transpose else 0, _ptr(x), _ptr(y)),
                       "xf_pack_apply")
        return y

    def solve(self, b, x0, x, opts, transpose=False):
        """A reference solver (XF_CG3, XF_BICG, XF_MR, XF_BICGSTAB) on Rn vectors
        (transpose: solve with A^T)."""
        st = SolveStats()
        rc = lib().xf_pack_solve(self.ctx.h, self.h, 1 if transpose else 0, _ptr(b), _ptr(x0), _ptr(x),
                                 C.byref(opts), C.byref(st))
        if rc not in (XF_OK,) and st.status == XF_OK:
            self.ctx.check(rc, "xf_pack_solve")
        return st

    # Rn vector algebra (TRnRelease1dSpace): every entry is an unknown
    def lincomb(self, p1, s1, p2, s2, res):
        self.ctx.check(lib().xf_rn_lincomb(self.ctx.h, self.n_unknowns, float(p1), _ptr(s1), float(p2), _ptr(s2),
                                           _ptr(res)), "xf_rn_lincomb")
        return res

    def dot(self, a, b):
        out = C.c_double()
        self.ctx.check(lib().xf_rn_dot(self.ctx.h, self.n_unknowns, _ptr(a), _ptr(b), C.byref(out)), "xf_rn_dot")
        return out.value

    def absmax(self, x):
        v, idx = C.c_double(), C.c_int32()
        self.ctx.check(lib().xf_rn_absmax(self.ctx.h, self.n_unknowns, _ptr(x), C.byref(v), C.byref(idx)),
                       "xf_rn_absmax")
        return v.value, idx.value

    def gather(self, field, vec):
        self.ctx.check(lib().xf_pack_gather(self.ctx.h, self.h, _ptr(field), _ptr(vec)), "xf_pack_gather")
        return vec

    def scatter(self, vec, field):
        self.ctx.check(lib().xf_pack_scatter(self.ctx.h, self.h, _ptr(vec), _ptr(field)), "xf_pack_scatter")
        return field

    def close(self):
        if getattr(self, "h", None) and self.ctx.h:
            lib().xf_pack_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ output files
def write_vtk_flow(ctx, path, grids, x, y, z, u, v, w, p=None, density=None, concentration=None, viscosity=None):
    """The reference's node-centred VTK of a flow state (xf_vtk_write_flow).
    grids: one Grid or (velocity, pressure, density, concentration) Grids whose
    point types mask the averages; fields are device tensors of the grid shape."""
    if isinstance(grids, Grid):
        grids = (grids,) * 4
    xs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z)]
    f = VtkFlow()
    f.mask_grid_u, f.mask_grid_p, f.mask_grid_d, f.mask_grid_c = (g.h.value for g in grids)
    f.x, f.y, f.z = (a.ctypes.data for a in xs)
    for name, t in (("u", u), ("v", v), ("w", w), ("p", p), ("density", density), ("concentration", concentration),
                    ("viscosity", viscosity)):
        setattr(f, name, None if t is None else grids[0]._vec(t).value)
    ctx.check(lib().xf_vtk_write_flow(ctx.h, str(path).encode(), C.byref(f)), "xf_vtk_write_flow")


def write_vtk_boundary(ctx, path, circles, in_circle, xyz, xyz0, vel, p, wss, time=0.0):
    """The reference's immersed-boundary VTK (xf_vtk_write_boundary); arrays are
    device tensors or host numpy arrays."""
    keep = []

    def ptr(a):
        if hasattr(a, "data_ptr"):
            return C.c_void_p(a.data_ptr())
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return C.c_void_p(a.ctypes.data)
    ctx.check(lib().xf_vtk_write_boundary(ctx.h, str(path).encode(), int(circles), int(in_circle), ptr(xyz), ptr(xyz0),
                                          ptr(vel), ptr(p), ptr(wss), float(time)), "xf_vtk_write_boundary")


def npy_save(ctx, path, a):
    """cnpy::npy_save of a device tensor or host numpy array (xf_npy_save)."""
    if hasattr(a, "data_ptr"):
        import torch
        kinds = {torch.float64: ("f", 8), torch.float32: ("f", 4), torch.int32: ("i", 4), torch.int64: ("i", 8),
                 torch.uint8: ("u", 1)}
        kind, ws = kinds[a.dtype]
        assert a.is_contiguous()
        shape, ptr, keep = tuple(a.shape), C.c_void_p(a.data_ptr()), a
    else:
        keep = np.ascontiguousarray(a)
        kind, ws = keep.dtype.kind, keep.dtype.itemsize
        shape, ptr = keep.shape, C.c_void_p(keep.ctypes.data)
    sh = (C.c_uint32 * max(len(shape), 1))(*shape)
    ctx.check(lib().xf_npy_save(ctx.h, str(path).encode(), ptr, kind.encode(), ws, len(shape), sh), "xf_npy_save")
    del keep
