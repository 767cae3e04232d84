"""ctypes binding of liboracle.so, the CPU oracle — TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker or the timed CPU baseline.  The product path
(xflow_amd / libxflow_hip.so) never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OK, EINVAL, EMAXIT, EBREAKDOWN, ENAN = 0, -1, -2, -3, -4


class XoGrid(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("l", C.c_int), ("m", C.c_int),
        ("hx", C.c_void_p), ("hy", C.c_void_p), ("hz", C.c_void_p),
        ("codes", C.c_void_p), ("space_mask", C.c_void_p),
        ("normals", C.c_void_p), ("has_normal", C.c_void_p), ("rho", C.c_void_p),
    ]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        d, i, p, i64 = C.c_double, C.c_int, C.c_void_p, C.c_int64
        gp = C.POINTER(XoGrid)
        sigs = {
            "xo_set_threads": (None, [i]),
            "xo_space_mask": (i, [i64, p, p]),
            "xo_rn_number": (i64, [i, i, i, p, p]),
            "xo_activity": (i, [i, i, i, p, p, p]),
            "xo_apply": (i, [gp, p, p, d, p]),
            "xo_apply_dot": (d, [gp, p, p]),
            "xo_lincomb": (None, [gp, d, p, d, p, p]),
            "xo_copy": (None, [gp, p, p]),
            "xo_dot": (d, [gp, p, p]),
            "xo_absmax": (d, [gp, p, p, p]),
            "xo_cg3": (i, [gp, p, p, p, d, i, p, p]),
            "xo_bicg": (i, [gp, p, p, p, d, i, p, p]),
            "xo_mr": (i, [gp, p, p, p, d, i, p, p]),
            "xo_bicgstab": (i, [gp, p, p, p, d, i, p, p]),
            "xo_cg_hs": (i, [gp, p, p, p, d, i, i, p, p]),
            "xo_cg_cg": (i, [gp, p, p, p, d, i, i, p, p]),
            "xo_pack": (C.c_int64, [gp, i, p, p, p]),
            "xo_pack_apply": (None, [C.c_int64, p, p, p, p, p]),
            "xo_pack_solve": (i, [i, C.c_int64, p, p, p, p, p, p, p, p, p, d, i, p, p]),
            "xo_thomas": (None, [i, p, p, p, p, p]),
            "xo_lod": (i, [gp, p, p, p, p, p, p, p, d, i]),
            "xo_ib_delta": (d, [d, d]),
            "xo_ib_index": (i, [d, i, i, d, i]),
            "xo_ib_interp": (None, [gp, i, i, p, p, p]),
            "xo_ib_spread": (None, [gp, i, i, p, p, p]),
            "xo_pressure_rhs": (i, [i, i, i, p, p, p, p, p, p, p, p, p, d, p]),
            "xo_velocity_correct": (i, [i, i, i, i, p, p, p, p, p, p, d, p]),
            "xo_divergence_max": (d, [i, i, i, p, p, p, p, p, p, p, p]),
            "xo_transport_velocity": (i, [i, i, i, i, p, p, p, p, p, p, p, p, p, p, p, p, p, p, p]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data


class Grid:
    """Holds contiguous copies of a case's arrays and the XoGrid view of them."""

    def __init__(self, hx, hy, hz, codes, space_mask=None, normals=None, has_normal=None, rho=None):
        self.hx = np.ascontiguousarray(hx, dtype=np.float64)
        self.hy = np.ascontiguousarray(hy, dtype=np.float64)
        self.hz = np.ascontiguousarray(hz, dtype=np.float64)
        self.codes = np.ascontiguousarray(codes, dtype=np.int32)
        self.shape = self.codes.shape
        if space_mask is None:
            space_mask = space_mask_of(self.codes)
        self.space_mask = np.ascontiguousarray(space_mask, dtype=np.int32)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float64)
        self.has_normal = None if has_normal is None else np.ascontiguousarray(has_normal, dtype=np.int32)
        self.rho = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64)
        n, l, m = self.shape
        self.s = XoGrid(n, l, m, _ptr(self.hx), _ptr(self.hy), _ptr(self.hz), _ptr(self.codes),
                        _ptr(self.space_mask), _ptr(self.normals), _ptr(self.has_normal), _ptr(self.rho))

    @property
    def ref(self):
        return C.byref(self.s)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n):
    lib().xo_set_threads(int(n))


def space_mask_of(codes):
    codes = np.ascontiguousarray(codes, dtype=np.int32)
    sm = np.zeros_like(codes)
    st = lib().xo_space_mask(codes.size, codes.ctypes.data, sm.ctypes.data)
    if st != OK:
        raise ValueError("unknown point type in mask")
    return sm


def rn_number(space_mask):
    sm = np.ascontiguousarray(space_mask, dtype=np.int32)
    rn = np.zeros_like(sm)
    n, l, m = sm.shape
    cnt = lib().xo_rn_number(n, l, m, sm.ctypes.data, rn.ctypes.data)
    if cnt < 0:
        raise ValueError("space mask must be 0/1")
    return rn, int(cnt)


def activity(space_mask, rn, count):
    sm = np.ascontiguousarray(space_mask, dtype=np.int32)
    rn = np.ascontiguousarray(rn, dtype=np.int32)
    act = np.zeros((count, 8), dtype=np.int32)
    n, l, m = sm.shape
    lib().xo_activity(n, l, m, sm.ctypes.data, rn.ctypes.data, act.ctypes.data)
    return act


def apply(g, x, b=None, sign_b=0.0, y=None):
    x = _f64(x)
    y = np.zeros_like(x) if y is None else _f64(y).copy()
    bb = None if b is None else _f64(b)
    st = lib().xo_apply(g.ref, x.ctypes.data, _ptr(bb), float(sign_b), y.ctypes.data)
    if st != OK:
        raise ValueError("operator undefined for a row of this mask")
    return y


def apply_dot(g, x):
    x = _f64(x)
    st = C.c_int(0)
    v = lib().xo_apply_dot(g.ref, x.ctypes.data, C.byref(st))
    if st.value != OK:
        raise ValueError("operator undefined for a row of this mask")
    return v


def dot(g, a, b):
    a, b = _f64(a), _f64(b)
    return lib().xo_dot(g.ref, a.ctypes.data, b.ctypes.data)


def absmax(g, x, rn):
    x = _f64(x)
    rn = np.ascontiguousarray(rn, dtype=np.int32)
    idx = C.c_int32(0)
    v = lib().xo_absmax(g.ref, x.ctypes.data, rn.ctypes.data, C.byref(idx))
    return v, idx.value


def _solve(fn, g, b, x0, eps, maxit, *extra):
    b, x0 = _f64(b), _f64(x0)
    x = np.zeros_like(x0)
    it = C.c_int(0)
    res = C.c_double(0)
    st = fn(g.ref, b.ctypes.data, x0.ctypes.data, x.ctypes.data, float(eps), int(maxit), *extra,
            C.byref(it), C.byref(res))
    return x, it.value, res.value, st


def cg3(g, b, x0, eps=1e-5, maxit=100000):
    return _solve(lib().xo_cg3, g, b, x0, eps, maxit)


def bicg(g, b, x0, eps=1e-8, maxit=100000):
    return _solve(lib().xo_bicg, g, b, x0, eps, maxit)


def mr(g, b, x0, eps=1e-7, maxit=1000000):
    return _solve(lib().xo_mr, g, b, x0, eps, maxit)


def bicgstab(g, b, x0, eps=1e-5, maxit=100000):
    return _solve(lib().xo_bicgstab, g, b, x0, eps, maxit)


def cg_hs(g, b, x0, rtol=1e-8, maxit=100000, project_mean=False):
    return _solve(lib().xo_cg_hs, g, b, x0, rtol, maxit, int(bool(project_mean)))


def pack(g, transpose=False):
    """T1dPackOperator rows: (counts[nunk], vals[nnz], cols[nnz], 0-based columns)."""
    nunk = int((np.asarray(g.space_mask) > 0).sum())
    counts = np.zeros(nunk, dtype=np.int32)
    vals = np.zeros(7 * max(nunk, 1))
    cols = np.zeros(7 * max(nunk, 1), dtype=np.int32)
    nnz = lib().xo_pack(g.ref, int(bool(transpose)), counts.ctypes.data, vals.ctypes.data, cols.ctypes.data)
    if nnz < 0:
        raise ValueError("xo_pack: a row throws in the reference or is empty")
    return counts, vals[:nnz].copy(), cols[:nnz].copy()


def pack_apply(counts, vals, cols, x):
    counts = np.ascontiguousarray(counts, dtype=np.int32)
    vals, x = _f64(vals), _f64(x)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    y = np.zeros(counts.size)
    lib().xo_pack_apply(counts.size, counts.ctypes.data, vals.ctypes.data, cols.ctypes.data, x.ctypes.data,
                        y.ctypes.data)
    return y


PACK_METHODS = {"cg3": 0, "bicg": 1, "mr": 2, "bicgstab": 3}
PACK_EPS = {"cg3": 1e-5, "bicg": 1e-8, "mr": 1e-7, "bicgstab": 1e-5}  # XSLAE.cpp:64, 166, 358, 669


def pack_solve(method, rows, rows_t, b, x0, eps=None, maxit=100000):
    """A reference solver on packed rows (rows, rows_t = pack(g), pack(g, True))
    over Rn vectors; returns (x, iterations, final residual norm, status)."""
    c, v, k = (np.ascontiguousarray(a) for a in rows)
    ct, vt, kt = (np.ascontiguousarray(a) for a in rows_t)
    c, ct, k, kt = (a.astype(np.int32, copy=False) for a in (c, ct, k, kt))
    v, vt, b, x0 = _f64(v), _f64(vt), _f64(b), _f64(x0)
    x = np.zeros(c.size)
    it, res = C.c_int(0), C.c_double(0)
    eps = PACK_EPS[method] if eps is None else eps
    st = lib().xo_pack_solve(PACK_METHODS[method], c.size, c.ctypes.data, v.ctypes.data, k.ctypes.data,
                             ct.ctypes.data, vt.ctypes.data, kt.ctypes.data, b.ctypes.data, x0.ctypes.data,
                             x.ctypes.data, eps, maxit, C.byref(it), C.byref(res))
    return x, it.value, res.value, st


def cg_cg(g, b, x0, rtol=1e-8, maxit=100000, project_mean=False):
    """Single-reduction (Chronopoulos-Gear) CG: the throughput solver's recurrence."""
    return _solve(lib().xo_cg_cg, g, b, x0, rtol, maxit, int(bool(project_mean)))


def thomas(u, a, b, c, f):
    u = _f64(u).copy()
    a, b, c, f = _f64(a), _f64(b), _f64(c), _f64(f)
    lib().xo_thomas(len(u) - 1, u.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data, f.ctypes.data)
    return u


def lod(g, f, f_prev, u, v, w, rp, kappa, tau, scheme):
    f = _f64(f).copy()
    arrs = [_f64(a) for a in (f_prev, u, v, w, rp, kappa)]
    st = lib().xo_lod(g.ref, f.ctypes.data, *[a.ctypes.data for a in arrs], float(tau), int(scheme))
    if st != OK:
        raise ValueError("transport step rejected the mask (sector borders)")
    return f


def ib_delta(r, h):
    return lib().xo_ib_delta(float(r), float(h))


def ib_index(coord, typ, axis, h, end_index):
    return lib().xo_ib_index(float(coord), int(typ), int(axis), float(h), int(end_index))


def ib_interp(g, axis, xyz, vel):
    xyz = _f64(xyz)
    vel = _f64(vel)
    U = np.zeros(xyz.shape[0])
    lib().xo_ib_interp(g.ref, int(axis), xyz.shape[0], xyz.ctypes.data, vel.ctypes.data, U.ctypes.data)
    return U


def ib_spread(g, axis, xyz, F):
    xyz, F = _f64(xyz), _f64(F)
    out = np.zeros(g.shape)
    lib().xo_ib_spread(g.ref, int(axis), xyz.shape[0], xyz.ctypes.data, F.ctypes.data, out.ctypes.data)
    return out


def pressure_rhs(codes, u, v, w, hxU, hyV, hzW, tau, rho=None, pbc=None):
    codes = np.ascontiguousarray(codes, dtype=np.int32)
    n, l, m = codes.shape
    u, v, w, hxU, hyV, hzW = [_f64(a) for a in (u, v, w, hxU, hyV, hzW)]
    rho = None if rho is None else _f64(rho)
    pbc = None if pbc is None else _f64(pbc)
    rhs = np.zeros(codes.shape)
    st = lib().xo_pressure_rhs(n, l, m, codes.ctypes.data, u.ctypes.data, v.ctypes.data, w.ctypes.data, _ptr(rho),
                               hxU.ctypes.data, hyV.ctypes.data, hzW.ctypes.data, _ptr(pbc), float(tau),
                               rhs.ctypes.data)
    if st != OK:
        raise ValueError("pressure rhs: unsupported point type")
    return rhs


def velocity_correct(codes, axis, us, p, hP, tau, rho=None, bc=None, out=None):
    codes = np.ascontiguousarray(codes, dtype=np.int32)
    n, l, m = codes.shape
    us, p, hP = _f64(us), _f64(p), _f64(hP)
    rho = None if rho is None else _f64(rho)
    bc = None if bc is None else _f64(bc)
    out = np.zeros(codes.shape) if out is None else _f64(out).copy()
    st = lib().xo_velocity_correct(n, l, m, int(axis), codes.ctypes.data, us.ctypes.data, p.ctypes.data, _ptr(rho),
                                   hP.ctypes.data, _ptr(bc), float(tau), out.ctypes.data)
    if st != OK:
        raise ValueError("velocity correction: border check failed")
    return out


def divergence_max(codes, u, v, w, hxU, hyV, hzW):
    codes = np.ascontiguousarray(codes, dtype=np.int32)
    n, l, m = codes.shape
    arrs = [_f64(a) for a in (u, v, w, hxU, hyV, hzW)]
    cell = C.c_int64(0)
    val = lib().xo_divergence_max(n, l, m, codes.ctypes.data, *[a.ctypes.data for a in arrs], C.byref(cell))
    return val, cell.value


def transport_velocity(axis, codes, u, v, w, vis=None, bc=None, h=None, normals=None):
    """xo_transport_velocity: the transport velocities (tu, tv, tw), viscosity
    tvis and boundary right-hand side rp of one transport equation, as
    TVPProblem::SolveBySplitting builds them (XProblem.cpp:5090-5303 and
    5357-5417 for velocity component axis 1..3 on its grid, 5889-5944 for the
    concentration, axis 0).  h = (hx, hy, hz) separator steps (needed with vis).
    Returns (tu, tv, tw, tvis, rp); raises ValueError where the reference throws."""
    codes = np.ascontiguousarray(codes, dtype=np.int32)
    n, l, m = codes.shape
    f = [_f64(a) for a in (u, v, w)]
    out = [np.zeros(codes.shape) for _ in range(5)]
    vis_a = _f64(vis) if vis is not None else None
    bc_a = _f64(bc) if bc is not None else None
    if h is None:
        h = [np.ones(n - 1), np.ones(l - 1), np.ones(m - 1)]
    hs = [_f64(x) for x in h]
    nrm = _f64(normals) if normals is not None else None
    ptr = lambda a: a.ctypes.data if a is not None else None
    st = lib().xo_transport_velocity(n, l, m, int(axis), codes.ctypes.data, ptr(nrm), *[x.ctypes.data for x in f],
                                     ptr(vis_a), ptr(bc_a), *[x.ctypes.data for x in hs],
                                     *[x.ctypes.data for x in out])
    if st != OK:
        raise ValueError("transport velocities: a point type or position the reference rejects")
    return tuple(out)


# ------------------------------------------------------------ output files
# Restatements of the reference's writers (test checkers, pure Python/numpy).

def npy_bytes(a):
    """cnpy::npy_save's bytes for a C-order array (cnpy.h:207-236 header, then
    the raw little-endian data)."""
    a = np.ascontiguousarray(a)
    shape = ", ".join(str(d) for d in a.shape) + ("," if a.ndim == 1 else "")
    d = "{'descr': '<%s%d', 'fortran_order': False, 'shape': (%s), }" % (a.dtype.kind, a.dtype.itemsize, shape)
    rem = 16 - (10 + len(d)) % 16
    d = d + " " * rem
    d = d[:-1] + "\n"
    return b"\x93NUMPY\x01\x00" + len(d).to_bytes(2, "little") + d.encode() + a.tobytes()


def vtk_boundary_text(circles, in_circle, xyz, xyz0, vel, p, wss):
    """TImmersedBoundaryProblem::OutputBoundary's file (XProblem.cpp:9568-9632)."""
    nn = circles * in_circle
    out = ["# vtk DataFile Version 1.0\nBound\nASCII\n\n", "DATASET UNSTRUCTURED_GRID\n", "POINTS %d double\n" % nn]
    out += ["%f %f %f\n" % tuple(xyz[q]) for q in range(nn)]
    ncell = in_circle * (circles - 1)
    out.append("CELLS %d %d\n" % (ncell, 5 * ncell))
    for cr in range(circles - 1):
        for i in range(in_circle):
            i1 = cr * in_circle + i
            i2, i3, i4 = i1 + 1, i1 + 1 + in_circle, i1 + in_circle
            if i == in_circle - 1:
                i2, i3 = cr * in_circle, (cr + 1) * in_circle
            out.append("%d %d %d %d %d\n" % (4, i1, i2, i3, i4))
    out.append("CELL_TYPES %d\n" % ncell)
    out += ["9\n"] * ncell
    out.append("POINT_DATA %d\n" % nn)
    out.append("VECTORS Velocity double\n")
    out += ["%f %f %f\n" % tuple(vel[q]) for q in range(nn)]
    out.append("VECTORS Displacement double\n")
    d = np.asarray(xyz0, dtype=np.float64) - np.asarray(xyz, dtype=np.float64)
    out += ["%f %f %f\n" % tuple(d[q]) for q in range(nn)]
    out.append("\n\nSCALARS Pressure float\nLOOKUP_TABLE default\n")
    out += ["%f\n" % v for v in p]
    out.append("\n\nSCALARS WSS float\nLOOKUP_TABLE default\n")
    out += ["%f\n" % v for v in wss]
    return "".join(out)


def vtk_node_average8(mask, f):
    """Nodes 1..n-2 (i fastest): 0 where mask is Fictive, else the sum of the
    eight cells (i..i+1, j..j+1, k..k+1) in the reference's order over the
    count of non-Fictive ones (XProblem.cpp:6208-6239)."""
    mask = np.asarray(mask)
    f = np.asarray(f, dtype=np.float64)
    s = (slice(1, -1),) * 3
    up = lambda di, dj, dk: f[1 + di:f.shape[0] - 1 + di, 1 + dj:f.shape[1] - 1 + dj, 1 + dk:f.shape[2] - 1 + dk]
    mu = lambda di, dj, dk: mask[1 + di:mask.shape[0] - 1 + di, 1 + dj:mask.shape[1] - 1 + dj,
                                 1 + dk:mask.shape[2] - 1 + dk] != 0
    order = [(1, 0, 0), (1, 1, 0), (1, 0, 1), (1, 1, 1), (0, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1)]
    tot = up(*order[0]).copy()
    for o in order[1:]:
        tot = tot + up(*o)
    avr = sum(mu(*o).astype(np.int64) for o in order)
    own = mask[s] != 0
    val = np.where(own, tot / np.where(avr > 0, avr, 1), 0.0)
    return np.transpose(val, (2, 1, 0)).reshape(-1)  # file order: k, j, i (i fastest)


def vtk_flow_text(codes, x, y, z, u, v, w, p=None, density=None, concentration=None, viscosity=None):
    """TVPProblem::SolveBySplitting's VTK (XProblem.cpp:6142-6379); codes: one
    array or (velocity, pressure, density, concentration) point types."""
    if not isinstance(codes, (tuple, list)):
        codes = (codes,) * 4
    cu, cp, cd, cc = (np.asarray(c) for c in codes)
    n, l, m = cu.shape
    nodes = (n - 2) * (l - 2) * (m - 2)
    zero = np.zeros(cu.shape)
    p, density, concentration, viscosity = (zero if a is None else np.asarray(a, dtype=np.float64)
                                            for a in (p, density, concentration, viscosity))
    out = ["# vtk DataFile Version 3.0\n", "Flow\nASCII\n\n", "DATASET STRUCTURED_GRID\n",
           "DIMENSIONS %d %d %d\n" % (n - 2, l - 2, m - 2), "POINTS %d double\n" % nodes]
    for k in range(1, m - 1):
        for j in range(1, l - 1):
            for i in range(1, n - 1):
                out.append("%f  %f  %f\n" % (x[i], y[j], z[k]))
    out.append("POINT_DATA %d\n\nVECTORS Velocity double\n" % nodes)
    u, v, w = (np.asarray(a, dtype=np.float64) for a in (u, v, w))
    I, J, K = slice(1, n - 1), slice(1, l - 1), slice(1, m - 1)
    Ip, Jp, Kp = slice(2, n), slice(2, l), slice(2, m)
    uu = (u[I, J, K] + u[I, Jp, K] + u[I, J, Kp] + u[I, Jp, Kp]) / 4   # XProblem.cpp:6185-6187
    vv = (v[I, J, K] + v[Ip, J, K] + v[I, J, Kp] + v[Ip, J, Kp]) / 4
    ww = (w[I, J, K] + w[Ip, J, K] + w[I, Jp, K] + w[Ip, Jp, K]) / 4
    own = cu[I, J, K] != 0
    vel = [np.transpose(np.where(own, a, 0.0), (2, 1, 0)).reshape(-1) for a in (uu, vv, ww)]
    out += ["%f %f %f\n " % (a, b, c) for a, b, c in zip(*vel)]
    for head, fmt, mask, f in (("\nSCALARS Pressure float\nLOOKUP_TABLE default\n", "%f\n ", cp, p),
                               ("\nSCALARS Density float\nLOOKUP_TABLE default\n", "%f\n", cd, density),
                               ("\nSCALARS Concentration float\nLOOKUP_TABLE default\n", "%f\n", cc, concentration),
                               ("\nSCALARS Viscosity float\nLOOKUP_TABLE default\n", "%f\n", cc, viscosity)):
        out.append(head)
        out += [fmt % val for val in vtk_node_average8(mask, f)]
    return "".join(out)
