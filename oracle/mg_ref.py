"""Test infrastructure: a numpy restatement of the V-cycle preconditioner of
XF_FCG_MGNN (xflow_amd/csrc/xf_mg.hip) without the network (the full-depth
cycle with an exact pseudo-inverse solve on the deepest grid), used only to
check the GPU cycle.  Not a reference routine: the reference has no
multigrid; the cycle's parts are pinned by what the mathematics fixes (a
smooth error is carried to the coarse grid and back by restriction and
prolongation, Jacobi sweeps reduce A-norm error) in tests/test_oracle_math.py.

FP32 fields, as on the device; the GPU result agrees to FP32 rounding
(different operation grouping is allowed for, not bit equality).
"""
import numpy as np

FICTIVE, ACTUAL = 0, 1
OMEGA, SWEEPS, COARSE_SWEEPS = np.float32(0.857142857), 4, 8
COARSE_CELLS = 512


def coefficients(h, cnt):
    """Per-node pairs (c-, c+) of the fast stencil: 1/(h[i-1] hh), 1/(h[i] hh), hh = (h[i-1] + h[i]) / 2."""
    c = np.zeros((cnt, 2), dtype=np.float64)
    for i in range(1, cnt - 1):
        hh = (h[i - 1] + h[i]) / 2
        c[i, 0] = 1.0 / (h[i - 1] * hh)
        c[i, 1] = 1.0 / (h[i] * hh)
    return c.astype(np.float32)


class Level:
    def __init__(self, codes, h):
        self.codes = codes
        self.shape = codes.shape
        n, l, m = codes.shape
        self.h = h
        self.c = [coefficients(np.full(s - 1, hh), s) for s, hh in zip(codes.shape, h)]
        self.actual = codes == ACTUAL
        self.unknown = self.actual  # throughput grids: every unknown row is Actual
        nf = codes != FICTIVE
        # open[axis][side]: neighbour on that side is not Fictive (side 0 = minus, 1 = plus)
        self.open = []
        for ax in range(3):
            minus = np.zeros(codes.shape, dtype=bool)
            plus = np.zeros(codes.shape, dtype=bool)
            sl_lo = [slice(None)] * 3
            sl_hi = [slice(None)] * 3
            sl_lo[ax] = slice(1, None)
            sl_hi[ax] = slice(None, -1)
            minus[tuple(sl_lo)] = nf[tuple(sl_hi)]
            plus[tuple(sl_hi)] = nf[tuple(sl_lo)]
            self.open.append((minus, plus))
        self.inner = np.zeros(codes.shape, dtype=bool)
        self.inner[1:-1, 1:-1, 1:-1] = True
        self.act_in = self.actual & self.inner
        d = np.zeros(codes.shape, dtype=np.float32)
        for ax in range(3):
            shp = [1, 1, 1]
            shp[ax] = codes.shape[ax]
            cm = self.c[ax][:, 0].reshape(shp)
            cp = self.c[ax][:, 1].reshape(shp)
            d += np.where(self.open[ax][0], cm, np.float32(0)) + np.where(self.open[ax][1], cp, np.float32(0))
        self.diag = d

    def apply(self, z):
        """A z at Actual interior cells, 0 elsewhere."""
        out = np.zeros_like(z)
        for ax in range(3):
            shp = [1, 1, 1]
            shp[ax] = self.shape[ax]
            cm = self.c[ax][:, 0].reshape(shp)
            cp = self.c[ax][:, 1].reshape(shp)
            zp = np.roll(z, -1, axis=ax)
            zm = np.roll(z, 1, axis=ax)
            t = np.where(self.open[ax][1], cp * (zp - z), np.float32(0)) - \
                np.where(self.open[ax][0], cm * (z - zm), np.float32(0))
            out = out + t
        return np.where(self.act_in, -out, np.float32(0)).astype(np.float32)

    def jacobi(self, r, z=None):
        wd = np.where(self.diag > 0, OMEGA / np.where(self.diag > 0, self.diag, 1), np.float32(0)).astype(np.float32)
        if z is None:
            return np.where(self.act_in, wd * r, np.float32(0)).astype(np.float32)
        return np.where(self.act_in, z + wd * (r - self.apply(z)), np.float32(0)).astype(np.float32)


def coarsen(codes):
    from xflow_amd.cases import coarsen_codes
    return coarsen_codes(codes).astype(np.int32)


def hierarchy(codes, h):
    """Every level down to the first whose grid (ring included) has at most
    COARSE_CELLS cells, or whose interior extents cannot be halved again
    (odd, or a half below 2)."""
    levels = [Level(np.asarray(codes, dtype=np.int32), h)]
    while True:
        n, l, m = (s - 2 for s in levels[-1].shape)
        if n % 2 or l % 2 or m % 2 or n // 2 < 2 or l // 2 < 2 or m // 2 < 2:
            break
        if levels[-1].codes.size <= COARSE_CELLS:
            break
        levels.append(Level(coarsen(levels[-1].codes), tuple(2 * x for x in levels[-1].h)))
    return levels


def dense_operator(L):
    """The level's A on its Actual interior cells as a dense matrix (column by
    column from apply), with the list of those cells."""
    unk = np.flatnonzero(L.act_in.ravel())
    A = np.zeros((unk.size, unk.size))
    for c, q in enumerate(unk):
        e = np.zeros(L.codes.size, dtype=np.float32)
        e[q] = 1
        A[:, c] = L.apply(e.reshape(L.shape)).ravel()[unk]
    return A, unk


def coarse_solve(L, r):
    """z = A^+ r on the deepest grid: the least-squares solution of least
    norm (numpy's SVD pseudo-inverse)."""
    if not hasattr(L, "pinv"):
        A, L.unk = dense_operator(L)
        L.pinv = np.linalg.pinv(A)
    z = np.zeros(L.codes.size)
    z[L.unk] = L.pinv @ np.asarray(r, dtype=np.float64).ravel()[L.unk]
    return z.reshape(L.shape).astype(np.float32)


def restrict(fine, t, coarse):
    n, l, m = (s - 2 for s in fine.shape)
    tu = np.where(fine.unknown, t, np.float32(0))[1:-1, 1:-1, 1:-1].reshape(n // 2, 2, l // 2, 2, m // 2, 2)
    cnt = fine.unknown[1:-1, 1:-1, 1:-1].reshape(n // 2, 2, l // 2, 2, m // 2, 2).sum(axis=(1, 3, 5))
    s = tu.sum(axis=(1, 3, 5), dtype=np.float32)
    rc = np.zeros(coarse.shape, dtype=np.float32)
    rc[1:-1, 1:-1, 1:-1] = np.where(cnt > 0, s / np.maximum(cnt, 1).astype(np.float32), np.float32(0))
    return rc


def prolong(fine, zc):
    inner = zc[1:-1, 1:-1, 1:-1]
    up = inner.repeat(2, 0).repeat(2, 1).repeat(2, 2)
    out = np.zeros(fine.shape, dtype=np.float32)
    out[1:-1, 1:-1, 1:-1] = up
    return np.where(fine.unknown & fine.inner, out, np.float32(0))


def vcycle(levels, k, r):
    L = levels[k]
    if k == len(levels) - 1:
        if L.codes.size <= COARSE_CELLS:
            return coarse_solve(L, r)
        z = L.jacobi(r)
        for _ in range(COARSE_SWEEPS - 1):
            z = L.jacobi(r, z)
        return z
    z = L.jacobi(r)
    for _ in range(SWEEPS - 1):
        z = L.jacobi(r, z)
    t = np.where(L.act_in, r - L.apply(z), np.float32(0)).astype(np.float32)
    rc = restrict(L, t, levels[k + 1])
    zc = vcycle(levels, k + 1, rc)
    z = (z + prolong(L, zc)).astype(np.float32)
    for _ in range(SWEEPS):
        z = L.jacobi(r, z)
    return z


def apply(codes, h, r):
    """z = M(r) for an FP32 residual r on the grid with point types `codes` and uniform steps h = (hx, hy, hz)."""
    levels = hierarchy(codes, h)
    return vcycle(levels, 0, np.asarray(r, dtype=np.float32)), len(levels)
