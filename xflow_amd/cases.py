"""Deterministic synthetic inputs for the pressure-Poisson workloads.

A pressure grid with n^3 cells is stored with a one-cell ring: arrays have
(n+2)^3 entries, indexed [i][j][k] with k contiguous, exactly the reference's
logical layout (XStatement.h:130-135).  Point types follow XMask.h:6-7 and
XOperator.h:13-18: 0 Fictive, 1 Actual, 3 Defined, 4 PreDefined, 5 Normal,
6 PreNormal.  A cell is solid (Fictive) iff its centre lies inside an
obstacle; the ring is Fictive, so the masked 7-point operator has Neumann
walls (XOperator.cpp:1002-1009).

Workloads (sizes are the cell counts per axis):
  * sphere Poisson, 128^3: solid sphere r = 0.2 at the cube centre;
  * random-obstacle Poisson, 256^3: 1-4 spheres or boxes per seed, radius or
    half-width uniform in [0.05, 0.2], centres uniform in [0.25, 0.75]^3;
  * right-hand side: face velocities i.i.d. N(0,1) (numpy PCG64, given seed),
    separable Gaussian smoothing with sigma = 3 cells truncated at 4 sigma and
    mirror boundaries, MAC divergence at the fluid cells, fluid mean removed.
"""
import numpy as np

FICTIVE, ACTUAL, BORDER, DEFINED, PREDEFINED, NORMAL, PRENORMAL = 0, 1, 2, 3, 4, 5, 6


def axis_steps(ncell, length=1.0):
    """Separator steps of a cell-centred axis with a ring: ncell+1 steps of h."""
    h = length / ncell
    return np.full(ncell + 1, h)


def cell_centres(ncell, length=1.0):
    h = length / ncell
    return (np.arange(ncell + 2) - 0.5) * h


def sphere_obstacle(centre, radius):
    return ("sphere", tuple(float(c) for c in centre), float(radius))


def box_obstacle(centre, half):
    return ("box", tuple(float(c) for c in centre), float(half))


def random_obstacles(seed):
    rng = np.random.default_rng(10_000 + seed)
    count = int(rng.integers(1, 5))
    obs = []
    for _ in range(count):
        kind = "sphere" if rng.random() < 0.5 else "box"
        centre = rng.uniform(0.25, 0.75, size=3)
        size = rng.uniform(0.05, 0.2)
        obs.append((kind, tuple(float(c) for c in centre), float(size)))
    return obs


def solid_mask(ncell, obstacles, shape=None, length=(1.0, 1.0, 1.0)):
    """Boolean (n+2)^3 array, True where a cell centre is inside an obstacle."""
    if shape is None:
        shape = (ncell, ncell, ncell)
    xs = [cell_centres(shape[a], length[a]) for a in range(3)]
    solid = np.zeros(tuple(s + 2 for s in shape), dtype=bool)
    for kind, c, r in obstacles:
        # Evaluate on the bounding box only.
        lo, hi = [], []
        for a in range(3):
            h = length[a] / shape[a]
            lo.append(max(0, int(np.floor((c[a] - r) / h))))
            hi.append(min(shape[a] + 2, int(np.ceil((c[a] + r) / h)) + 2))
        X = xs[0][lo[0]:hi[0], None, None]
        Y = xs[1][None, lo[1]:hi[1], None]
        Z = xs[2][None, None, lo[2]:hi[2]]
        if kind == "sphere":
            inside = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 <= r * r
        else:
            inside = (np.abs(X - c[0]) <= r) & (np.abs(Y - c[1]) <= r) & (np.abs(Z - c[2]) <= r)
        solid[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] |= inside
    return solid


def neumann_codes(ncell, obstacles, shape=None, length=(1.0, 1.0, 1.0)):
    """Point types: Actual fluid interior, Fictive ring and solids."""
    if shape is None:
        shape = (ncell, ncell, ncell)
    codes = np.zeros(tuple(s + 2 for s in shape), dtype=np.uint8)
    codes[1:-1, 1:-1, 1:-1] = ACTUAL
    if obstacles:
        codes[solid_mask(ncell, obstacles, shape, length)] = FICTIVE
    codes[0, :, :] = codes[-1, :, :] = 0
    codes[:, 0, :] = codes[:, -1, :] = 0
    codes[:, :, 0] = codes[:, :, -1] = 0
    return codes


def _any_pairs(a, axis):
    """Halve one axis of a ringed boolean array: ring entries stay, interior
    entries 2I-1 and 2I (1-based) combine with 'or'."""
    a = np.moveaxis(a, axis, 0)
    out = np.concatenate([a[:1], a[1:-1:2] | a[2:-1:2], a[-1:]], axis=0)
    return np.moveaxis(out, 0, axis)


def coarsen_codes(codes):
    """Point types of the grid halved per axis, as the multigrid hierarchy of
    the flexible-CG preconditioner forms them: a coarse cell is Actual if any
    of its children (8 inside, fewer on the ring) is, Defined if any child is
    another non-Fictive type, Fictive otherwise."""
    assert all((s_ - 2) % 2 == 0 for s_ in codes.shape)

    def reduce(x):
        for ax in range(3):
            x = _any_pairs(x, ax)
        return x
    ca = reduce(codes == ACTUAL)
    cb = reduce((codes != ACTUAL) & (codes != FICTIVE))
    out = np.zeros(ca.shape, dtype=np.uint8)
    out[cb] = DEFINED
    out[ca] = ACTUAL
    return out


def _gauss_kernel(sigma, truncate=4.0):
    r = int(truncate * sigma + 0.5)
    x = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (x / sigma) ** 2)
    return k / k.sum(), r


def _smooth_axis(a, axis, kern, r):
    """Convolution along one axis with mirror ('reflect', d c b a | a b c d) ends."""
    pad = [(0, 0)] * a.ndim
    pad[axis] = (r, r)
    p = np.pad(a, pad, mode="symmetric")
    out = np.zeros_like(a)
    n = a.shape[axis]
    for t, w in enumerate(kern):
        sl = [slice(None)] * a.ndim
        sl[axis] = slice(t, t + n)
        out += w * p[tuple(sl)]
    return out


def smooth(a, sigma=3.0):
    kern, r = _gauss_kernel(sigma)
    for axis in range(a.ndim):
        a = _smooth_axis(a, axis, kern, r)
    return a


def divergence_rhs(codes, seed, sigma=3.0, norm=None, h=None):
    """MAC divergence of a smoothed random face-velocity field at the fluid
    cells, fluid mean removed; optionally scaled to a given 2-norm."""
    shape = tuple(s - 2 for s in codes.shape)
    if h is None:
        h = [1.0 / s for s in shape]
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    u = smooth(rng.standard_normal((nx + 1, ny, nz)), sigma)
    v = smooth(rng.standard_normal((nx, ny + 1, nz)), sigma)
    w = smooth(rng.standard_normal((nx, ny, nz + 1)), sigma)
    div = (u[1:] - u[:-1]) / h[0] + (v[:, 1:] - v[:, :-1]) / h[1] + (w[:, :, 1:] - w[:, :, :-1]) / h[2]
    b = np.zeros(codes.shape)
    b[1:-1, 1:-1, 1:-1] = div
    fluid = codes == ACTUAL
    b[~fluid] = 0.0
    b[fluid] -= b[fluid].mean()
    if norm is not None:
        b *= norm / np.linalg.norm(b)
    return b


def poisson_case(ncell, obstacles, seed, norm=None):
    codes = neumann_codes(ncell, obstacles)
    b = divergence_rhs(codes, seed, norm=norm)
    h = axis_steps(ncell)
    return {"codes": codes, "hx": h, "hy": h.copy(), "hz": h.copy(), "b": b}


def sphere_case(ncell=128, seed=0, norm=None):
    """Solid sphere r = 0.2 at the cube centre (the 128^3 CG/BiCG/MR workload)."""
    return poisson_case(ncell, [sphere_obstacle((0.5, 0.5, 0.5), 0.2)], seed, norm)


def obstacle_case(ncell=256, seed=0, norm=None):
    """1-4 random spheres/boxes (the 256^3 mixed-precision / learned workloads)."""
    return poisson_case(ncell, random_obstacles(seed), seed, norm)


def icosphere(subdivisions=4, radius=1.0, centre=(0.0, 0.0, 0.0)):
    """Vertices and triangles of an icosphere: the icosahedron with each
    triangle split into four `subdivisions` times, midpoints pushed to the
    sphere (12 -> 42 -> 162 -> 642 -> 2562 vertices)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                v = verts[a] + verts[b]
                verts.append(v / np.linalg.norm(v))
                cache[key] = len(verts) - 1
            return cache[key]
        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.asarray(centre, dtype=np.float64)
    return V, np.array(faces, dtype=np.int64)


def vertex_areas(V, F):
    """Lumped (one third of each adjacent triangle) surface area per vertex."""
    a = np.zeros(len(V))
    tri = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    area = 0.5 * np.linalg.norm(tri, axis=1)
    for q in range(3):
        np.add.at(a, F[:, q], area / 3.0)
    return a
