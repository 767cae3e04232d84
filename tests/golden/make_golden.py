"""Writes the golden fixtures in tests/golden/*.npz from the compiled reference.

Run in a container that has /root/reference (the recipe in oracle/Makefile
compiles it into oracle/_ref/ref_harness):

    make -C oracle ref && python tests/golden/make_golden.py

Each fixture holds the inputs given to the reference and the outputs it
produced.  Inputs are synthetic and seeded; nothing from the reference tree is
stored except what its code computed.
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
from xfb import read_xfb, write_xfb  # noqa: E402
from xflow_amd import cases  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def space_mask_rule(codes):
    """XProblem.cpp:4155-4182 (restated; checked bit-exactly by the oracle tests)."""
    sm = np.zeros(codes.shape, dtype=np.int32)
    sm[(codes == 1) | (codes == 4) | (codes == 5) | (codes == 6) | (codes == 8)] = 1
    return sm


def run(cmd, inputs):
    with tempfile.TemporaryDirectory() as td:
        fi, fo = os.path.join(td, "in.xfb"), os.path.join(td, "out.xfb")
        write_xfb(fi, inputs)
        p = subprocess.run([HARNESS, cmd, fi, fo], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
        if p.returncode != 0:
            raise RuntimeError("harness %s failed: %s" % (cmd, p.stderr))
        return read_xfb(fo)


def coords_from_steps(h, start=0.0):
    x = np.zeros(len(h) + 1)
    x[0] = start
    x[1:] = start + np.cumsum(h)
    return x


def save(name, inputs, outputs):
    d = {"in_" + k: np.asarray(v) for k, v in inputs.items()}
    d.update({"out_" + k: np.asarray(v) for k, v in outputs.items()})
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
    print("wrote", name, {k: v.shape for k, v in outputs.items()})


def geometry(x, y, z):
    """Separator steps exactly as TRandomSeparator forms them (XSeparator.cpp:81-99)."""
    return np.diff(x), np.diff(y), np.diff(z)


def mixed_border_mask(n, l, m, rng):
    """Actual interior with a Fictive block, every face a different border type
    with its outward axis-aligned normal, edges and corners Fictive."""
    codes = np.zeros((n, l, m), dtype=np.int32)
    codes[1:-1, 1:-1, 1:-1] = 1
    codes[n // 2 - 1:n // 2 + 1, l // 2 - 1:l // 2 + 2, m // 2:m // 2 + 2] = 0
    normals = np.zeros((n, l, m, 3))
    has = np.zeros((n, l, m), dtype=np.int32)
    faces = [  # (axis, side, code, normal sign)
        (0, 0, 4, -1.0), (0, n - 1, 6, 1.0),
        (1, 0, 3, -1.0), (1, l - 1, 6, 1.0),
        (2, 0, 0, -1.0), (2, m - 1, 4, 1.0),
    ]
    for axis, side, code, sgn in faces:
        sl = [slice(1, -1)] * 3
        sl[axis] = side
        codes[tuple(sl)] = code
        nv = np.zeros(3)
        nv[axis] = sgn
        normals[tuple(sl)] = nv
        has[tuple(sl)] = 1 if code in (4, 6) else 0
    return codes, normals.reshape(-1, 3), has


def case_op():
    rng = np.random.default_rng(1)
    n, l, m = 12, 10, 9
    x = coords_from_steps(rng.uniform(0.05, 0.15, n - 1))
    y = coords_from_steps(rng.uniform(0.05, 0.15, l - 1), 0.3)
    z = coords_from_steps(rng.uniform(0.05, 0.15, m - 1), -0.2)
    codes, normals, has = mixed_border_mask(n, l, m, rng)
    sm = space_mask_rule(codes)
    inp = {"x": x, "y": y, "z": z, "codes": codes, "normals": normals, "has_normal": has,
           "space_mask": sm, "xvec": rng.standard_normal((n, l, m)),
           "bvec": rng.standard_normal((n, l, m)), "yinit": np.zeros((n, l, m))}
    save("op_mixed", inp, run("op", inp))
    # uniform sphere-obstacle operator (the Poisson workloads' operator, small)
    nc = 16
    codes = cases.neumann_codes(nc, [cases.sphere_obstacle((0.5, 0.5, 0.5), 0.2)]).astype(np.int32)
    xx = cases.cell_centres(nc)
    inp = {"x": xx, "y": xx.copy(), "z": xx.copy(), "codes": codes, "space_mask": space_mask_rule(codes),
           "xvec": rng.standard_normal(codes.shape), "bvec": rng.standard_normal(codes.shape),
           "yinit": np.zeros(codes.shape)}
    save("op_sphere16", inp, run("op", inp))


def case_pack():
    """T1dPackOperator of the Laplace operator on the mixed-border mask (PreDefined
    and PreNormal rows make it non-symmetric) and on the uniform sphere mask."""
    rng = np.random.default_rng(6)
    n, l, m = 12, 10, 9
    x = coords_from_steps(rng.uniform(0.05, 0.15, n - 1))
    y = coords_from_steps(rng.uniform(0.05, 0.15, l - 1), 0.3)
    z = coords_from_steps(rng.uniform(0.05, 0.15, m - 1), -0.2)
    codes, normals, has = mixed_border_mask(n, l, m, rng)
    inp = {"x": x, "y": y, "z": z, "codes": codes, "normals": normals, "has_normal": has,
           "space_mask": space_mask_rule(codes), "xvec": rng.standard_normal((n, l, m))}
    save("pack_mixed", inp, run("pack", inp))
    nc = 10
    codes = cases.neumann_codes(nc, [cases.sphere_obstacle((0.5, 0.5, 0.5), 0.25)]).astype(np.int32)
    xx = cases.cell_centres(nc)
    inp = {"x": xx, "y": xx.copy(), "z": xx.copy(), "codes": codes, "space_mask": space_mask_rule(codes),
           "xvec": rng.standard_normal(codes.shape)}
    save("pack_sphere10", inp, run("pack", inp))


def case_pack_solve():
    """BiCG, BiCGStab and minimal residuals on the packed operator of the
    mixed-border mask: BiCG's shadow sequence uses the transposed rows."""
    rng = np.random.default_rng(8)
    n, l, m = 12, 10, 9
    x = coords_from_steps(rng.uniform(0.05, 0.15, n - 1))
    y = coords_from_steps(rng.uniform(0.05, 0.15, l - 1), 0.3)
    z = coords_from_steps(rng.uniform(0.05, 0.15, m - 1), -0.2)
    codes, normals, has = mixed_border_mask(n, l, m, rng)
    sm = space_mask_rule(codes)
    for solver, norm in (("pack_bicg", 1.0), ("pack_bicgstab", 1e3), ("pack_mr", 1e-2)):
        b = rng.standard_normal((n, l, m)) * (sm > 0)
        b *= norm / np.linalg.norm(b)
        inp = {"x": x, "y": y, "z": z, "codes": codes, "normals": normals, "has_normal": has, "space_mask": sm,
               "x0": np.zeros((n, l, m)), "bvec": b}
        save(solver + "_mixed", inp, run(solver, inp))


def case_pack_solve_spd():
    """BiCGStab and minimal residuals on the packed operator of the SPD
    Dirichlet problems the grid-solver fixtures use (same inputs): smooth
    convergence, so iteration counts are stable under rounding (on the
    non-symmetric mask a 1e-15 relative change of b moves them by tens)."""
    for name, solver, nc, norm in (("pack_bicgstab_dirichlet12", "pack_bicgstab", 12, 1e8),
                                   ("pack_mr_dirichlet8", "pack_mr", 8, 10.0)):
        codes = dirichlet_codes(nc)
        b = cases.divergence_rhs(codes, seed=7, sigma=1.5)
        b[codes == 1] += 0.3
        b *= norm / np.linalg.norm(b)
        xx = cases.cell_centres(nc)
        inp = {"x": xx, "y": xx.copy(), "z": xx.copy(), "codes": codes, "space_mask": space_mask_rule(codes),
               "x0": np.zeros(codes.shape), "bvec": b}
        save(name, inp, run(solver, inp))


def case_io():
    """cnpy::npy_save and TImmersedBoundaryProblem::OutputBoundary: the files'
    bytes (stored as uint8)."""
    rng = np.random.default_rng(9)
    for name, a in (("npy_f8_3d", rng.standard_normal((3, 4, 5))), ("npy_f8_1d", rng.standard_normal(7) * 1e3)):
        out = run("npy", {"data": a})
        save(name, {"data": a}, {"bytes": np.asarray(out["bytes"], dtype=np.uint8)})
    circles, per = 4, 6
    nn = circles * per
    xyz = rng.uniform(-1, 1, (nn, 3))
    inp = {"xyz": xyz, "xyz0": xyz + 1e-3 * rng.standard_normal((nn, 3)), "vel": rng.standard_normal((nn, 3)),
           "p": rng.standard_normal(nn) * 10, "wss": np.abs(rng.standard_normal(nn)),
           "circles": np.array([circles], dtype=np.int32), "in_circle": np.array([per], dtype=np.int32),
           "step": np.array([12], dtype=np.int32), "dt": np.array([0.005])}
    out = run("vtk_boundary", inp)
    save("vtk_boundary", inp, {"bytes": np.asarray(out["bytes"], dtype=np.uint8)})


def case_divop():
    rng = np.random.default_rng(2)
    n, l, m = 11, 9, 10
    x = coords_from_steps(rng.uniform(0.05, 0.15, n - 1))
    y = coords_from_steps(rng.uniform(0.05, 0.15, l - 1))
    z = coords_from_steps(rng.uniform(0.05, 0.15, m - 1))
    codes = np.zeros((n, l, m), dtype=np.int32)
    codes[1:-1, 1:-1, 1:-1] = 1
    codes[4:6, 3:5, 4:7] = 0
    normals = np.zeros((n, l, m, 3))
    has = np.zeros((n, l, m), dtype=np.int32)
    codes[0, 1:-1, 1:-1] = 4
    normals[0, 1:-1, 1:-1] = (-1, 0, 0)
    has[0, 1:-1, 1:-1] = 1
    codes[1:-1, l - 1, 1:-1] = 3
    inp = {"x": x, "y": y, "z": z, "codes": codes, "normals": normals.reshape(-1, 3), "has_normal": has,
           "space_mask": space_mask_rule(codes), "xvec": rng.standard_normal((n, l, m)),
           "bvec": rng.standard_normal((n, l, m)), "yinit": np.zeros((n, l, m)),
           "rho": rng.uniform(0.5, 2.0, (n, l, m))}
    save("divop_mixed", inp, run("divop", inp))


def dirichlet_codes(nc):
    """Actual interior, Defined (Dirichlet, value 0) ring on the x faces,
    Fictive (Neumann) elsewhere, and a sphere hole: an SPD system."""
    codes = cases.neumann_codes(nc, [cases.sphere_obstacle((0.5, 0.5, 0.5), 0.2)]).astype(np.int32)
    codes[0, 1:-1, 1:-1] = 3
    codes[-1, 1:-1, 1:-1] = 3
    return codes


def case_solvers():
    for name, solver, nc, norm, neumann in [
        ("cg_dirichlet16", "cg", 16, 1e3, False),
        ("cg_neumann20", "cg", 20, 1e3, True),
        ("bicg_dirichlet12", "bicg", 12, 1.0, False),
        ("mr_dirichlet8", "mr", 8, 10.0, False),
        ("bicgstab_dirichlet12", "bicgstab", 12, 1e8, False),
    ]:
        if neumann:
            codes = cases.neumann_codes(nc, [cases.sphere_obstacle((0.5, 0.5, 0.5), 0.2)]).astype(np.int32)
        else:
            codes = dirichlet_codes(nc)
        b = cases.divergence_rhs(codes, seed=7, sigma=1.5)
        if not neumann:
            b[codes == 1] += 0.3  # not zero-mean: a genuinely Dirichlet problem
        b *= norm / np.linalg.norm(b)
        xx = cases.cell_centres(nc)
        inp = {"x": xx, "y": xx.copy(), "z": xx.copy(), "codes": codes, "space_mask": space_mask_rule(codes),
               "x0": np.zeros(codes.shape), "bvec": b}
        save(name, inp, run(solver, inp))


def case_thomas():
    rng = np.random.default_rng(3)
    M = 24
    a = rng.uniform(-1, 1, M + 1)
    c = rng.uniform(-1, 1, M + 1)
    b = 3.0 + rng.uniform(0, 1, M + 1)
    f = rng.standard_normal(M + 1)
    u = np.zeros(M + 1)
    u[0], u[M] = 0.7, -1.3
    inp = {"a": a, "b": b, "c": c, "f": f, "u": u}
    save("thomas24", inp, run("thomas", inp))


def lod_mask(n, l, m):
    """Transport mask: Actual interior, an obstacle of Defined cells, and each
    face a different border type (sector ends must be border points,
    XProblem.cpp:2951-2961)."""
    codes = np.full((n, l, m), 3, dtype=np.int32)
    codes[1:-1, 1:-1, 1:-1] = 1
    codes[0, :, :] = 4
    codes[-1, :, :] = 5
    codes[:, 0, :] = 3
    codes[:, -1, :] = 6
    codes[:, :, 0] = 7
    codes[:, :, -1] = 4
    codes[n // 2:n // 2 + 2, l // 2:l // 2 + 2, 2:4] = 3
    return codes


def case_lod():
    rng = np.random.default_rng(4)
    n, l, m = 10, 9, 8
    x = coords_from_steps(rng.uniform(0.08, 0.12, n - 1))
    y = coords_from_steps(rng.uniform(0.08, 0.12, l - 1))
    z = coords_from_steps(rng.uniform(0.08, 0.12, m - 1))
    codes = lod_mask(n, l, m)
    for scheme in (0, 1):
        inp = {"x": x, "y": y, "z": z, "codes": codes, "space_mask": space_mask_rule(codes),
               "f": rng.standard_normal((n, l, m)), "f_prev": rng.standard_normal((n, l, m)),
               "u": rng.standard_normal((n, l, m)), "v": rng.standard_normal((n, l, m)),
               "w": rng.standard_normal((n, l, m)), "rp": rng.standard_normal((n, l, m)),
               "kappa": rng.uniform(0.01, 0.1, (n, l, m)), "tau": np.array([0.01]),
               "scheme": np.array([scheme], dtype=np.int32)}
        save("lod_scheme%d" % scheme, inp, run("lod", inp))


def case_vp_masks():
    """The reference's own masks: a TVPProblem with the live cylinder-channel
    statement (TSimpleWithCylinderPressureMaskGenerator with the pressure
    condition of TCanalIBMWithElasticBoundary, XProblem.cpp:10340-10373)
    builds its U, V, W, P grids (XStatement.cpp:7717-7977) and space models
    (LocalPrepareVariables, XProblem.cpp:4116-4206); on P also the activity
    table and A x for a seeded x."""
    base = {"dims": np.array([31, 25, 25], dtype=np.int32), "lengths": np.array([1.5, 1.2, 1.2]),
            "p_in": np.array([6.0]), "p_out": np.array([1.0])}
    shape = run("vp_masks", base)["P_codes"].shape
    inp = dict(base)
    inp["xvec"] = np.random.default_rng(11).standard_normal(shape)
    save("vp_cylinder", inp, run("vp_masks", inp))


def case_ib():
    rng = np.random.default_rng(5)
    nc = 16
    h = 1.0 / nc
    xx = np.arange(nc + 2) * h
    codes = np.ones((nc + 2,) * 3, dtype=np.int32)
    nn = 25
    nodes = rng.uniform(0.2, 0.8, (nn, 3))
    nodes[0] = (0.02, 0.5, 0.97)   # near the walls: exercises the index clamp
    nodes[1] = (0.5, 0.5, 0.5)     # exactly on grid lines
    for axis in (1, 2, 3):
        inp = {"x": xx, "y": xx.copy(), "z": xx.copy(), "codes": codes, "axis": np.array([axis], dtype=np.int32),
               "nodes": nodes, "force_n": rng.standard_normal(nn), "vel": rng.standard_normal(codes.shape),
               "zero": np.zeros(codes.shape)}
        save("ib_axis%d" % axis, inp, run("ib", inp))


if __name__ == "__main__":
    if not os.path.exists(HARNESS):
        sys.exit("build the reference harness first: make -C oracle ref")
    CASES = {"op": case_op, "divop": case_divop, "solvers": case_solvers, "thomas": case_thomas, "lod": case_lod,
             "ib": case_ib, "pack": case_pack, "pack_solve": case_pack_solve, "pack_solve_spd": case_pack_solve_spd, "io": case_io,
             "vp_masks": case_vp_masks}
    # python make_golden.py [case ...]: only the named cases (all by default)
    for name in sys.argv[1:] or list(CASES):
        CASES[name]()
