"""Output-file restatements in the oracle against the compiled reference's
files (tests/golden: cnpy::npy_save and TImmersedBoundaryProblem::OutputBoundary
bytes), and the flow VTK's node averages against closed forms.

The flow VTK (TVPProblem::SolveBySplitting, XProblem.cpp:6132-6379) is written
inside the reference's time loop and cannot be driven on its own: its
restatement is pinned by the closed forms below (linear fields are reproduced
exactly at the nodes of an all-Actual uniform grid; next to Fictive cells the
eight-cell sum over the non-Fictive count gives the hand-computed value)."""
import numpy as np
import pytest

import xf_oracle as xo
from conftest import golden


@pytest.mark.parametrize("name", ["npy_f8_3d", "npy_f8_1d"])
def test_npy_bytes_match_cnpy(name):
    ins, outs = golden(name)
    assert xo.npy_bytes(ins["data"]) == bytes(outs["bytes"])


def test_vtk_boundary_text_matches_reference():
    ins, outs = golden("vtk_boundary")
    txt = xo.vtk_boundary_text(int(ins["circles"][0]), int(ins["in_circle"][0]), ins["xyz"], ins["xyz0"], ins["vel"],
                               ins["p"], ins["wss"])
    assert txt.encode() == bytes(outs["bytes"])


def test_vtk_flow_linear_fields_exact():
    n = 7
    h = 0.25
    x = np.arange(n) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    codes = np.ones((n, n, n), dtype=np.int32)
    p = 1.0 + 2.0 * X - 3.0 * Y + 0.5 * Z          # cell values; node (i,j,k) sits at the cells' corner mean
    node = xo.vtk_node_average8(codes, p)
    # the mean of the eight cells (i..i+1, ...) of a linear field is its value at (x_i + h/2, ...)
    I = np.arange(1, n - 1)
    Kk, Jj, Ii = np.meshgrid(I, I, I, indexing="ij")
    want = 1.0 + 2.0 * (Ii * h + h / 2) - 3.0 * (Jj * h + h / 2) + 0.5 * (Kk * h + h / 2)
    assert np.allclose(node, want.reshape(-1), rtol=0, atol=1e-14)
    # velocity: mean of the four faces around the node; u = x(i), v = y(j), w = z(k)
    # are constant over the faces averaged, so the node gets (x_i, y_j, z_k)
    txt = xo.vtk_flow_text(codes, x, x, x, X.copy(), Y.copy(), Z.copy(), p=p)
    block = txt.split("VECTORS Velocity double\n")[1].split("\nSCALARS")[0]
    vel = np.array([[float(t) for t in ln.split()] for ln in block.split("\n ") if ln.strip()])
    want_v = np.stack([x[Ii].reshape(-1), x[Jj].reshape(-1), x[Kk].reshape(-1)], axis=1)
    assert np.array_equal(vel, want_v)
    assert txt.startswith("# vtk DataFile Version 3.0\nFlow\nASCII\n\nDATASET STRUCTURED_GRID\nDIMENSIONS 5 5 5\n")


def test_vtk_flow_fictive_count():
    """Next to Fictive cells the sum still takes all eight cells and divides by
    the non-Fictive count (XProblem.cpp:6216-6239)."""
    n = 4
    codes = np.ones((n, n, n), dtype=np.int32)
    codes[2, 2, 2] = 0
    f = np.arange(n ** 3, dtype=np.float64).reshape(n, n, n)
    vals = xo.vtk_node_average8(codes, f).reshape(n - 2, n - 2, n - 2)  # [k][j][i]
    # node (1,1,1): its eight cells (1..2)^3 include the Fictive (2,2,2): sum of 8 over 7
    cells = [f[a, b, c] for a in (1, 2) for b in (1, 2) for c in (1, 2)]
    assert vals[0, 0, 0] == pytest.approx(sum(cells) / 7, rel=1e-15)
    # node (2,2,2) is Fictive: 0
    assert vals[1, 1, 1] == 0.0
