"""Output files written by the library: cnpy .npy and the immersed-boundary VTK
byte-identical to the compiled reference's files (tests/golden), the flow VTK
byte-identical to the oracle's restatement (node averages computed on the GPU,
text formatted on the host with the reference's printf formats)."""
import numpy as np
import pytest
import torch

import xflow_amd as xf
import xf_oracle as xo
from conftest import golden
from xflow_amd import cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = xf.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", ["npy_f8_3d", "npy_f8_1d"])
@pytest.mark.parametrize("on_device", [True, False])
def test_npy_save_bytes(ctx, tmp_path, name, on_device):
    ins, outs = golden(name)
    a = torch.from_numpy(ins["data"]).cuda() if on_device else ins["data"]
    path = tmp_path / "a.npy"
    xf.npy_save(ctx, path, a)
    assert path.read_bytes() == bytes(outs["bytes"])


def test_vtk_boundary_bytes(ctx, tmp_path):
    ins, outs = golden("vtk_boundary")
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    path = tmp_path / "b.vtk"
    xf.write_vtk_boundary(ctx, path, int(ins["circles"][0]), int(ins["in_circle"][0]), dev(ins["xyz"]),
                          dev(ins["xyz0"]), dev(ins["vel"]), dev(ins["p"]), dev(ins["wss"]))
    assert path.read_bytes() == bytes(outs["bytes"])


@pytest.mark.parametrize("shape", [(12, 10, 9), (34, 20, 27)])
def test_vtk_flow_matches_oracle(ctx, tmp_path, shape):
    rng = np.random.default_rng(3)
    codes = cases.neumann_codes(None, [cases.sphere_obstacle((0.5, 0.5, 0.5), 0.25)], shape=shape).astype(np.int32)
    codes_p = codes.copy()
    codes_p[1:4, 1:4, 1:4] = 0  # a different pressure mask
    h = [cases.axis_steps(s) for s in shape]
    x, y, z = (np.concatenate([[0.0], np.cumsum(hh)]) for hh in h)
    g = xf.Grid(ctx, h[0], h[1], h[2], codes)
    gp = xf.Grid(ctx, h[0], h[1], h[2], codes_p)
    f = {k: rng.standard_normal(codes.shape) for k in ("u", "v", "w", "p", "density", "concentration", "viscosity")}
    dev = {k: torch.from_numpy(v).cuda() for k, v in f.items()}
    path = tmp_path / "flow.vtk"
    xf.write_vtk_flow(ctx, path, (g, gp, g, g), x, y, z, dev["u"], dev["v"], dev["w"], p=dev["p"],
                      density=dev["density"], concentration=dev["concentration"], viscosity=dev["viscosity"])
    want = xo.vtk_flow_text((codes, codes_p, codes, codes), x, y, z, f["u"], f["v"], f["w"], p=f["p"],
                            density=f["density"], concentration=f["concentration"], viscosity=f["viscosity"])
    got = path.read_text()
    assert got == want
    # fields left out are written as zeros
    path2 = tmp_path / "flow2.vtk"
    xf.write_vtk_flow(ctx, path2, g, x, y, z, dev["u"], dev["v"], dev["w"])
    assert path2.read_text() == xo.vtk_flow_text(codes, x, y, z, f["u"], f["v"], f["w"])
    g.close()
    gp.close()
