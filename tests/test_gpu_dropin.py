"""The drop-in C++ classes (xflow_amd/host) driven by reference-style host
code (tests/host/dropin_poisson.cpp and dropin_pack.cpp, built by
__graft_entry__.build())."""
import json
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def test_dropin_manufactured_poisson():
    exe = os.path.join(ROOT, "tests", "host", "dropin_poisson")
    assert os.path.exists(exe), "build first: __graft_entry__.build()"
    p = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=240)
    print(p.stdout, p.stderr)
    res = json.loads(p.stdout.strip().splitlines()[-1])
    assert p.returncode == 0, res
    assert res["trunc24"] < res["trunc12"] / 3
    assert res["err_cg24"] < 5e-3


def test_dropin_packed_operator():
    """T1dPackOperator on the drop-in classes (tests/host/dropin_pack.cpp): the
    Rn path of the reference's Poisson demo and BiCG with the transposed rows."""
    exe = os.path.join(ROOT, "tests", "host", "dropin_pack")
    assert os.path.exists(exe), "build first: __graft_entry__.build()"
    p = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=240)
    print(p.stdout, p.stderr)
    res = json.loads(p.stdout.strip().splitlines()[-1])
    assert p.returncode == 0, res
    assert res["row_diff"] == 0 and res["prod_diff"] < 1e-13
    assert abs(res["it3"] - res["itRn"]) <= 2
