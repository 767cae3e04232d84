"""The drop-in C++ classes (xflow_amd/host) driven by reference-style host
code (tests/host/dropin_poisson.cpp, dropin_pack.cpp and dropin_pressure.cpp,
built by __graft_entry__.build())."""
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


def test_dropin_pressure_solve_call_forms():
    """The pressure solve of a projection step with the call forms of
    TVPProblem::SolveBySplitting (XProblem.cpp:5516-5590) on the drop-in
    classes (tests/host/dropin_pressure.cpp): withDensity, p.EndIndex,
    Assign, the packed branch through Rn vectors and a T1dPackOperator, the
    grid branch with UpdateA/UpdateRP/UpdateP on a later step, the
    six-argument TSLAE, Property accessors.  Both branches reach BiCGStab's
    max-norm tolerance on a variable-density channel and agree."""
    exe = os.path.join(ROOT, "tests", "host", "dropin_pressure")
    assert os.path.exists(exe), "build first: __graft_entry__.build()"
    p = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=240)
    print(p.stdout, p.stderr)
    res = json.loads(p.stdout.strip().splitlines()[-1])
    assert p.returncode == 0 and res["failures"] == 0, res
    assert res["res_pack"] <= 1e-4 and res["res_grid"] <= 1e-4 and res["res_step2"] <= 1e-4
    assert res["res_fused"] <= 1e-4  # SolveByFastBiConjugateGradientsStab on the step-2 system
    assert res["pack_vs_grid"] <= 1e-5 and res["six_vs_three"] == 0
