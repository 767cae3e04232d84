"""The C ABI library loads (no GPU needed to load it) and exports exactly the
entry points include/xflow.h declares; the Python binding covers them all."""
import os
import re
import subprocess

from conftest import ROOT


def declared():
    txt = open(os.path.join(ROOT, "include", "xflow.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(xf_[a-z0-9_]+)\s*\(", txt))


def test_library_loads_and_exports_every_declared_symbol():
    from xflow_amd import _lib
    L = _lib.lib()
    names = declared()
    assert len(names) > 20
    for n in sorted(names):
        assert hasattr(L, n), "missing export %s" % n
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\b(xf_[a-z0-9_]+)\b", out))
    assert names <= exported
    assert exported - names == set(), "undeclared exports: %s" % sorted(exported - names)


def test_binding_table_matches_header():
    from xflow_amd import _lib
    assert set(_lib.SIGNATURES) == declared()


def test_product_package_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "xflow_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".h", ".cpp")):
                s = open(os.path.join(dp, f)).read()
                assert "xf_oracle" not in s and "liboracle" not in s, f
