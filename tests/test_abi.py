"""The C-ABI library loads without a GPU and exports every declared entry point."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            src = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*)\s+(ofh_\w+)\s*\(", src, re.M))
    return names


def test_library_exports_all_declared_symbols():
    import ofeli_amd as oa
    L = oa.lib()
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    from ofeli_amd import _abi
    assert declared_symbols() == set(_abi.SIGNATURES)


def test_host_side_entry_points_without_gpu():
    """Mesh and numbering entry points are host code: callable on the CPU."""
    import numpy as np
    from ofeli_amd import api
    coord, code, elem = api.mesh2d(4, 3)
    assert coord.shape == (20, 2) and elem.shape == (24, 3)
    dof, neq = api.number_equations(code)
    assert neq == 3 * 2
    x, w = api.gauss(2)
    assert np.allclose(x, [-1 / np.sqrt(3), 1 / np.sqrt(3)]) and np.allclose(w, 1.0)
