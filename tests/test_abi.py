"""The C ABI library loads and exports every symbol include/gmrf_c.h declares (no GPU calls)."""

import os
import re

from conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "gmrf_c.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gmrf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_functions():
    names = declared_functions()
    assert "gmrf_solver_factorize" in names and "gmrf_batch_selinv" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    from inla_gpu import _lib

    missing = [n for n in declared_functions() if not hasattr(_lib.lib, n)]
    assert not missing, missing
    unbound = [n for n in declared_functions() if n not in _lib.SIGNATURES]
    assert not unbound, unbound


def test_status_enum_matches_header():
    from inla_gpu import _lib

    text = open(os.path.join(ROOT, "include", "gmrf_c.h")).read()
    for name in ["OK", "SINGULAR", "NOT_POSITIVE_DEFINITE", "SOLVE_FAILURE", "INVALID_ARGUMENT",
                 "OUT_OF_MEMORY", "DEVICE_ERROR"]:
        m = re.search(rf"GMRF_{name}\s*=\s*(\d+)", text)
        assert m and int(m.group(1)) == getattr(_lib, f"GMRF_{name}")
