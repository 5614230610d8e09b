import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "rust-inla_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

# Achieved parity errors, one row per (test, quantity): written before each
# assertion so that a failing comparison still reports what it measured.
_ERRORS = []


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture
def record_error(request):
    """record_error(quantity, achieved, bar, kind) logs the achieved error of one
    comparison (kind: "abs" or "rel") for the end-of-session parity table."""

    def rec(quantity, achieved, bar, kind="rel"):
        _ERRORS.append({"test": request.node.nodeid.split("::", 1)[-1], "quantity": quantity,
                        "achieved": float(achieved), "bar": float(bar), "kind": kind})
        return float(achieved)

    return rec


def pytest_terminal_summary(terminalreporter):
    if not _ERRORS:
        return
    tr = terminalreporter
    tr.section("achieved parity errors")
    for e in _ERRORS:
        flag = "ok" if e["achieved"] <= e["bar"] else "OVER"
        tr.write_line(f"{e['test'][:70]:70s} {e['quantity']:24s} {e['kind']:3s} "
                      f"{e['achieved']:.3e} <= {e['bar']:.0e} {flag}")
    path = os.environ.get("PARITY_ERRORS_JSON")
    if path:
        with open(path, "w") as f:
            json.dump(_ERRORS, f, indent=1)
