"""CPU-side checks of the C ABI: the library loads here (no GPU needed) and
exports every entry point include/qsv.h declares; the Python binding lists the
same set.  No compute call is made."""
import ctypes
import os
import re

import pytest

from qpanda_amd import engine as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "qsv.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(qsv_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "qsv_create" in names and "qsv_run_circuit" in names and len(names) > 30


def test_library_exports_every_declared_symbol():
    if not os.path.exists(E.LIB_PATH):
        pytest.fail("libqsv.so not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(E.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(E.EXPORTED) == declared()


def test_rng_entry_points_are_pure_host_code():
    # Park-Miller KAT through the engine's own generator (no device involved).
    s = 1
    for _ in range(10000):
        s = E.pm_next(s)
    assert s == 1043618065 and E.pm_jump(1, 10000) == 1043618065


def sdk_declared():
    src = open(os.path.join(ROOT, "include", "qsv_sdk.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src)) - {"extern"})


def test_sdk_library_exports_every_declared_symbol():
    from qpanda_amd import sdk
    names = sdk_declared()
    assert {"loadFileAPI", "setComputeUnitAPI", "runAPI", "getResultAPI", "getQuantumStateAPI",
            "qrunes_parse_file", "qpanda_sdk_run"} <= set(names)
    lib = sdk.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the binding declares a signature for every entry point
    assert [n for n in names if getattr(lib, n).argtypes is None] == []
