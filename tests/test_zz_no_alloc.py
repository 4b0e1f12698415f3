"""The allocation-free readout (qsv_readout_noalloc) through the C ABI: its
FP64 results equal the handle-based readouts (qsv_norm2, qsv_prob0,
qsv_expect_zz_chain), which share its kernels and summation order, on two
calls in a row, the second on a smaller state; bad arguments are rejected."""
import ctypes as C
import sys

import numpy as np
import pytest

from qpanda_amd import circuits as Circ
from qpanda_amd import engine as E

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc(sv, what, q, out):
    """One call through the C ABI on the state's stream; the result is read
    back from the device double (amplitude 0 of the 1-qubit state `out`)."""
    rc = E.lib().qsv_readout_noalloc(C.c_void_p(sv.device_ptr()), sv.n, sv.precision, what, q,
                                     C.c_void_p(out.device_ptr()), C.c_void_p(sv.stream()))
    assert rc == 0, E.lib().qsv_last_error()
    sv.sync()
    return float(out.get()[0].real)


@pytest.mark.parametrize("prec", [E.C128, E.C64])
def test_noalloc_readout_matches_handle_readouts(prec):
    with E.StateVector(1, E.C128) as out:
        for n, seed in ((20, 5), (14, 6)):          # the second call on a smaller state
            say("noalloc readout prec", prec, "n", n)
            with E.StateVector(n, prec) as sv:
                sv.set(Circ.random_gaussian_state(n, seed) * 1.7)
                q = n // 3
                got = {
                    "norm2": noalloc(sv, E.READ_NORM2, 0, out),
                    "prob0": noalloc(sv, E.READ_PROB0, q, out),
                    "zz": noalloc(sv, E.READ_ZZ_CHAIN, 0, out),
                }
                ref = {"norm2": sv.norm2(), "prob0": sv.prob0(q), "zz": sv.expect_zz_chain()}
                say("  got", got, "ref", ref)
                assert got == ref
                assert abs(got["norm2"] - 1.7 ** 2) < 1e-5


def test_noalloc_readout_rejects_bad_arguments():
    host = (C.c_double * 8)()
    with E.StateVector(12, E.C128) as sv, E.StateVector(1, E.C128) as out:
        lib = E.lib()
        dp, op, st = C.c_void_p(sv.device_ptr()), C.c_void_p(out.device_ptr()), C.c_void_p(sv.stream())
        assert lib.qsv_readout_noalloc(dp, 35, E.C128, E.READ_NORM2, 0, op, st) == -1      # n beyond 34
        assert lib.qsv_readout_noalloc(dp, 12, 7, E.READ_NORM2, 0, op, st) == -1           # precision
        assert lib.qsv_readout_noalloc(dp, 12, E.C128, 9, 0, op, st) == -1                 # what
        assert lib.qsv_readout_noalloc(dp, 12, E.C128, E.READ_PROB0, 12, op, st) == -2     # qubit
        assert lib.qsv_readout_noalloc(C.cast(host, C.c_void_p), 12, E.C128, E.READ_NORM2, 0, op, st) == -1
        assert lib.qsv_readout_noalloc(dp, 12, E.C128, E.READ_NORM2, 0, C.cast(host, C.c_void_p), st) == -1
        assert lib.qsv_readout_noalloc(None, 12, E.C128, E.READ_NORM2, 0, op, st) == -1
        # still usable afterwards
        assert abs(noalloc(sv, E.READ_NORM2, 0, out) - 1.0) < 1e-12
