"""ctypes binding of libqpanda_sdk.so: the QRunes front end and the QPandaSDK
facade on the MI355X engine (include/qsv_sdk.h)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqpanda_sdk.so")

# QPanda's QError values (QuantumInstructionHandle/QError.h)
Q_ERROR_NONE, UNDEFINE_ERROR, Q_PARAMETER_ERROR, QBIT_ERROR = 2, 3, 4, 5
INIT_STATE_ERROR, SET_COMPUTE_UNIT_ERROR, RUN_PROGRAM_ERROR, NOT_READ_FILE_ERROR = 7, 9, 10, 13
CPU, GPU = 1, 2

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run qpanda_amd/build.sh)")
        L = C.CDLL(LIB_PATH)
        vp, I, cp, ip = C.c_void_p, C.c_int, C.c_char_p, C.POINTER(C.c_int)
        sigs = {
            "loadFileAPI": ([cp], I), "setComputeUnitAPI": ([I], I), "runAPI": ([I], I),
            "getFileNameAPI": ([cp, ip], I), "getResultAPI": ([cp, ip], I), "getQuantumStateAPI": ([cp, ip], I),
            "qrunes_parse_file": ([cp, cp, ip], I), "qrunes_parse_text": ([cp, cp, ip], I),
            "qpanda_sdk_create": ([], vp), "qpanda_sdk_destroy": ([vp], None),
            "qpanda_sdk_load_file": ([vp, cp], I), "qpanda_sdk_set_compute_unit": ([vp, I], I),
            "qpanda_sdk_set_seed": ([vp, C.c_int32], None), "qpanda_sdk_set_resample_disabled": ([vp, I], None),
            "qpanda_sdk_run": ([vp, I], I), "qpanda_sdk_last_run_sampled": ([vp], I),
            "qpanda_sdk_get_result": ([vp, cp, ip], I), "qpanda_sdk_get_state": ([vp, cp, ip], I),
            "qpanda_sdk_get_amplitudes": ([vp, C.c_void_p, C.c_int64], C.c_int64),
            "qpanda_sdk_save_state": ([vp, cp, I], I),
        }
        for name, (args, res) in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _string(fn, *args) -> str:
    n = C.c_int(0)
    fn(*args, None, C.byref(n))
    buf = C.create_string_buffer(max(1, n.value))
    n2 = C.c_int(n.value)
    fn(*args, buf, C.byref(n2))
    return buf.raw[: n.value].decode("latin-1")


def _nodes(text: str):
    out = []
    for line in text.splitlines():
        head, tail = line.split(" | ")
        parts = head.split()
        kv = dict(x.split("=", 1) for x in tail.split())
        out.append({"kind": parts[0], "qubits": [int(q) for q in parts[1:]], "angle": float(kv["angle"]),
                    "dagger": kv["dagger"] == "1",
                    "controls": [int(c) for c in kv["ctrl"].split(",") if c != ""], "creg": int(kv["creg"])})
    return out


def parse_text(text: str):
    """(qubits or -1, flattened instruction list) of a QRunes program."""
    raw = text.encode("latin-1")
    n = C.c_int(0)
    r = lib().qrunes_parse_text(raw, None, C.byref(n))
    if r < 0:
        return r, []
    buf = C.create_string_buffer(max(1, n.value))
    n2 = C.c_int(n.value)
    lib().qrunes_parse_text(raw, buf, C.byref(n2))
    return r, _nodes(buf.raw[: n.value].decode("latin-1"))


def parse_file(path: str):
    raw = path.encode()
    n = C.c_int(0)
    r = lib().qrunes_parse_file(raw, None, C.byref(n))
    if r < 0:
        return r, []
    buf = C.create_string_buffer(max(1, n.value))
    n2 = C.c_int(n.value)
    lib().qrunes_parse_file(raw, buf, C.byref(n2))
    return r, _nodes(buf.raw[: n.value].decode("latin-1"))


class SDK:
    """One QPandaSDK instance (handle-based; the reference's process-wide
    instance is reached through loadFileAPI / runAPI / ...)."""

    def __init__(self):
        self.h = lib().qpanda_sdk_create()

    def close(self):
        if self.h:
            lib().qpanda_sdk_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def load_file(self, path: str) -> int:
        return lib().qpanda_sdk_load_file(self.h, path.encode())

    def set_compute_unit(self, unit: int) -> int:
        return lib().qpanda_sdk_set_compute_unit(self.h, unit)

    def set_seed(self, seed: int) -> None:
        lib().qpanda_sdk_set_seed(self.h, seed)

    def set_resample_disabled(self, off: bool) -> None:
        lib().qpanda_sdk_set_resample_disabled(self.h, int(off))

    def run(self, repeat: int) -> int:
        return lib().qpanda_sdk_run(self.h, repeat)

    def last_run_sampled(self) -> bool:
        return bool(lib().qpanda_sdk_last_run_sampled(self.h))

    def result(self) -> str:
        return _string(lib().qpanda_sdk_get_result, self.h)

    def state_text(self) -> str:
        return _string(lib().qpanda_sdk_get_state, self.h)

    def save_state(self, path: str, binary: bool) -> int:
        """path.fullstate (state text) or path.fullstate.binary ({f64, f64} records)."""
        return lib().qpanda_sdk_save_state(self.h, path.encode(), int(binary))

    def amplitudes(self):
        """complex128 numpy array of the last run's final state."""
        import numpy as np
        n = lib().qpanda_sdk_get_amplitudes(self.h, None, 0)
        out = np.zeros(n, np.complex128)
        if n:
            lib().qpanda_sdk_get_amplitudes(self.h, out.ctypes.data, n)
        return out
