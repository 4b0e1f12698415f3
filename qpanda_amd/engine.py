"""Python binding of libqsv.so (the C ABI declared in include/qsv.h).

The shared library is built in-tree by qpanda_amd/build.sh; loading fails
loudly when it is missing -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import circuits as _circ

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqsv.so")

C128, C64 = 0, 1
READ_NORM2, READ_PROB0, READ_ZZ_CHAIN = 0, 1, 2   # qsv_readout_noalloc
FILE_BINARY, FILE_TEXT = 0, 1

EXPORTED = [
    "qsv_create", "qsv_destroy", "qsv_reset_zero", "qsv_num_qubits", "qsv_precision",
    "qsv_set_state", "qsv_get_state", "qsv_save_state", "qsv_load_state", "qsv_device_ptr", "qsv_stream", "qsv_sync",
    "qsv_set_gate_tiling",
    "qsv_h", "qsv_x", "qsv_apply_1q", "qsv_diag", "qsv_rx", "qsv_ry", "qsv_rz",
    "qsv_cphase", "qsv_cr", "qsv_iswap", "qsv_sqiswap", "qsv_cswap", "qsv_apply_2q",
    "qsv_apply_block", "qsv_prob0", "qsv_norm2", "qsv_collapse", "qsv_measure",
    "qsv_marginals", "qsv_pmeasure", "qsv_expect_zz_chain", "qsv_run_circuit",
    "qsv_sample_shots", "qsv_sample_shots_copy", "qsv_gather", "qsv_pm_next", "qsv_pm_jump", "qsv_pm_seed_from_tm", "qsv_version",
    "qsv_last_error", "qsv_device_count", "qsv_copy_state", "qsv_inner", "qsv_fuse_plan", "qsv_window_plan",
    "qsv_window_plan_beam", "qsv_permute_qubits",
    "qsv_cm_create", "qsv_cm_destroy", "qsv_cm_nparams", "qsv_cm_get_params", "qsv_cm_set_params",
    "qsv_cm_block_features", "qsv_cm_forward", "qsv_cm_gradients", "qsv_cm_train", "qsv_set_costmodel",
    "qsv_timing_log", "qsv_batch_create", "qsv_batch_destroy", "qsv_batch_autotune", "qsv_batch_run",
    "qsv_batch_graph_run", "qsv_readout_noalloc",
]

CM_FEATURES = 12
DEFAULT_COSTMODEL = os.path.join(HERE, "costmodel_params.f32")


class QsvError(RuntimeError):
    pass


class RunOpts(C.Structure):
    _fields_ = [("exact", C.c_int), ("fuse_k", C.c_int), ("split", C.c_int),
                ("scheduler", C.c_int), ("profile", C.c_int), ("tile", C.c_int),
                ("group", C.c_int), ("pass_tb", C.c_int), ("pass_kernel", C.c_int),
                ("fixed_low", C.c_int)]


class RunStats(C.Structure):
    _fields_ = [("gates", C.c_int64), ("launches", C.c_int64), ("blocks", C.c_int64),
                ("tie_resolutions", C.c_int64), ("device_ms", C.c_double),
                ("model_ms_pred", C.c_double), ("class_launches", C.c_int64 * 8),
                ("class_ms", C.c_double * 8), ("class_bytes", C.c_double * 8),
                ("class_flops", C.c_double * 8), ("plan_ms", C.c_double), ("fuse_k", C.c_int64),
                ("planner", C.c_int64)]

    CLASS_NAMES = ["dense1", "diag", "perm", "phase", "block64", "fused", "readout", "pass"]

    def classes(self) -> dict:
        out = {}
        for i, nm in enumerate(self.CLASS_NAMES):
            if self.class_launches[i]:
                ms = self.class_ms[i]
                out[nm] = {"launches": int(self.class_launches[i]), "ms": ms,
                           "bytes": self.class_bytes[i], "flops": self.class_flops[i],
                           "gbs": self.class_bytes[i] / (ms * 1e6) if ms > 0 else 0.0}
        return out


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QsvError(f"{LIB_PATH} is not built: run qpanda_amd/build.sh "
                           "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp, I, D, U64 = C.c_void_p, C.c_int, C.c_double, C.c_uint64
        P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        sig = {
            "qsv_create": ([I, I, C.POINTER(vp)], I),
            "qsv_destroy": ([vp], I),
            "qsv_reset_zero": ([vp], I),
            "qsv_num_qubits": ([vp], I),
            "qsv_precision": ([vp], I),
            "qsv_set_state": ([vp, P], I),
            "qsv_get_state": ([vp, P], I),
            "qsv_save_state": ([vp, C.c_char_p, I, I], I),
            "qsv_load_state": ([vp, C.c_char_p, I], I),
            "qsv_device_ptr": ([vp, C.POINTER(vp)], I),
            "qsv_stream": ([vp, C.POINTER(vp)], I),
            "qsv_sync": ([vp], I),
            "qsv_set_gate_tiling": ([vp, I], I),
            "qsv_h": ([vp, I, U64, I], I),
            "qsv_x": ([vp, I, U64], I),
            "qsv_apply_1q": ([vp, I, P, U64], I),
            "qsv_diag": ([vp, I, P, U64], I),
            "qsv_rx": ([vp, I, D, I, U64], I),
            "qsv_ry": ([vp, I, D, I, U64], I),
            "qsv_rz": ([vp, I, D, I, U64], I),
            "qsv_cphase": ([vp, I, I, D, D, U64], I),
            "qsv_cr": ([vp, I, I, D, I], I),
            "qsv_iswap": ([vp, I, I, I], I),
            "qsv_sqiswap": ([vp, I, I, I, I], I),
            "qsv_cswap": ([vp, I, I, I], I),
            "qsv_apply_2q": ([vp, I, I, P], I),
            "qsv_apply_block": ([vp, C.POINTER(I), I, P, I], I),
            "qsv_prob0": ([vp, I, C.POINTER(D)], I),
            "qsv_norm2": ([vp, C.POINTER(D)], I),
            "qsv_collapse": ([vp, I, I, D], I),
            "qsv_measure": ([vp, I, D, C.POINTER(I)], I),
            "qsv_marginals": ([vp, C.POINTER(I), I, P], I),
            "qsv_pmeasure": ([vp, C.POINTER(I), I, np.ctypeslib.ndpointer(dtype=np.uint64), P,
                              C.POINTER(I)], I),
            "qsv_expect_zz_chain": ([vp, C.POINTER(D)], I),
            "qsv_readout_noalloc": ([vp, I, I, I, I, vp, vp], I),
            "qsv_run_circuit": ([vp, vp, C.c_int64, vp, C.POINTER(RunOpts), C.POINTER(C.c_int32),
                                 vp, C.POINTER(RunStats)], I),
            "qsv_sample_shots": ([vp, C.POINTER(I), I, C.c_int64, C.POINTER(C.c_int32), C.c_int64,
                                  np.ctypeslib.ndpointer(dtype=np.uint64), C.POINTER(C.c_int64)], I),
            "qsv_gather": ([vp, np.ctypeslib.ndpointer(dtype=np.uint64), C.c_int64,
                            np.ctypeslib.ndpointer(dtype=np.float64)], I),
            "qsv_sample_shots_copy": ([vp, C.POINTER(I), I, C.c_int64, C.POINTER(C.c_int32), C.c_int64,
                                       np.ctypeslib.ndpointer(dtype=np.uint64), C.POINTER(C.c_int64)], I),
            "qsv_pm_next": ([C.c_int32], C.c_int32),
            "qsv_pm_jump": ([C.c_int32, U64], C.c_int32),
            "qsv_pm_seed_from_tm": ([I] * 6, C.c_int32),
            "qsv_version": ([], C.c_char_p),
            "qsv_last_error": ([], C.c_char_p),
            "qsv_device_count": ([C.POINTER(I)], I),
            "qsv_fuse_plan": ([vp, C.c_int64, vp, I, I, C.c_int64, vp, vp, vp, vp,
                               C.POINTER(C.c_int64)], I),
            "qsv_window_plan": ([vp, C.c_int64, vp, I, I, I, I, C.c_int64, vp, vp, vp, vp, vp,
                                 C.POINTER(C.c_int64)], I),
            "qsv_window_plan_beam": ([vp, C.c_int64, vp, I, I, I, I, I, I, I, I, I, I, C.c_int64, vp, vp, vp, vp, vp,
                                      C.POINTER(C.c_int64)], I),
            "qsv_cm_create": ([U64, C.POINTER(vp)], I),
            "qsv_cm_destroy": ([vp], I),
            "qsv_cm_nparams": ([], I),
            "qsv_cm_get_params": ([vp, vp], I),
            "qsv_cm_set_params": ([vp, vp], I),
            "qsv_cm_block_features": ([I, I, I, I, C.POINTER(I), vp], I),
            "qsv_cm_forward": ([vp, vp, C.c_int64, vp], I),
            "qsv_cm_gradients": ([vp, vp, vp, C.c_int64, vp, C.POINTER(D)], I),
            "qsv_cm_train": ([vp, vp, vp, C.c_int64, I, C.c_float, C.POINTER(D)], I),
            "qsv_set_costmodel": ([vp, vp], I),
            "qsv_timing_log": ([vp, vp, vp, C.c_int64, C.POINTER(C.c_int64)], I),
            "qsv_batch_create": ([vp, vp, C.c_int64, vp, C.POINTER(vp)], I),
            "qsv_batch_destroy": ([vp], I),
            "qsv_batch_autotune": ([vp, C.POINTER(I), C.POINTER(D)], I),
            "qsv_batch_run": ([vp, I, vp, vp, C.POINTER(D)], I),
            "qsv_batch_graph_run": ([vp, I, vp, C.POINTER(D)], I),
            "qsv_copy_state": ([vp, vp], I),
            "qsv_inner": ([vp, vp, C.POINTER(D), C.POINTER(D)], I),
            "qsv_permute_qubits": ([vp, C.POINTER(I)], I),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise QsvError(f"qsv error {rc}: {lib().qsv_last_error().decode()}")


def _ints(xs):
    return (C.c_int * len(xs))(*[int(x) for x in xs])


def _mat(u) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(u, np.complex128)).view(np.float64).ravel()


def mask(qubits) -> int:
    m = 0
    for q in qubits:
        m |= 1 << int(q)
    return m


class StateVector:
    """A 2^n-amplitude state resident in HBM."""

    def __init__(self, n: int, precision: int = C128):
        h = C.c_void_p()
        _check(lib().qsv_create(n, precision, C.byref(h)))
        self.h = h
        self.n = n
        self.precision = precision

    def close(self):
        if self.h:
            lib().qsv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state ----
    def reset(self):
        _check(lib().qsv_reset_zero(self.h))

    def set(self, amps):
        _check(lib().qsv_set_state(self.h, _mat(amps)))

    def get(self) -> np.ndarray:
        out = np.empty(2 << self.n, np.float64)
        _check(lib().qsv_get_state(self.h, out))
        return out.view(np.complex128)

    def save(self, path: str, text: bool = False, digits: int = 0):
        """State file: binary {f64 re, f64 im} records, or the getQState text
        with `digits` significant digits (0: 6, as the reference prints)."""
        _check(lib().qsv_save_state(self.h, path.encode(), FILE_TEXT if text else FILE_BINARY, digits))

    def load(self, path: str, text: bool = False):
        _check(lib().qsv_load_state(self.h, path.encode(), FILE_TEXT if text else FILE_BINARY))

    def sync(self):
        _check(lib().qsv_sync(self.h))

    def set_gate_tiling(self, on: bool = True):
        """Queue gate sweeps and apply runs of gates whose qubits share a
        64 KiB tile in one sweep (bit-identical state); any readout flushes."""
        _check(lib().qsv_set_gate_tiling(self.h, 1 if on else 0))

    def attach_costmodel(self, model: "CostModel | None"):
        self._cm = model                      # keep the model alive while attached
        _check(lib().qsv_set_costmodel(self.h, model.h if model is not None else None))

    def timing_log(self):
        """(features [count x 12], measured ms [count]) of the last profiled fused run."""
        cnt = C.c_int64()
        _check(lib().qsv_timing_log(self.h, None, None, 0, C.byref(cnt)))
        f = np.zeros((cnt.value, CM_FEATURES), np.float32)
        ms = np.zeros(cnt.value, np.float32)
        _check(lib().qsv_timing_log(self.h, _fptr(f), _fptr(ms), cnt.value, C.byref(cnt)))
        return f, ms

    def copy_from(self, other: "StateVector"):
        _check(lib().qsv_copy_state(self.h, other.h))

    def permute_qubits(self, perm) -> None:
        """Move qubit q to position perm[q] (complex64, >= 13 qubits; bit-exact data movement)."""
        p = np.ascontiguousarray(perm, dtype=np.int32)
        if p.shape != (self.n,):
            raise ValueError("perm must list a position for every qubit")
        _check(lib().qsv_permute_qubits(self.h, p.ctypes.data_as(C.POINTER(C.c_int))))

    def inner(self, other: "StateVector") -> complex:
        re, im = C.c_double(), C.c_double()
        _check(lib().qsv_inner(self.h, other.h, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def fidelity(self, other: "StateVector") -> float:
        """|<self|other>|^2 / (<self|self> <other|other>), all in FP64 on the device."""
        z = self.inner(other)
        return abs(z) ** 2 / (self.norm2() * other.norm2())

    def device_ptr(self) -> int:
        p = C.c_void_p()
        _check(lib().qsv_device_ptr(self.h, C.byref(p)))
        return p.value

    def stream(self) -> int:
        p = C.c_void_p()
        _check(lib().qsv_stream(self.h, C.byref(p)))
        return p.value or 0

    # ---- gates ----
    def h_gate(self, t, ctrl=0, exact=False): _check(lib().qsv_h(self.h, t, ctrl, int(exact)))
    def x(self, t, ctrl=0): _check(lib().qsv_x(self.h, t, ctrl))
    def u1(self, t, u, ctrl=0): _check(lib().qsv_apply_1q(self.h, t, _mat(u), ctrl))
    def diag(self, t, d0, d1, ctrl=0):
        _check(lib().qsv_diag(self.h, t, np.array([d0.real, d0.imag, d1.real, d1.imag]), ctrl))
    def rx(self, t, th, dagger=0, ctrl=0): _check(lib().qsv_rx(self.h, t, th, dagger, ctrl))
    def ry(self, t, th, dagger=0, ctrl=0): _check(lib().qsv_ry(self.h, t, th, dagger, ctrl))
    def rz(self, t, th, dagger=0, ctrl=0): _check(lib().qsv_rz(self.h, t, th, dagger, ctrl))
    def cphase(self, a, b, z, ctrl=0): _check(lib().qsv_cphase(self.h, a, b, z.real, z.imag, ctrl))
    def cr(self, a, b, th, dagger=0): _check(lib().qsv_cr(self.h, a, b, th, dagger))
    def iswap(self, a, b, dagger=0): _check(lib().qsv_iswap(self.h, a, b, dagger))
    def sqiswap(self, a, b, dagger=0, exact=False): _check(lib().qsv_sqiswap(self.h, a, b, dagger, int(exact)))
    def cswap(self, c, a, b): _check(lib().qsv_cswap(self.h, c, a, b))
    def u2(self, qa, qb, u): _check(lib().qsv_apply_2q(self.h, qa, qb, _mat(u)))

    def block(self, qubits, u, split=2):
        _check(lib().qsv_apply_block(self.h, _ints(qubits), len(qubits), _mat(u), int(split)))

    # ---- readout ----
    def prob0(self, q) -> float:
        v = C.c_double()
        _check(lib().qsv_prob0(self.h, q, C.byref(v)))
        return v.value

    def norm2(self) -> float:
        v = C.c_double()
        _check(lib().qsv_norm2(self.h, C.byref(v)))
        return v.value

    def expect_zz_chain(self) -> float:
        v = C.c_double()
        _check(lib().qsv_expect_zz_chain(self.h, C.byref(v)))
        return v.value

    def readout_noalloc(self, what: int, out_ptr: int, q: int = 0) -> None:
        """Queue qsv_readout_noalloc on this state's stream; the FP64 result
        lands in the device double at out_ptr (no allocation in the call)."""
        _check(lib().qsv_readout_noalloc(self.device_ptr(), self.n, self.precision, int(what), int(q),
                                         C.c_void_p(out_ptr), C.c_void_p(self.stream())))

    def collapse(self, q, outcome, scale):
        _check(lib().qsv_collapse(self.h, q, outcome, scale))

    def measure(self, q, r) -> int:
        o = C.c_int()
        _check(lib().qsv_measure(self.h, q, r, C.byref(o)))
        return o.value

    def marginals(self, qubits) -> np.ndarray:
        out = np.empty(1 << len(qubits), np.float64)
        _check(lib().qsv_marginals(self.h, _ints(qubits), len(qubits), out))
        return out

    def pmeasure(self, qubits):
        idx = np.zeros(10, np.uint64)
        prob = np.zeros(10, np.float64)
        cnt = C.c_int()
        _check(lib().qsv_pmeasure(self.h, _ints(qubits), len(qubits), idx, prob, C.byref(cnt)))
        return list(zip(idx[:cnt.value].tolist(), prob[:cnt.value].tolist()))

    def gather(self, idx) -> np.ndarray:
        """Amplitudes at the given basis indices (complex128)."""
        idx = np.ascontiguousarray(idx, np.uint64)
        out = np.zeros(2 * max(1, idx.size), np.float64)
        _check(lib().qsv_gather(self.h, idx if idx.size else np.zeros(1, np.uint64), idx.size, out))
        return out[: 2 * idx.size].view(np.complex128)

    def sample_shots(self, qubits, gate_draws: int, rng_state: int, shots: int, method: str = "tree"):
        """Reference shot loop from the current state (which is left unchanged):
        each shot skips `gate_draws` draws, then measures `qubits` in order.
        method "tree": prefix-sum sampler (no state copy per shot); "copy": copy
        and measure per shot, as the reference's loop.  Both give the same
        outcomes.  Returns (words, rng_state, ties)."""
        out = np.zeros(max(shots, 1), np.uint64)
        rs = C.c_int32(rng_state)
        ties = C.c_int64(0)
        fn = {"tree": lib().qsv_sample_shots, "copy": lib().qsv_sample_shots_copy}[method]
        _check(fn(self.h, _ints(qubits), len(qubits), gate_draws, C.byref(rs), shots, out, C.byref(ties)))
        return out[:shots], rs.value, ties.value

    def run(self, circ: "_circ.Circuit", exact=False, fuse_k=0, split=2, scheduler=0,
            rng_state=None, outcomes=None, profile=False, tile=0, group=0, pass_tb=0,
            pass_kernel=0, fixed_low=False) -> RunStats:
        """scheduler: 0 greedy blocks (one sweep each), 1 learned (needs an attached
        cost model), 2 greedy blocks grouped into multi-block sweeps of <= group,
        3 window planner (one sweep per window of pass_tb tile bits, 0 = 13).
        pass_kernel: 0 or 1 (one workgroup per tile).  fixed_low: window
        schedules keep qubits 0..3 in every sweep (True), relabel qubits
        between sweeps from 28 qubits (False) or from 14 qubits (-1)."""
        g = np.ascontiguousarray(circ.gates)
        m = np.ascontiguousarray(circ.mats, np.complex128)
        opts = RunOpts(int(exact), int(fuse_k), int(split), int(scheduler), int(profile), int(tile),
                       int(group), int(pass_tb), int(pass_kernel), int(fixed_low))
        st = RunStats()
        rs = C.c_int32(rng_state if rng_state is not None else 1)
        self._last_rng = None
        _check(lib().qsv_run_circuit(self.h, g.ctypes.data_as(C.c_void_p), g.shape[0],
                                     m.ctypes.data_as(C.c_void_p) if m.size else None,
                                     C.byref(opts), C.byref(rs) if rng_state is not None else None,
                                     outcomes.ctypes.data_as(C.c_void_p) if outcomes is not None else None,
                                     C.byref(st)))
        self._last_rng = rs.value
        return st

    @property
    def last_rng(self):
        return self._last_rng


def _fptr(a):
    return a.ctypes.data_as(C.c_void_p)


def block_features(n, kq, split, tile, qubits) -> np.ndarray:
    f = np.zeros(CM_FEATURES, np.float32)
    _check(lib().qsv_cm_block_features(n, kq, int(split), tile, _ints(qubits), _fptr(f)))
    return f


class CostModel:
    """The fused-block cost model: an MLP whose forward, backward and Adam
    steps run as HIP kernels on the device."""

    def __init__(self, seed: int = 1):
        h = C.c_void_p()
        _check(lib().qsv_cm_create(seed, C.byref(h)))
        self.h = h

    @staticmethod
    def nparams() -> int:
        return lib().qsv_cm_nparams()

    def close(self):
        if self.h:
            lib().qsv_cm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.nparams(), np.float32)
        _check(lib().qsv_cm_get_params(self.h, _fptr(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams()
        _check(lib().qsv_cm_set_params(self.h, _fptr(p)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, CM_FEATURES)
        y = np.zeros(x.shape[0], np.float32)
        _check(lib().qsv_cm_forward(self.h, _fptr(x), x.shape[0], _fptr(y)))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, CM_FEATURES)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.nparams(), np.float32)
        loss = C.c_double()
        _check(lib().qsv_cm_gradients(self.h, _fptr(x), _fptr(t), x.shape[0], _fptr(g), C.byref(loss)))
        return g, loss.value

    def train(self, x, t, epochs: int, lr: float = 3e-3) -> float:
        x = np.ascontiguousarray(x, np.float32).reshape(-1, CM_FEATURES)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_double()
        _check(lib().qsv_cm_train(self.h, _fptr(x), _fptr(t), x.shape[0], epochs, lr, C.byref(loss)))
        return loss.value

    def save(self, path: str = DEFAULT_COSTMODEL):
        self.get_params().tofile(path)

    @classmethod
    def load(cls, path: str = DEFAULT_COSTMODEL) -> "CostModel":
        m = cls(1)
        m.set_params(np.fromfile(path, np.float32))
        return m


class Batch:
    """Many hardware-efficient-ansatz instances evaluated in one persistent launch
    (see qsv_batch_* in include/qsv.h).  angles: list of (layers, n, 2) arrays."""

    def __init__(self, nq, layers, angles):
        self.count = len(nq)
        nq = np.ascontiguousarray(nq, np.int32)
        ly = np.ascontiguousarray(layers, np.int32)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float64).ravel() for a in angles]))
        h = C.c_void_p()
        _check(lib().qsv_batch_create(_fptr(nq), _fptr(ly), self.count, _fptr(flat), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().qsv_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self):
        t, ms = C.c_int(), C.c_double()
        _check(lib().qsv_batch_autotune(self.h, C.byref(t), C.byref(ms)))
        return t.value, ms.value

    def run(self, threads: int = 0):
        zz = np.zeros(self.count, np.float64)
        lat = np.zeros(self.count, np.float64)
        ms = C.c_double()
        _check(lib().qsv_batch_run(self.h, threads, _fptr(zz), _fptr(lat), C.byref(ms)))
        return zz, lat, ms.value

    def graph_run(self, replays: int = 3):
        zz = np.zeros(self.count, np.float64)
        ms = C.c_double()
        _check(lib().qsv_batch_graph_run(self.h, replays, _fptr(zz), C.byref(ms)))
        return zz, ms.value


def hea_batch(count: int, seed: int, nmin: int = 10, nmax: int = 22, lmin: int = 1, lmax: int = 8):
    """Seeded configuration: qubits uniform in [nmin, nmax], layers uniform in
    [lmin, lmax], angles uniform in [-pi, pi]."""
    rng = _circ.Xoshiro256(seed)
    nq, ly, ang = [], [], []
    for _ in range(count):
        n = nmin + rng.below(nmax - nmin + 1)
        L = lmin + rng.below(lmax - lmin + 1)
        a = np.array([[[np.pi * (2 * rng.uniform() - 1) for _ in range(2)] for _ in range(n)]
                      for _ in range(L)]).reshape(L, n, 2)
        nq.append(n)
        ly.append(L)
        ang.append(a)
    return nq, ly, ang


def hea_circuit(n: int, angles) -> "_circ.Circuit":
    """The gate list of one batch instance (RY then RZ per qubit, CNOT ladder)."""
    b = _circ._Builder(n)
    for layer in np.asarray(angles).reshape(-1, n, 2):
        for q in range(n):
            b.add(_circ.RY, q, theta=float(layer[q, 0]))
            b.add(_circ.RZ, q, theta=float(layer[q, 1]))
        for q in range(n - 1):
            b.add(_circ.CNOT, q, q + 1)
    return b.build(f"hea_n{n}")


def fuse_plan(circ: "_circ.Circuit", max_k: int, exact: bool = False):
    """Greedy fusion plan (host code; no device needed).  Returns a list of
    (qubits, matrix, ngates) blocks, or (None, None, record_index) for
    MEASURE/RESET records, in emission order."""
    g = np.ascontiguousarray(circ.gates)
    m = np.ascontiguousarray(circ.mats, np.complex128)
    cap = max(1, g.shape[0])
    ks = np.zeros(cap, np.int32)
    qs = np.zeros(cap * 5, np.int32)
    mats = np.zeros(cap * 2048, np.float64)
    ngs = np.zeros(cap, np.int64)
    nb = C.c_int64()
    _check(lib().qsv_fuse_plan(g.ctypes.data_as(C.c_void_p), g.shape[0],
                               m.ctypes.data_as(C.c_void_p) if m.size else None, max_k, int(exact), cap,
                               ks.ctypes.data_as(C.c_void_p), qs.ctypes.data_as(C.c_void_p),
                               mats.ctypes.data_as(C.c_void_p), ngs.ctypes.data_as(C.c_void_p),
                               C.byref(nb)))
    out = []
    for b in range(nb.value):
        k = int(ks[b])
        if k == 0:
            out.append((None, None, int(ngs[b])))
            continue
        d = 1 << k
        u = mats[b * 2048: b * 2048 + 2 * d * d].view(np.complex128).reshape(d, d).copy()
        out.append((qs[b * 5: b * 5 + k].tolist(), u, int(ngs[b])))
    return out


def window_plan(circ: "_circ.Circuit", max_k: int, tb: int = 13, exact: bool = False, beam=None):
    """Window plan (host code; no device needed): list of (qubits, matrix,
    ngates, window) blocks, or (None, None, record_index, window) for
    MEASURE/RESET records, in emission order.  beam = (width, branch, first,
    delay[, fuse_tries[, relabel]]) plans with the beam search
    (qsv_window_plan_beam)."""
    g = np.ascontiguousarray(circ.gates)
    m = np.ascontiguousarray(circ.mats, np.complex128)
    cap = max(1, g.shape[0])
    ks = np.zeros(cap, np.int32)
    qs = np.zeros(cap * 5, np.int32)
    mats = np.zeros(cap * 2048, np.float64)
    ngs = np.zeros(cap, np.int64)
    win = np.zeros(cap, np.int32)
    nb = C.c_int64()
    outs = (cap, ks.ctypes.data_as(C.c_void_p), qs.ctypes.data_as(C.c_void_p), mats.ctypes.data_as(C.c_void_p),
            ngs.ctypes.data_as(C.c_void_p), win.ctypes.data_as(C.c_void_p), C.byref(nb))
    mp = m.ctypes.data_as(C.c_void_p) if m.size else None
    if beam is None:
        _check(lib().qsv_window_plan(g.ctypes.data_as(C.c_void_p), g.shape[0], mp, max_k, int(exact), tb,
                                     circ.n, *outs))
    else:
        width, branch, first, delay, tries, relabel = (tuple(int(v) for v in beam) + (0, 0))[:6]
        _check(lib().qsv_window_plan_beam(g.ctypes.data_as(C.c_void_p), g.shape[0], mp, max_k, int(exact), tb,
                                          circ.n, width, branch, first, delay, tries, relabel, *outs))
    out = []
    for b in range(nb.value):
        k = int(ks[b])
        if k == 0:
            out.append((None, None, int(ngs[b]), int(win[b])))
            continue
        d = 1 << k
        u = mats[b * 2048: b * 2048 + 2 * d * d].view(np.complex128).reshape(d, d).copy()
        out.append((qs[b * 5: b * 5 + k].tolist(), u, int(ngs[b]), int(win[b])))
    return out


def device_count() -> int:
    c = C.c_int()
    _check(lib().qsv_device_count(C.byref(c)))
    return c.value


def pm_next(s):
    return lib().qsv_pm_next(s)


def pm_jump(s, k):
    return lib().qsv_pm_jump(s, k)
