"""ctypes wrapper of liboracle.so, the CPU restatement of QPanda's CPU simulator.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker.  The engine never imports
this module.  Each function names the reference loop it restates (paths
relative to the QPanda source tree).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def build_ref(ref_root: str = "/root/reference") -> bool:
    """Compile the reference CPU simulator into oracle/_ref (only where the
    reference tree exists)."""
    src = os.path.join(ref_root, "QPandaSDK", "QuantumInstructionHandle")
    if not os.path.isdir(src):
        return False
    subprocess.check_call(["make", "-s", "-C", HERE, "ref", f"REF={src}"])
    return True


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        I = C.c_int
        D = C.c_double
        L.orc_pm_next.restype = C.c_int32
        L.orc_pm_next.argtypes = [C.c_int32]
        L.orc_pm_jump.restype = C.c_int32
        L.orc_pm_jump.argtypes = [C.c_int32, C.c_uint64]
        L.orc_pm_seed_from_tm.restype = C.c_int32
        L.orc_pm_seed_from_tm.argtypes = [I] * 6
        L.orc_init_zero.argtypes = [P, I]
        L.orc_h.argtypes = [P, I, I, I]
        L.orc_x.argtypes = [P, I, I]
        for nm in ("orc_rx", "orc_ry", "orc_rz"):
            getattr(L, nm).argtypes = [P, I, I, D, I]
        L.orc_u1.argtypes = [P, I, I, P]
        L.orc_ctrl_1q.argtypes = [P, I, C.POINTER(C.c_int), I, I, D, I, C.c_void_p, I]
        L.orc_cnot.argtypes = [P, I, I, I]
        L.orc_cr.argtypes = [P, I, I, I, D, I]
        L.orc_cz_exact.argtypes = [P, I, I, I]
        L.orc_iswap.argtypes = [P, I, I, I, I]
        L.orc_sqiswap.argtypes = [P, I, I, I, I, I]
        L.orc_toffoli.argtypes = [P, I, I, I, I]
        L.orc_cswap.argtypes = [P, I, I, I, I]
        L.orc_block.argtypes = [P, I, C.POINTER(C.c_int), I, P]
        L.orc_u2.argtypes = [P, I, I, I, P]
        L.orc_prob0.restype = D
        L.orc_prob0.argtypes = [P, I, I]
        L.orc_measure.restype = I
        L.orc_measure.argtypes = [P, I, I, D]
        L.orc_marginals.argtypes = [P, I, C.POINTER(C.c_int), I, P]
        L.orc_norm2.restype = D
        L.orc_norm2.argtypes = [P, I]
        L.orc_expect_zz_chain.restype = D
        L.orc_expect_zz_chain.argtypes = [P, I]
        L.orc_run.restype = I
        L.orc_run.argtypes = [P, I, C.c_void_p, C.c_int64, C.c_void_p, I,
                              C.POINTER(C.c_int32), C.c_void_p]
        PD = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.orc_mlp_nparams.restype = C.c_int
        L.orc_mlp_forward.argtypes = [PD, PD, C.c_int64, PD]
        L.orc_mlp_loss.restype = C.c_double
        L.orc_mlp_loss.argtypes = [PD, PD, PD, C.c_int64]
        L.orc_mlp_grad.restype = C.c_double
        L.orc_mlp_grad.argtypes = [PD, PD, PD, C.c_int64, PD]
        L.orc_sample_shots.argtypes = [P, I, C.POINTER(C.c_int), I, C.c_int64,
                                       C.POINTER(C.c_int32), C.c_int64,
                                       np.ctypeslib.ndpointer(dtype=np.uint64), P]
        _lib = L
    return _lib


def _ints(xs):
    arr = (C.c_int * len(xs))(*[int(x) for x in xs])
    return arr


class State:
    """FP64 state on the host, interleaved (re, im) like std::complex<double>."""

    def __init__(self, n: int, exact: bool = False):
        self.n = n
        self.exact = int(bool(exact))
        self.buf = np.zeros(2 << n, np.float64)
        lib().orc_init_zero(self.buf, n)

    @property
    def psi(self) -> np.ndarray:
        return self.buf.view(np.complex128)

    def set(self, amps) -> None:
        self.buf[:] = np.ascontiguousarray(amps, np.complex128).view(np.float64)

    # --- gates (CPUQuantumGates.cpp loops) ---
    def h(self, q): lib().orc_h(self.buf, self.n, q, self.exact)
    def x(self, q): lib().orc_x(self.buf, self.n, q)
    def rx(self, q, t, dagger=0): lib().orc_rx(self.buf, self.n, q, t, dagger)
    def ry(self, q, t, dagger=0): lib().orc_ry(self.buf, self.n, q, t, dagger)
    def rz(self, q, t, dagger=0): lib().orc_rz(self.buf, self.n, q, t, dagger)
    def cnot(self, c, t): lib().orc_cnot(self.buf, self.n, c, t)
    def cr(self, a, b, t, dagger=0): lib().orc_cr(self.buf, self.n, a, b, t, dagger)
    def cz_exact(self, a, b): lib().orc_cz_exact(self.buf, self.n, a, b)
    def iswap(self, a, b, dagger=0): lib().orc_iswap(self.buf, self.n, a, b, dagger)
    def sqiswap(self, a, b, dagger=0): lib().orc_sqiswap(self.buf, self.n, a, b, dagger, self.exact)
    def toffoli(self, a, b, t): lib().orc_toffoli(self.buf, self.n, a, b, t)
    def cswap(self, c, a, b): lib().orc_cswap(self.buf, self.n, c, a, b)

    def u1(self, q, u):
        lib().orc_u1(self.buf, self.n, q, np.ascontiguousarray(u, np.complex128).view(np.float64).ravel())

    def u2(self, qa, qb, u):
        lib().orc_u2(self.buf, self.n, qa, qb, np.ascontiguousarray(u, np.complex128).view(np.float64).ravel())

    def block(self, qubits, u):
        k = len(qubits)
        lib().orc_block(self.buf, self.n, _ints(qubits), k,
                        np.ascontiguousarray(u, np.complex128).view(np.float64).ravel())

    _CTRL_OPS = {"H": 0, "X": 1, "RX": 2, "RY": 3, "RZ": 4, "U1": 5}

    def controlled(self, op: str, ctrl_then_target, theta=0.0, dagger=0, u=None):
        """Controlled form with the reference's list convention: controls, then
        the target last (Instruction.cpp:62)."""
        lst = list(ctrl_then_target)
        ub = None
        if u is not None:
            arr = np.ascontiguousarray(u, np.complex128).view(np.float64).ravel()
            ub = arr.ctypes.data_as(C.c_void_p)
        lib().orc_ctrl_1q(self.buf, self.n, _ints(lst), len(lst), self._CTRL_OPS[op], theta,
                          dagger, ub, self.exact)

    # --- readout ---
    def prob0(self, q) -> float:
        return lib().orc_prob0(self.buf, self.n, q)

    def measure(self, q, r) -> int:
        return lib().orc_measure(self.buf, self.n, q, r)

    def marginals(self, qubits) -> np.ndarray:
        out = np.zeros(1 << len(qubits), np.float64)
        lib().orc_marginals(self.buf, self.n, _ints(qubits), len(qubits), out)
        return out

    def norm2(self) -> float:
        return lib().orc_norm2(self.buf, self.n)

    def expect_zz_chain(self) -> float:
        return lib().orc_expect_zz_chain(self.buf, self.n)

    def run(self, circ, rng_state=None, outcomes=None) -> int:
        """Apply a qpanda_amd.circuits.Circuit; returns the updated RNG state."""
        g = np.ascontiguousarray(circ.gates)
        m = np.ascontiguousarray(circ.mats, np.complex128)
        rs = C.c_int32(rng_state if rng_state is not None else 1)
        outp = None
        if outcomes is not None:
            outp = outcomes.ctypes.data_as(C.c_void_p)
        rc = lib().orc_run(self.buf, self.n, g.ctypes.data_as(C.c_void_p), g.shape[0],
                           m.ctypes.data_as(C.c_void_p) if m.size else None, self.exact,
                           C.byref(rs) if rng_state is not None else None, outp)
        if rc != 0:
            raise ValueError("oracle: bad gate record")
        return rs.value


def mlp_forward(params, x):
    """FP64 forward pass of the cost-model network (oracle/mlp_oracle.c)."""
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 12)
    y = np.zeros(x.shape[0], np.float64)
    lib().orc_mlp_forward(p, x.ravel(), x.shape[0], y)
    return y


def mlp_loss(params, x, t):
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 12)
    return lib().orc_mlp_loss(p, x.ravel(), np.ascontiguousarray(t, np.float64), x.shape[0])


def mlp_grad(params, x, t):
    """(gradient, loss) of 0.5 * mean((y - t)^2), FP64 chain rule."""
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 12)
    g = np.zeros(p.size, np.float64)
    loss = lib().orc_mlp_grad(p, x.ravel(), np.ascontiguousarray(t, np.float64), x.shape[0], g)
    return g, loss


def adam_update(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step with bias correction (Kingma & Ba, ICLR 2015, Algorithm 1),
    FP64: returns the new (p, m, v).  `step` counts from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mhat = m / (1 - b1 ** step)
    vhat = v / (1 - b2 ** step)
    return p - lr * mhat / (np.sqrt(vhat) + eps), m, v


def mlp_train(params, x, t, epochs, lr):
    """Full-batch Adam on 0.5 * mean((y - t)^2) from zero moments: the training
    loop of the cost model, restated in FP64.  Returns the parameters."""
    p = np.array(params, np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for step in range(1, epochs + 1):
        g, _ = mlp_grad(p, x, t)
        p, m, v = adam_update(p, g, m, v, step, lr)
    return p


def pm_next(s: int) -> int:
    return lib().orc_pm_next(s)


def pm_jump(s: int, k: int) -> int:
    return lib().orc_pm_jump(s, k)


def pm_seed_from_tm(y, mo, d, h, mi, se) -> int:
    return lib().orc_pm_seed_from_tm(y, mo, d, h, mi, se)


def sample_shots(final_state: np.ndarray, n: int, mq, gate_draws: int, rng_state: int,
                 shots: int):
    """Reference shot loop from a fixed pre-measurement state; returns (bits, rng)."""
    final = np.ascontiguousarray(final_state, np.complex128).view(np.float64).ravel()
    out = np.zeros(shots, np.uint64)
    scratch = np.zeros_like(final)
    rs = C.c_int32(rng_state)
    lib().orc_sample_shots(final, n, _ints(mq), len(mq), gate_draws, C.byref(rs), shots, out, scratch)
    return out, rs.value


def run_circuit(circ, exact=False, init=None, rng_state=None):
    st = State(circ.n, exact)
    if init is not None:
        st.set(init)
    rs = st.run(circ, rng_state)
    return st, rs


def ref_available() -> bool:
    return os.path.exists(os.path.join(REF_DIR, "qpanda_ref"))


def ref_run(circ_path: str, out_prefix: str, seed: int, shots: int, omp: bool = False) -> None:
    exe = os.path.join(REF_DIR, "qpanda_ref_omp" if omp else "qpanda_ref")
    subprocess.check_call([exe, "run", circ_path, out_prefix, str(seed), str(shots)])


def _ref_env(threads):
    env = dict(os.environ)
    if threads is not None:
        env["OMP_NUM_THREADS"] = str(threads)
        env["OMP_PROC_BIND"] = "close"
    return env


def ref_time(circ_path: str, stride: int, max_seconds: float, omp: bool = True,
             threads: int | None = None) -> dict:
    """Time QPanda's own CPU simulator on records 0, stride, 2*stride, ... of a
    circuit file until max_seconds has passed (allocation excluded).  omp:
    the OpenMP build with `threads` threads bound close; else the as-shipped
    build (no OpenMP, QPandaSDK/CMakeLists.txt:65-71), one thread."""
    import json
    exe = os.path.join(REF_DIR, "qpanda_ref_omp" if omp else "qpanda_ref")
    out = subprocess.check_output([exe, "time", circ_path, str(stride), str(max_seconds)],
                                  env=_ref_env(threads if omp else None))
    return json.loads(out.decode().strip().splitlines()[-1])


def ref_state_digest(circ_path: str, records: int, omp: bool, threads: int | None = None) -> str:
    """FNV-1a digest of the reference simulator's state after the first
    `records` records: equal digests mean bit-identical states, which checks
    the OpenMP build's racy CNOT / Toffoli loops (CPUQuantumGates.cpp:678/688,
    :1391/1414) against the 1-thread build."""
    import json
    exe = os.path.join(REF_DIR, "qpanda_ref_omp" if omp else "qpanda_ref")
    out = subprocess.check_output([exe, "check", circ_path, str(records)], env=_ref_env(threads if omp else None))
    return json.loads(out.decode().strip().splitlines()[-1])["fnv1a"]


# ---- QRunes programs: executor and run loop -------------------------------

def _qrunes_node(st: "State", nd: dict, rng: int, regs: dict, pm: list):
    """Apply one flattened QRunes instruction the way the reference executor
    does (Instruction.cpp:40-827 dispatch, CPUQuantumGates draws): returns
    (ok, rng).  Every gate call draws once and applies when the draw exceeds
    error_rate = 0; the controlled iSWAP / sqiSWAP overloads are no-ops that
    draw nothing (CPUQuantumGates.cpp:914-922, 1054-1062); a measurement draws
    once; Reset measures and, on 1, applies NOT (one more draw)."""
    k, qs, ctl = nd["kind"], nd["qubits"], list(nd["controls"])
    dag = 1 if nd["dagger"] else 0
    if any(q >= st.n for q in qs):
        return False, rng

    def draw():
        nonlocal rng
        rng = pm_next(rng)
        return rng / 2147483647.0

    if k == "H":
        draw()
        if ctl:
            st.controlled("H", ctl + [qs[0]])
        else:
            st.h(qs[0])
    elif k in ("RX", "RY", "RZ"):
        draw()
        if ctl:
            st.controlled(k, ctl + [qs[0]], theta=nd["angle"], dagger=dag)
        else:
            getattr(st, k.lower())(qs[0], nd["angle"], dag)
    elif k == "NOT":
        draw()
        if ctl:
            st.controlled("X", ctl + [qs[0]])
        else:
            st.x(qs[0])
    elif k == "CNOT":                 # controlled form: NOT on the target (CPUQuantumGates.cpp:750-757)
        draw()
        if ctl:
            st.controlled("X", ctl + [qs[0], qs[1]])
        else:
            st.cnot(qs[0], qs[1])
    elif k == "TOFFOLI":              # Instruction.cpp:378-386
        draw()
        if ctl:
            st.controlled("X", ctl + qs)
        else:
            st.toffoli(qs[0], qs[1], qs[2])
    elif k == "CR":                   # Instruction.cpp:517-526: controls ignored
        draw()
        st.cr(qs[0], qs[1], nd["angle"], dag)
    elif k == "ISWAP":                # Instruction.cpp:558-577: branches swapped
        if ctl:
            draw()
            st.iswap(qs[0], qs[1], dag)
    elif k == "SQISWAP":
        if not ctl:
            draw()
            st.sqiswap(qs[0], qs[1], dag)
    elif k == "RESET":                # CPUQuantumGates.cpp:126-136
        if st.measure(qs[0], draw()):
            draw()
            st.x(qs[0])
    elif k == "MEASURE":
        regs[nd["creg"]] = st.measure(qs[0], draw())
    elif k == "PMEASURE":
        m = st.marginals(qs)
        pm[:] = [len(qs), m]
    else:
        return False, rng
    return True, rng


def qrunes_run(n: int, nodes, repeat: int, rng: int):
    """QPandaSDK::run (QPandaAPI.cpp:123-205) over the oracle state: returns
    (result lines as getResult prints them, final state, generator state).
    PMEASURE results come back as {index: probability} for the caller to
    compare (the reference keeps the top 10 after an unstable sort)."""
    counts = {}
    regs = {}
    pm = []
    st = State(n)
    for run in range(repeat):
        st = State(n)
        for nd in nodes:
            ok, rng = _qrunes_node(st, nd, rng, regs, pm)
            if not ok:
                raise ValueError("oracle: instruction failed")
        if pm:
            break
        key = "".join(str(regs[c]) for c in sorted(regs))[::-1]
        counts[key] = counts.get(key, 0) + 1
    if pm:
        return {i: float(p) for i, p in enumerate(pm[1])}, st, rng
    lines = "".join(f"{k} {_cpp_g(v / repeat)}\n" for k, v in sorted(counts.items()))
    return lines, st, rng


def _cpp_g(x: float) -> str:
    """std::ostream's default formatting of a double (printf %g, 6 digits)."""
    return "%g" % x
