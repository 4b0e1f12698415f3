"""Seeded circuit generators and the circuit interchange format.

A circuit is a numpy structured array of 32-byte gate records (the layout of
``qsv_gate`` in include/qsv.h) plus a matrix pool of complex128 entries used by
dense U1/U2 records.  The same records drive the engine, the CPU oracle and the
reference driver, so every path sees identical gates and angles.

Randomness: SplitMix64 seeding a xoshiro256** stream (Vigna), so the sequence
is defined by the seed alone and independent of numpy's generators.

Workloads (gate distributions of the benchmark configurations):
  * ``random_layers``: per layer every qubit gets one of H/RX/RY/RZ (uniform),
    angle uniform in [0, 2*pi); then a Fisher-Yates perfect pairing of the
    qubits, each pair a CNOT (control = first) or CZ with equal odds, or with
    probability ``haar_frac`` a Haar-random two-qubit unitary (Mezzadri's QR of
    a complex Ginibre matrix).  With odd n one qubit per layer stays unpaired.
  * ``ghz``: H(0) then CNOT(i, i+1).
  * ``qft``: for i = n-1..0: H(i), then CR(i, j, 2*pi / 2^(i-j+1)) for j < i, no
    final swaps (the order of the inverse QFT in
    testMain/testCPUQuantumGates.cpp:756-763, forward angles).
  * ``hea``: hardware-efficient ansatz layers of RY, RZ on every qubit plus a
    CNOT ladder (i, i+1).
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field

import numpy as np

H, X, RX, RY, RZ, CNOT, CZ, CR, ISWAP, SQISWAP, TOFFOLI, CSWAP, U1, U2, MEASURE, RESET = range(1, 17)
KIND_NAMES = {H: "H", X: "X", RX: "RX", RY: "RY", RZ: "RZ", CNOT: "CNOT", CZ: "CZ", CR: "CR",
              ISWAP: "ISWAP", SQISWAP: "SQISWAP", TOFFOLI: "TOFFOLI", CSWAP: "CSWAP", U1: "U1",
              U2: "U2", MEASURE: "MEASURE", RESET: "RESET"}

GATE_DTYPE = np.dtype([("kind", "<i4"), ("dagger", "<i4"), ("q0", "<i4"), ("q1", "<i4"),
                       ("q2", "<i4"), ("mat", "<i4"), ("theta", "<f8")])
assert GATE_DTYPE.itemsize == 32

MASK64 = (1 << 64) - 1


class Xoshiro256:
    """xoshiro256** seeded through SplitMix64 (both by S. Vigna, public domain)."""

    def __init__(self, seed: int):
        x = seed & MASK64
        s = []
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & MASK64
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
            s.append(z ^ (z >> 31))
        self.s = s

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & MASK64

    def next_u64(self) -> int:
        s = self.s
        result = (self._rotl((s[1] * 5) & MASK64, 7) * 9) & MASK64
        t = (s[1] << 17) & MASK64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self) -> float:
        """Double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def below(self, k: int) -> int:
        """Integer in [0, k) (multiply-shift; bias < 2^-32 for the k used here)."""
        return ((self.next_u64() >> 32) * k) >> 32

    def normal(self) -> float:
        """Standard normal by Box-Muller (one value per call)."""
        u1 = 1.0 - self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


@dataclass
class Circuit:
    n: int
    gates: np.ndarray
    mats: np.ndarray = field(default_factory=lambda: np.zeros(0, np.complex128))
    name: str = ""

    @property
    def ngates(self) -> int:
        return int(self.gates.shape[0])

    def counts(self) -> dict:
        kinds, c = np.unique(self.gates["kind"], return_counts=True)
        return {KIND_NAMES[int(k)]: int(v) for k, v in zip(kinds, c)}

    def prefix(self, g: int) -> "Circuit":
        return Circuit(self.n, self.gates[:g].copy(), self.mats, self.name + f"[:{g}]")

    def save(self, path: str) -> None:
        """QSVC file: magic, int32 n, int64 ngates, int64 nmat, records, matrices."""
        with open(path, "wb") as f:
            f.write(b"QSVC")
            f.write(struct.pack("<iqq", self.n, self.ngates, int(self.mats.shape[0])))
            f.write(np.ascontiguousarray(self.gates).tobytes())
            f.write(np.ascontiguousarray(self.mats, dtype=np.complex128).tobytes())

    @staticmethod
    def load(path: str) -> "Circuit":
        with open(path, "rb") as f:
            assert f.read(4) == b"QSVC"
            n, ng, nm = struct.unpack("<iqq", f.read(20))
            gates = np.frombuffer(f.read(32 * ng), dtype=GATE_DTYPE).copy()
            mats = np.frombuffer(f.read(16 * nm), dtype=np.complex128).copy()
        return Circuit(n, gates, mats)


class _Builder:
    def __init__(self, n: int):
        self.n = n
        self.rows = []
        self.mats = []
        self.nmat = 0

    def add(self, kind, q0=0, q1=0, q2=0, theta=0.0, dagger=0, mat=None):
        mi = -1
        if mat is not None:
            mi = self.nmat
            flat = np.asarray(mat, np.complex128).reshape(-1)
            self.mats.append(flat)
            self.nmat += flat.size
        self.rows.append((kind, dagger, q0, q1, q2, mi, theta))

    def build(self, name="") -> Circuit:
        g = np.array(self.rows, dtype=GATE_DTYPE) if self.rows else np.zeros(0, GATE_DTYPE)
        m = np.concatenate(self.mats) if self.mats else np.zeros(0, np.complex128)
        return Circuit(self.n, g, m, name)


def haar_unitary(rng: Xoshiro256, d: int) -> np.ndarray:
    """Haar-random d x d unitary: QR of a complex Ginibre matrix with the
    phases of R's diagonal folded into Q (Mezzadri, Notices AMS 54, 2007)."""
    z = np.array([[complex(rng.normal(), rng.normal()) for _ in range(d)] for _ in range(d)])
    z /= math.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph[None, :]


def random_layers(n: int, depth: int, seed: int, haar_frac: float = 0.0,
                  measure_all: bool = False) -> Circuit:
    rng = Xoshiro256(seed)
    b = _Builder(n)
    for _ in range(depth):
        for q in range(n):
            k = rng.below(4)
            theta = 2.0 * math.pi * rng.uniform()
            if k == 0:
                b.add(H, q)
            else:
                b.add((RX, RY, RZ)[k - 1], q, theta=theta)
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = rng.below(i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for p in range(n // 2):
            a, c = perm[2 * p], perm[2 * p + 1]
            if haar_frac > 0.0 and rng.uniform() < haar_frac:
                b.add(U2, a, c, mat=haar_unitary(rng, 4))
            elif rng.below(2) == 0:
                b.add(CNOT, a, c)
            else:
                b.add(CZ, a, c)
    if measure_all:
        for q in range(n):
            b.add(MEASURE, q, q)
    return b.build(f"random_n{n}_d{depth}_s{seed}")


def ghz(n: int) -> Circuit:
    b = _Builder(n)
    b.add(H, 0)
    for i in range(n - 1):
        b.add(CNOT, i, i + 1)
    return b.build(f"ghz_n{n}")


def qft(n: int) -> Circuit:
    b = _Builder(n)
    for i in range(n - 1, -1, -1):
        b.add(H, i)
        for j in range(i - 1, -1, -1):
            b.add(CR, i, j, theta=2.0 * math.pi / (1 << (i - j + 1)))
    return b.build(f"qft_n{n}")


def hea(n: int, layers: int, seed: int) -> Circuit:
    rng = Xoshiro256(seed)
    b = _Builder(n)
    for _ in range(layers):
        for q in range(n):
            b.add(RY, q, theta=math.pi * (2.0 * rng.uniform() - 1.0))
            b.add(RZ, q, theta=math.pi * (2.0 * rng.uniform() - 1.0))
        for q in range(n - 1):
            b.add(CNOT, q, q + 1)
    return b.build(f"hea_n{n}_l{layers}_s{seed}")


def to_qrunes(circ: Circuit, measure: bool = True) -> str:
    """QRunes text of a circuit over the QRunes gate set (QRunes/README.md;
    tokens of QRunesParser/lex.yy.cpp:914-1079): H, NOT, RX/RY/RZ, CNOT, CR
    (CZ is written as CR with angle pi), ISWAP, SQISWAP, TOFFOLI, with
    MEASURE q,$q on every qubit when `measure`.  Angles are written as fixed
    point decimals with 25 fractional digits, which read back as the same
    doubles; daggered records use a DAGGER block."""
    out = [f"QINIT {circ.n}", f"CREG {circ.n}"]
    for g in circ.gates:
        k, q0, q1, q2, th = int(g["kind"]), int(g["q0"]), int(g["q1"]), int(g["q2"]), float(g["theta"])
        ang = f'"{th:.25f}"'
        if k == H:
            line = f"H {q0}"
        elif k == X:
            line = f"NOT {q0}"
        elif k in (RX, RY, RZ):
            line = f"{KIND_NAMES[k]} {q0},{ang}"
        elif k == CNOT:
            line = f"CNOT {q0},{q1}"
        elif k == CZ:
            line = f'CR {q0},{q1},"pi"'
        elif k == CR:
            line = f"CR {q0},{q1},{ang}"
        elif k == ISWAP:
            line = f'ISWAP {q0},{q1},"0"'
        elif k == SQISWAP:
            line = f"SQISWAP {q0},{q1}"
        elif k == TOFFOLI:
            line = f"TOFFOLI {q0},{q1},{q2}"
        else:
            raise ValueError(f"{KIND_NAMES.get(k, k)} has no QRunes statement")
        out += ["DAGGER", line, "ENDDAGGER"] if int(g["dagger"]) else [line]
    if measure:
        out += [f"MEASURE {q},${q}" for q in range(circ.n)]
    return "\n".join(out) + "\n"


def random_gaussian_state(n: int, seed: int) -> np.ndarray:
    """Normalised state with i.i.d. Gaussian real and imaginary parts, from a
    seeded numpy PCG64 stream (vectorised; 2^26 amplitudes in about a second)."""
    g = np.random.Generator(np.random.PCG64(seed))
    psi = g.standard_normal(1 << n) + 1j * g.standard_normal(1 << n)
    psi /= np.sqrt(np.vdot(psi, psi).real)
    return psi


def all_gate_kinds(n: int, seed: int, reps: int = 3, u1: bool = False) -> Circuit:
    """Every uncontrolled gate kind, both dagger settings, random qubits/angles.
    u1: also general 2x2 records (U1, a Haar-random unitary each); the
    reference cannot run them (its unitarySingleQubitGate addresses qubit
    qn - 1, CPUQuantumGates.cpp:1188), so the fixtures made by the reference
    leave them out."""
    rng = Xoshiro256(seed)
    b = _Builder(n)

    def pick(k):
        qs = list(range(n))
        out = []
        for _ in range(k):
            out.append(qs.pop(rng.below(len(qs))))
        return out

    for _ in range(reps):
        for q in range(n):
            b.add(H, q)
        for dag in (0, 1):
            for kind in (RX, RY, RZ):
                b.add(kind, pick(1)[0], theta=2 * math.pi * rng.uniform(), dagger=dag)
            a, c = pick(2)
            b.add(CR, a, c, theta=2 * math.pi * rng.uniform(), dagger=dag)
            a, c = pick(2)
            b.add(ISWAP, a, c, dagger=dag)
            a, c = pick(2)
            b.add(SQISWAP, a, c, dagger=dag)
        b.add(X, pick(1)[0])
        a, c = pick(2)
        b.add(CNOT, a, c)
        a, c = pick(2)
        b.add(CZ, a, c)
        a, c, t = pick(3)
        b.add(TOFFOLI, a, c, t)
        a, c, t = pick(3)
        b.add(CSWAP, a, c, t)
        if u1:
            for _ in range(2):
                b.add(U1, pick(1)[0], mat=haar_unitary(rng, 2))
    return b.build(f"allkinds_n{n}_s{seed}" + ("_u1" if u1 else ""))
