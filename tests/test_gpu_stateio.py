"""State files and whole-state host copies (qsv_io.cpp) on the device.

Bar: binary files are the state's FP64 amplitudes byte for byte (complex64
states: their FP32 values widened exactly); binary and 17-digit text files read
back bit-exactly; the 6-digit text equals the getQState text that the
reference prints (std::ostream default, %g) for the same amplitudes; states of
several 4 Mi-amplitude chunks round-trip; malformed files are rejected."""
import os
import sys

import numpy as np
import pytest

from qpanda_amd import engine as E
from qpanda_amd import sdk

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    v[::7] = 0.0                                   # exact zeros print as "0"
    v[3] = -0.0                                    # and a negative zero as "-0"
    return v / np.linalg.norm(v)


def reference_text(amps, digits=6):
    return "".join(f"state[{i}].real = {a.real:.{digits}g} state[{i}].imag = {a.imag:.{digits}g}\n"
                   for i, a in enumerate(amps))


@pytest.mark.parametrize("n,prec", [(10, E.C128), (10, E.C64), (24, E.C128), (23, E.C64)])
def test_binary_round_trip(tmp_path, n, prec):
    say("binary", n, prec)
    v = random_state(n, n)
    path = str(tmp_path / "s.bin")
    with E.StateVector(n, prec) as a, E.StateVector(n, prec) as b:
        a.set(v)
        want = v if prec == E.C128 else v.astype(np.complex64).astype(np.complex128)
        assert np.array_equal(a.get(), want)               # whole-state copies (chunked for c64)
        a.save(path)
        assert os.path.getsize(path) == 16 << n
        assert np.array_equal(np.fromfile(path, np.complex128), want)
        b.load(path)
        assert np.array_equal(b.get(), want)


def test_text_matches_reference_format_and_round_trips(tmp_path):
    n = 9
    v = random_state(n, 5)
    p6, p17 = str(tmp_path / "s6.txt"), str(tmp_path / "s17.txt")
    with E.StateVector(n) as a, E.StateVector(n) as b:
        a.set(v)
        a.save(p6, text=True)
        with open(p6) as f:
            assert f.read() == reference_text(v)
        a.save(p17, text=True, digits=17)
        b.load(p17, text=True)
        assert np.array_equal(b.get(), v)
        b.load(p6, text=True)                              # 6 digits: within the printed precision
        assert np.abs(b.get() - v).max() < 5e-6 * np.abs(v).max()


def test_multi_chunk_text(tmp_path):
    n = 23                                                 # two 4 Mi-amplitude chunks
    v = random_state(n, 7)
    path = str(tmp_path / "big.txt")
    with E.StateVector(n) as a, E.StateVector(n) as b:
        a.set(v)
        a.save(path, text=True, digits=17)
        b.load(path, text=True)
        assert np.array_equal(b.get(), v)


def test_malformed_files_are_rejected(tmp_path):
    n = 4
    v = random_state(n, 1)
    with E.StateVector(n) as a:
        a.set(v)
        with pytest.raises(E.QsvError):
            a.load(str(tmp_path / "missing.bin"))
        short = tmp_path / "short.bin"
        short.write_bytes(v.tobytes()[:-16])
        with pytest.raises(E.QsvError, match="bytes"):
            a.load(str(short))
        txt = reference_text(v, 17)
        bad = tmp_path / "bad.txt"
        bad.write_text(txt.replace("state[5].real", "state[6].real", 1))
        with pytest.raises(E.QsvError, match="amplitude 5"):
            a.load(str(bad), text=True)
        bad.write_text(txt + "state[16].real = 0 state[16].imag = 0\n")
        with pytest.raises(E.QsvError, match="longer"):
            a.load(str(bad), text=True)
        bad.write_text("".join(txt.splitlines(True)[:10]))
        with pytest.raises(E.QsvError, match="ends"):
            a.load(str(bad), text=True)
        with pytest.raises(E.QsvError):
            a.save("", text=False)


def test_sdk_state_files(tmp_path):
    prog = os.path.join(os.path.dirname(__file__), "golden", "qrunes", "shor_15_sdk.txt")
    with sdk.SDK() as s:
        assert s.save_state(str(tmp_path / "x"), True) == sdk.RUN_PROGRAM_ERROR     # before a run
        assert s.load_file(prog) == sdk.Q_ERROR_NONE
        assert s.set_compute_unit(sdk.GPU) == sdk.Q_ERROR_NONE
        s.set_seed(99)
        assert s.run(10) == sdk.Q_ERROR_NONE
        base = str(tmp_path / "shor")
        assert s.save_state(base, False) == sdk.Q_ERROR_NONE
        assert s.save_state(base, True) == sdk.Q_ERROR_NONE
        assert s.save_state("", True) == sdk.Q_PARAMETER_ERROR
        with open(base + ".fullstate") as f:
            assert f.read() == s.state_text()              # the getQuantumState text
        assert np.array_equal(np.fromfile(base + ".fullstate.binary", np.complex128), s.amplitudes())
