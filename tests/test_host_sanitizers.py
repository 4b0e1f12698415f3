"""CPU checks of the host C++ (no GPU):

* the greedy and window planners (qsv_fusion.cpp, qsv_window.cpp) and the
  QRunes front end (QRunes.cpp) built with AddressSanitizer and
  UndefinedBehaviorSanitizer and driven by tests/cpp/host_sanitize_driver.cpp
  (random gate lists of every record kind, the reference's example programs,
  generated and byte-mutated programs); the device-side symbols the planners'
  translation units reference are stubs (tests/cpp/host_sanitize_stubs.cpp);
* the QuantumGates adapter (HipQuantumGates.cpp) compiled against the
  reference's own QuantumGates.h, where the reference tree is present: the
  adapter's redeclared interface is the reference's."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_QIH = "/root/reference/QPandaSDK/QuantumInstructionHandle"


def _gxx():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    return gxx


def test_planners_and_qrunes_under_asan_ubsan(tmp_path):
    gxx = _gxx()
    exe = str(tmp_path / "host_sanitize")
    srcs = [os.path.join(ROOT, "tests", "cpp", "host_sanitize_driver.cpp"),
            os.path.join(ROOT, "tests", "cpp", "host_sanitize_stubs.cpp"),
            os.path.join(ROOT, "qpanda_amd", "csrc", "qsv_window.cpp"),
            os.path.join(ROOT, "qpanda_amd", "csrc", "qsv_fusion.cpp"),
            os.path.join(ROOT, "qpanda_amd", "cpp", "QRunes.cpp")]
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "qpanda_amd", "cpp"), "-o", exe] + srcs
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    progs = sorted(os.path.join(ROOT, "tests", "golden", "qrunes", f)
                   for f in os.listdir(os.path.join(ROOT, "tests", "golden", "qrunes")) if f.endswith(".txt"))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, "200"] + progs, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "0 failures" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]


def test_adapter_compiles_against_the_reference_interface(tmp_path):
    gxx = _gxx()
    if not os.path.isfile(os.path.join(REF_QIH, "QuantumGates.h")):
        pytest.skip("reference tree not present")
    # the adapter's own QuantumGates.h must not shadow the reference's: copy
    # the adapter (own files) where the only QuantumGates.h on the path is the
    # reference's, keeping its relative include of include/qsv.h
    d = tmp_path / "a" / "cpp"
    d.mkdir(parents=True)
    for f in ("HipQuantumGates.h", "HipQuantumGates.cpp"):
        shutil.copy(os.path.join(ROOT, "qpanda_amd", "cpp", f), d / f)
    os.symlink(os.path.join(ROOT, "include"), tmp_path / "include")
    cmd = [gxx, "-std=c++14", "-fsyntax-only", "-w", "-fpermissive", "-I" + REF_QIH, str(d / "HipQuantumGates.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
