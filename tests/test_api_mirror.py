"""The C++ mirror of quantr's API (include/quantr.hpp) run the way the reference's own tests
use the Rust API: tests/cpp/api_mirror.cpp restates circuit.rs's layout and error tests (CPU)
and tests/qft.rs, tests/grovers.rs, the custom-gate, register and post-selection cases (GPU)."""
import os
import subprocess
import sys

import pytest

import helpers as H

SRC = os.path.join(H.ROOT, "tests", "cpp", "api_mirror.cpp")
EXE = os.path.join(H.ROOT, "tests", "_build", "api_mirror")
LIBDIR = os.path.join(H.ROOT, "quantr_amd", "lib")


def build():
    deps = [SRC, os.path.join(H.ROOT, "include", "quantr.hpp"), os.path.join(H.ROOT, "include", "quantr_hip.h"),
            os.path.join(LIBDIR, "libquantr_hip.so")]
    if not os.path.exists(os.path.join(LIBDIR, "libquantr_hip.so")):
        pytest.fail("libquantr_hip.so is not built")
    if not os.path.exists(EXE) or os.path.getmtime(EXE) < max(os.path.getmtime(f) for f in deps):
        os.makedirs(os.path.dirname(EXE), exist_ok=True)
        subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(H.ROOT, "include"), SRC,
                               "-o", EXE, "-L" + LIBDIR, "-lquantr_hip", "-Wl,-rpath," + LIBDIR])
    return EXE


def run_mode(mode, timeout):
    exe = build()
    r = subprocess.run([exe, mode], capture_output=True, text=True, timeout=timeout)
    sys.stdout.write(r.stdout)
    sys.stdout.flush()
    fails = [l for l in r.stdout.splitlines() if l.startswith("FAIL")]
    assert r.returncode == 0 and not fails, (r.returncode, fails, r.stderr[-2000:])
    return r.stdout


def test_mirror_builder_and_errors():
    out = run_mode("cpu", 60)
    assert out.count("ok ") >= 16


def test_mirror_printer_diagrams():
    """Printer output equals the diagrams asserted by the reference's printer tests."""
    import json
    exe = build()
    out = subprocess.run([exe, "diagrams"], capture_output=True, timeout=60).stdout.decode("utf-8")
    fixture = json.load(open(os.path.join(H.ROOT, "tests", "golden", "printer_diagrams.json"), encoding="utf-8"))
    for case in fixture["cases"]:
        start = out.index("<<<" + case["name"] + "\n") + len(case["name"]) + 4
        got = out[start:out.index(">>>", start)]
        assert got == case["diagram"], (case["name"], got)


@pytest.mark.gpu
def test_mirror_simulations_gpu():
    out = run_mode("gpu", 240)
    for name in ("simple_qft", "grovers_3qubit", "x3sudoko", "custom_gates", "toffoli_gates", "custom_register",
                 "clone_and_simulate_keeps_circuit", "post_select", "measure_all_without_cache", "adjoint_function"):
        assert "ok " + name in out
