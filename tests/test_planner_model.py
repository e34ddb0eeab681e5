"""The fused-sweep planner, executed on a CPU model of the tile kernel, against the oracle.

tests/cpp/tile_model.cpp runs the engine's own lowering and planner and replays each tile
stage with the kernel's per-thread index arithmetic, so planner mistakes (a dropped gate, a
reordered dependency, a wrong slot) fail here without a GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import helpers as H
from quantr_amd import workloads as W

SO = os.path.join(H.ROOT, "tests", "_build", "libtile_model.so")


def model():
    srcs = [os.path.join(H.ROOT, "tests", "cpp", "tile_model.cpp"),
            os.path.join(H.ROOT, "quantr_amd", "csrc", "lower.cpp"),
            os.path.join(H.ROOT, "quantr_amd", "csrc", "planner.cpp")]
    hdrs = [os.path.join(H.ROOT, "quantr_amd", "csrc", f) for f in ("engine.hpp", "kernels.hpp")]
    newest = max(os.path.getmtime(f) for f in srcs + hdrs)
    if not os.path.exists(SO) or os.path.getmtime(SO) < newest:
        os.makedirs(os.path.dirname(SO), exist_ok=True)
        subprocess.check_call(["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-fPIC", "-shared",
                               "-I" + os.path.join(H.ROOT, "include"),
                               "-I" + os.path.join(H.ROOT, "quantr_amd", "csrc"), "-x", "hip",
                               "--offload-arch=gfx950"] + srcs + ["-o", SO],
                              stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    lib = C.CDLL(SO)
    lib.model_run.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                              C.POINTER(C.c_uint64)]
    lib.model_error.restype = C.c_char_p
    return lib


def run_model(n, program, tile_bits, low_bits=4, reg_bits=4):
    lib = model()
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    p = program.c_struct()
    sw = C.c_uint64()
    rc = lib.model_run(n, psi.ctypes.data, C.byref(p), tile_bits, low_bits, reg_bits, C.byref(sw))
    assert rc == 0, lib.model_error()
    return psi, sw.value


@pytest.mark.parametrize("tile_bits", [1, 2, 3, 4, 5, 6, 7, 9, 12])
def test_model_random(tile_bits):
    rng = np.random.default_rng(1000 + tile_bits)
    for trial in range(12):
        n = int(rng.integers(1, 12))
        gates, customs = H.random_gates(n, 50, rng, with_custom=(trial % 3 == 0))
        p = H.to_program(gates, customs)
        got, _ = run_model(n, p, tile_bits, low_bits=int(rng.integers(0, 5)))
        want = H.oracle_simulate(n, p)
        assert np.max(np.abs(got - want)) < 1e-12, (trial, n)


def test_model_brickwork_and_qft():
    n = 14
    p = W.to_program(W.brickwork(n, depth=12, seed=3))
    got, sweeps = run_model(n, p, 12)
    want = H.oracle_simulate(n, p)
    assert np.max(np.abs(got - want)) < 1e-12
    assert sweeps < len(p) / 5
    gates, _ = W.qft10(10, 2)
    p = W.to_program(gates)
    got, _ = run_model(10, p, 8)
    assert np.max(np.abs(got - H.oracle_simulate(10, p))) < 1e-12


def test_model_swap_heavy():
    rng = np.random.default_rng(44)
    for trial in range(10):
        n = int(rng.integers(3, 11))
        gates, _ = H.random_gates(n, 60, rng, kinds=["Swap", "H", "CNot", "T", "Rx"])
        p = H.to_program(gates)
        got, _ = run_model(n, p, min(n, 7))
        assert np.max(np.abs(got - H.oracle_simulate(n, p))) < 1e-12, trial


@pytest.mark.parametrize("reg_bits", [3, 5])
@pytest.mark.parametrize("tile_bits", [10, 11, 12, 13])
def test_model_register_bits(tile_bits, reg_bits):
    rng = np.random.default_rng(2000 + tile_bits)
    for trial in range(4):
        n = int(rng.integers(tile_bits, 15))
        gates, customs = H.random_gates(n, 60, rng, with_custom=(trial % 2 == 0))
        p = H.to_program(gates, customs)
        got, _ = run_model(n, p, tile_bits, low_bits=3, reg_bits=reg_bits)
        want = H.oracle_simulate(n, p)
        assert np.max(np.abs(got - want)) < 1e-12, (trial, n)
