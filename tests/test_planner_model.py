"""The fused-sweep planner, executed on a CPU model of the tile kernel, against the oracle.

tests/cpp/tile_model.cpp runs the engine's own lowering and planner and replays each tile
stage with the kernel's per-thread index arithmetic, so planner mistakes (a dropped gate, a
reordered dependency, a wrong slot) fail here without a GPU."""
import ctypes as C
import math
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
    lib.model_adjoint.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                  C.c_uint64, C.POINTER(C.c_uint64)]
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


@pytest.mark.parametrize("rot", ["0", "1"])
def test_model_rotation_forms(rot, monkeypatch):
    """Rx and Ry as rotations with cos or sin divided out (4 fused multiply-adds per pair, one
    variant per sign of |cos| - |sin|; complex64's default) and as plain 2x2 forms (complex128's
    default), run through the launch coefficients, against the oracle. The angles take both
    variants, the crossover at pi/2 and the zeros of cos and sin."""
    monkeypatch.setenv("QR_ROT_FORMS", rot)
    n = 9
    angles = [0.0, 1e-12, 0.3, math.pi / 4, math.pi / 2, math.pi / 2 + 1e-9, 2.0, math.pi, 4.2,
              3 * math.pi / 2, 5.9, 2 * math.pi]
    gates = []
    for layer in range(6):
        for w in range(n):
            kind = ("Rx", "Ry", "H")[(w + layer) % 3]
            gates.append((kind, w, (), angles[(5 * w + layer) % len(angles)] if kind != "H" else 0.0, 0))
        for w in range(layer % 2, n - 1, 2):
            gates.append(("CNot", w + 1, (w,), 0.0, 0))
    p = H.to_program(gates)
    want = H.oracle_simulate(n, p)
    for tb in (3, 6, 9):
        got, _ = run_model(n, p, tb)
        assert np.max(np.abs(got - want)) < 1e-12, tb


def test_model_whole_state_stage_limit():
    """A tile covering the whole state stops at the per-stage op limit; the split keeps order."""
    rng = np.random.default_rng(45)
    n = 6
    gates, customs = H.random_gates(n, 400, rng, with_custom=True)
    p = H.to_program(gates, customs)
    got, sweeps = run_model(n, p, 12)
    assert np.max(np.abs(got - H.oracle_simulate(n, p))) < 1e-12
    assert sweeps >= 4


@pytest.mark.parametrize("cap", [1, 3, 9, 40])
def test_model_stage_caps(cap):
    """Capped stages (the planner's ops-per-stage limit, which stage balancing varies) defer
    their tail to later sweeps without reordering dependent ops."""
    lib = model()
    lib.model_run_capped.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                     C.POINTER(C.c_uint64)]
    rng = np.random.default_rng(300 + cap)
    for trial in range(6):
        n = int(rng.integers(7, 12))
        gates, customs = H.random_gates(n, 120, rng, with_custom=(trial % 2 == 0))
        p = H.to_program(gates, customs)
        psi = np.zeros(1 << n, dtype=np.complex128)
        psi[0] = 1.0
        sw = C.c_uint64()
        ps = p.c_struct()
        rc = lib.model_run_capped(n, psi.ctypes.data, C.byref(ps), 6, 3, cap, C.byref(sw))
        assert rc == 0, lib.model_error()
        assert np.max(np.abs(psi - H.oracle_simulate(n, p))) < 1e-12, (cap, trial)
        if cap == 1:
            assert sw.value >= len(gates) // 2


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


# ---- adjoint sweep (dual-state stages with gradient taps) ----------------------------------

def pauli_apply(n, psi, terms):
    """O psi for O = sum_t c_t P_t by direct index arithmetic (wire w is index bit n-1-w)."""
    idx = np.arange(1 << n)
    out = np.zeros_like(psi)
    for c, fac in terms:
        v = psi.copy()
        for w, p in fac:  # rightmost factor acts first; factors on distinct wires commute
            bit = 1 << (n - 1 - w)
            b = (idx & bit) != 0
            if p == "Z":
                v = np.where(b, -v, v)
            elif p == "X":
                v = v[idx ^ bit]
            else:  # (Y v)_j = -i (-1)^{b_j} v_{j^bit}
                v = -1j * np.where(b, -1.0, 1.0) * v[idx ^ bit]
        out += c * v
    return out


def unitary_customs(gates, customs, rng):
    for i in customs:
        k = 1 + len(gates[i][2])
        customs[i] = (W.random_unitary(k, int(rng.integers(1 << 30))).T.copy(), np.ones(1 << k, dtype=np.uint8))
    return customs


def shift_gradients(n, gates, customs, obs):
    """Parameter-shift rule through the oracle: exact for exp(-i theta G / 2), G^2 = I;
    the global phase gate has zero gradient."""
    from quantr_amd import _capi
    o = _capi.Observable(obs)
    out = []
    for i, (kind, pos, ctl, arg, argi) in enumerate(gates):
        if kind not in ("Rx", "Ry", "Rz", "Phase"):
            continue
        if kind == "Phase":
            out.append(0.0)
            continue
        vals = []
        for sgn in (+1, -1):
            gg = list(gates)
            gg[i] = (kind, pos, ctl, arg + sgn * np.pi / 2, argi)
            psi = H.oracle_simulate(n, H.to_program(gg, customs))
            vals.append(H.oracle_expectation(n, psi, o))
        out.append(0.5 * (vals[0] - vals[1]))
    return np.array(out)


def run_model_adjoint(n, gates, customs, obs, tile_bits, low_bits):
    lib = model()
    p = H.to_program(gates, customs)
    psi = H.oracle_simulate(n, p).astype(np.complex128)
    lam = pauli_apply(n, psi, obs)
    npar = sum(1 for g in gates if g[0] in ("Rx", "Ry", "Rz", "Phase"))
    grads = np.zeros(npar)
    ps = p.c_struct()
    sw = C.c_uint64()
    rc = lib.model_adjoint(n, psi.ctypes.data, lam.ctypes.data, C.byref(ps), tile_bits, low_bits,
                           grads.ctypes.data, npar, C.byref(sw))
    assert rc == 0, lib.model_error()
    return grads, psi, sw.value


def random_observable(n, rng, terms=5):
    obs = []
    for _ in range(terms):
        m = int(rng.integers(1, min(n, 3) + 1))
        wires = rng.permutation(n)[:m]
        obs.append((float(rng.uniform(-1, 1)), [(int(w), "XYZ"[int(rng.integers(3))]) for w in wires]))
    return obs


@pytest.mark.parametrize("tile_bits", [1, 2, 3, 5, 7, 10])
def test_model_adjoint_random(tile_bits):
    """Dual-state fused reverse sweep: gradients equal the parameter-shift rule, and psi is
    returned to |0...0>."""
    rng = np.random.default_rng(3000 + tile_bits)
    for trial in range(6):
        n = int(rng.integers(2, 9))
        gates, customs = H.random_gates(n, 30, rng, with_custom=(trial % 2 == 0))
        customs = unitary_customs(gates, customs, rng)
        obs = random_observable(n, rng)
        got, psi0, _ = run_model_adjoint(n, gates, customs, obs, tile_bits, int(rng.integers(0, 4)))
        want = shift_gradients(n, gates, customs, obs)
        assert np.max(np.abs(got - want)) < 1e-10, (trial, n)
        e0 = np.zeros(1 << n, dtype=np.complex128)
        e0[0] = 1.0
        assert np.max(np.abs(psi0 - e0)) < 1e-10


def test_model_adjoint_long_whole_state():
    rng = np.random.default_rng(46)
    n = 5
    gates, customs = H.random_gates(n, 200, rng, with_custom=True)
    customs = unitary_customs(gates, customs, rng)
    obs = random_observable(n, rng)
    got, _, sweeps = run_model_adjoint(n, gates, customs, obs, 10, 2)
    assert np.max(np.abs(got - shift_gradients(n, gates, customs, obs))) < 1e-10
    assert sweeps >= 3


def test_model_adjoint_hea():
    n, L = 9, 3
    gates, _ = W.hea(n, L, seed=11)
    obs = W.hea_observable(n)
    got, _, sweeps = run_model_adjoint(n, gates, {}, obs, 6, 3)
    want = shift_gradients(n, gates, {}, obs)
    assert np.max(np.abs(got - want)) < 1e-10
    assert sweeps < len(gates) / 3


@pytest.mark.parametrize("diag", [(0.0, np.sqrt(2.0)), (np.sqrt(2.0), 0.0), (0.0, 0.0), (1e-300, 1.0),
                                  (0.5j, 3.0), (2.0, 0.5)])
@pytest.mark.parametrize("controls", [(), (0,)])
def test_model_custom_diagonal_zero_entries(diag, controls):
    """Custom diagonal gates with zero or tiny entries (post-selection, the zero gate) inside a
    fused tile stage: no entry is ever divided by a zero or tiny one."""
    n = 3
    images = np.zeros((1 << (1 + len(controls)), 1 << (1 + len(controls))), dtype=np.complex128)
    dim = images.shape[0]
    for j in range(dim):
        images[j, j] = 1.0
    images[dim - 2, dim - 2] = diag[0]
    images[dim - 1, dim - 1] = diag[1]
    p = H.Program()
    for w in range(n):
        p.add("H", w)
    p.add_custom(1, list(controls), images, np.ones(dim, dtype=np.uint8))
    p.add("Rx", 2, (), 0.3)
    p.add("T", 1)
    want = H.oracle_simulate(n, p)
    for tile_bits in (2, 3):
        got, _ = run_model(n, p, tile_bits)
        assert np.all(np.isfinite(got)), (diag, controls, tile_bits)
        assert np.max(np.abs(got - want)) < 1e-12, (diag, controls, tile_bits)


def run_model_opts(n, program, tile_bits, low_bits, reg_bits=4, max_ops=0, lds_split=0, dyn_low=0, seed=0,
                   tie_slack=0):
    lib = model()
    lib.model_run_opts.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_uint32, C.c_int, C.POINTER(C.c_uint64)]
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    p = program.c_struct()
    sw = C.c_uint64()
    rc = lib.model_run_opts(n, psi.ctypes.data, C.byref(p), tile_bits, low_bits, reg_bits, max_ops, lds_split,
                            dyn_low, seed, tie_slack, C.byref(sw))
    assert rc == 0, lib.model_error()
    return psi, sw.value


def run_model_split(n, program, tile_bits, low_bits, reg_bits=4, max_ops=0):
    return run_model_opts(n, program, tile_bits, low_bits, reg_bits, max_ops, lds_split=1)


@pytest.mark.parametrize("tile_bits,low_bits,reg_bits", [(13, 4, 4), (13, 3, 4), (12, 3, 4), (10, 3, 3), (11, 4, 5)])
def test_model_split_remaps(tile_bits, low_bits, reg_bits):
    """Remaps through half the LDS in two rounds (the split position's bit is the top LDS
    address bit; first-round reads still to be overwritten go through temporaries): every
    round writes and reads each LDS slot once, and the states match the oracle."""
    rng = np.random.default_rng(3300 + tile_bits * 10 + low_bits + reg_bits)
    for trial in range(6):
        n = int(rng.integers(tile_bits, tile_bits + 3))
        gates, customs = H.random_gates(n, 90, rng, with_custom=(trial % 3 == 0))
        p = H.to_program(gates, customs)
        got, _ = run_model_split(n, p, tile_bits, low_bits, reg_bits)
        want = H.oracle_simulate(n, p)
        assert np.max(np.abs(got - want)) < 1e-12, (trial, n)
    n = tile_bits + 2
    p = W.to_program(W.brickwork(n, depth=8, seed=5))
    got, _ = run_model_split(n, p, tile_bits, low_bits, reg_bits)
    assert np.max(np.abs(got - H.oracle_simulate(n, p))) < 1e-12


@pytest.mark.parametrize("tile_bits,low_bits,lds_split", [(12, 4, 1), (12, 3, 0), (10, 4, 0), (6, 2, 0), (13, 4, 1)])
def test_model_moving_layout(tile_bits, low_bits, lds_split):
    """The moving layout: stages store bits the next stage uses at the physical low bits, the
    layout is restored before standalone ops and at the end; states match the oracle."""
    rng = np.random.default_rng(4400 + tile_bits * 10 + low_bits)
    for trial in range(8):
        n = int(rng.integers(tile_bits + 1, tile_bits + 5))
        gates, customs = H.random_gates(n, 120, rng, with_custom=(trial % 2 == 0))
        p = H.to_program(gates, customs)
        got, _ = run_model_opts(n, p, tile_bits, low_bits, lds_split=lds_split, dyn_low=1)
        want = H.oracle_simulate(n, p)
        assert np.max(np.abs(got - want)) < 1e-12, (trial, n)
    n = tile_bits + 4
    p = W.to_program(W.brickwork(n, depth=10, seed=9))
    got, sweeps = run_model_opts(n, p, tile_bits, low_bits, lds_split=lds_split, dyn_low=1)
    assert np.max(np.abs(got - H.oracle_simulate(n, p))) < 1e-12
    _, fixed = run_model_opts(n, p, tile_bits, low_bits, lds_split=lds_split, dyn_low=0)
    assert sweeps <= fixed + 1, (sweeps, fixed)


@pytest.mark.parametrize("seed,tie_slack", [(1, 0), (5, 0), (7, 0), (3, 1), (29, 1)])
def test_model_seeded_ties(seed, tie_slack):
    """Seeded tie rules in the greedy bit choice, near-ties included (tie_slack 1), give other
    plans of the same program (the balanced planner tries many); each must still compute the
    circuit."""
    rng = np.random.default_rng(4500 + seed)
    for trial in range(6):
        tile_bits, low_bits = [(12, 3), (10, 4), (13, 4)][trial % 3]
        n = int(rng.integers(tile_bits + 1, tile_bits + 4))
        gates, customs = H.random_gates(n, 120, rng, with_custom=(trial % 2 == 0))
        p = H.to_program(gates, customs)
        want = H.oracle_simulate(n, p)
        for max_ops in (0, 40):
            got, _ = run_model_opts(n, p, tile_bits, low_bits, lds_split=1, dyn_low=1, seed=seed, max_ops=max_ops,
                                    tie_slack=tie_slack)
            assert np.max(np.abs(got - want)) < 1e-12, (trial, n, max_ops)
    n = 16
    p = W.to_program(W.brickwork(n, depth=12, seed=seed))
    got, _ = run_model_opts(n, p, 12, 3, lds_split=1, dyn_low=1, seed=seed, tie_slack=tie_slack)
    assert np.max(np.abs(got - H.oracle_simulate(n, p))) < 1e-12
