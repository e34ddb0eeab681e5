"""GPU parity: the HIP engine through the C ABI against the oracle and the golden vectors.

Bars: the unfused reference-order path (fused=0) is bitwise equal to the oracle; the fused
path is within 1e-12 max amplitude error (complex128); complex64 within 2e-5; bins identical
for the same injected draws.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import helpers as H
from quantr_amd import _capi
from quantr_amd import workloads as W

pytestmark = pytest.mark.gpu

TOL_FUSED = 1e-12
TOL_C64 = 2e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if _capi.device_count() < 1:
        pytest.fail("no GPU visible: the HIP engine cannot run")


def run_gpu(n, program, fused=1, precision=_capi.QR_C128, register=None, tile_bits=0, jit=0):
    st = _capi.State(n, precision)
    st.set_options(fused=fused, tile_bits=tile_bits, jit=jit)
    if register is not None:
        st.set_amplitudes(register)
    st.apply(program)
    out = st.amplitudes()
    st.close()
    return out


def nested_gpu(k, program, reg):
    return run_gpu(k, program, fused=0, register=reg)


@pytest.mark.parametrize("fused", [0, 1])
def test_golden_vectors_gpu(fused):
    d = H.golden()
    for case in d["cases"]:
        if "expected" not in case:
            continue
        say("golden", case["name"], "fused", fused)
        p = H.program_from_fixture(case["program"], d["tables"], nested_gpu)
        reg = H.cvec(case["register"]) if "register" in case else None
        got = run_gpu(case["n"], p, fused=fused, register=reg)
        want = H.cvec(case["expected"])
        assert np.max(np.abs(got - want)) < 1e-12, case["name"]
        orc = H.oracle_simulate(case["n"], p, register=reg)
        if fused == 0:
            assert np.array_equal(got, orc), case["name"]
        else:
            assert np.max(np.abs(got - orc)) < TOL_FUSED, case["name"]


def test_random_circuits_exact_bitwise():
    rng = np.random.default_rng(101)
    for trial in range(30):
        n = int(rng.integers(1, 13))
        gates, customs = H.random_gates(n, 40, rng, with_custom=True)
        say("exact trial", trial, "n", n)
        p = H.to_program(gates, customs)
        got = run_gpu(n, p, fused=0)
        want = H.oracle_simulate(n, p)
        assert np.array_equal(got, want), trial


def test_random_circuits_fused():
    rng = np.random.default_rng(202)
    for trial in range(40):
        n = int(rng.integers(1, 16))
        gates, customs = H.random_gates(n, 60, rng, with_custom=(trial % 2 == 0))
        say("fused trial", trial, "n", n)
        p = H.to_program(gates, customs)
        got = run_gpu(n, p, fused=1)
        want = H.oracle_simulate(n, p)
        assert np.max(np.abs(got - want)) < TOL_FUSED, trial


@pytest.mark.parametrize("tile_bits", [2, 3, 4, 5, 6, 8, 10, 11, 12])
def test_tile_widths(tile_bits):
    rng = np.random.default_rng(300 + tile_bits)
    n = 14
    gates, _ = H.random_gates(n, 120, rng)
    say("tile width", tile_bits)
    p = H.to_program(gates)
    got = run_gpu(n, p, fused=1, tile_bits=tile_bits)
    want = H.oracle_simulate(n, p)
    assert np.max(np.abs(got - want)) < TOL_FUSED


@pytest.mark.parametrize("low_bits", [3, 4])
def test_split_lds_tiles(low_bits):
    """Tiles of 2^13 complex128 amplitudes (128 KB) exchange through half the LDS in two rounds
    in compiled stages; the interpreter runs the same plans through the whole tile in LDS."""
    rng = np.random.default_rng(1300 + low_bits)
    for trial in range(6):
        n = int(rng.integers(13, 18))
        gates, customs = H.random_gates(n, 150, rng, with_custom=(trial % 3 == 0))
        p = H.to_program(gates, customs)
        want = H.oracle_simulate(n, p)
        for jit in (1, 0):
            say("split tiles trial", trial, "n", n, "low bits", low_bits, "jit", jit)
            st = _capi.State(n)
            st.set_options(fused=1, tile_bits=13, low_bits=low_bits, jit=jit)
            st.apply(p)
            got = st.amplitudes()
            st.close()
            assert np.max(np.abs(got - want)) < TOL_FUSED, (trial, jit)
    n = 18
    p = W.to_program(W.brickwork(n, depth=12, seed=18))
    want = H.oracle_simulate(n, p, threads=8)
    st = _capi.State(n)
    st.set_options(fused=1, tile_bits=13, low_bits=low_bits, jit=1)
    st.apply(p)
    assert np.max(np.abs(st.amplitudes() - want)) < TOL_FUSED
    st.close()


@pytest.mark.parametrize("tile_bits,low_bits", [(12, 4), (13, 4), (12, 3)])
def test_moving_layout(tile_bits, low_bits, monkeypatch):
    """Plans whose stages move bits into the physical low bits (restored to the identity before
    standalone ops and at the end), compiled and interpreted, against the oracle."""
    monkeypatch.setenv("QR_DYN_LOW", "1")  # the default; set explicitly for clarity
    rng = np.random.default_rng(1500 + tile_bits * 10 + low_bits)
    for trial in range(6):
        n = int(rng.integers(tile_bits + 1, tile_bits + 5))
        gates, customs = H.random_gates(n, 150, rng, with_custom=(trial % 2 == 0))
        p = H.to_program(gates, customs)
        want = H.oracle_simulate(n, p)
        for prec, tol in ((_capi.QR_C128, TOL_FUSED), (_capi.QR_C64, TOL_C64)):
            for jit in (1, 0):
                say("moving layout trial", trial, "n", n, "prec", prec, "jit", jit)
                st = _capi.State(n, prec)
                st.set_options(fused=1, tile_bits=tile_bits, low_bits=low_bits, jit=jit)
                st.apply(p)
                got = st.amplitudes()
                st.close()
                assert np.max(np.abs(got - want)) < tol, (trial, prec, jit)


def test_identity_layout_knob(monkeypatch):
    """QR_DYN_LOW=0 plans with the identity layout; same results."""
    monkeypatch.setenv("QR_DYN_LOW", "0")
    rng = np.random.default_rng(1600)
    n = 15
    gates, customs = H.random_gates(n, 120, rng, with_custom=True)
    p = H.to_program(gates, customs)
    got = run_gpu(n, p, fused=1, jit=1)
    assert np.max(np.abs(got - H.oracle_simulate(n, p))) < TOL_FUSED


@pytest.mark.parametrize("diag", [(0.0, np.sqrt(2.0)), (np.sqrt(2.0), 0.0), (0.0, 0.0), (1e-300, 1.0)])
@pytest.mark.parametrize("controls", [(), (0,)])
def test_custom_diagonal_zero_entries_gpu(diag, controls):
    """Custom diagonal gates with zero or tiny entries (post-selection onto |1>, the zero gate)
    in fused tile stages, interpreted and compiled: finite and equal to the oracle."""
    n = 3
    dim = 1 << (1 + len(controls))
    images = np.eye(dim, dtype=np.complex128)
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
        for jit in (0, 1):
            got = run_gpu(n, p, fused=1, tile_bits=tile_bits, jit=jit)
            assert np.all(np.isfinite(got)), (diag, controls, tile_bits, jit)
            assert np.max(np.abs(got - want)) < TOL_FUSED, (diag, controls, tile_bits, jit)


@pytest.mark.parametrize("prec,tol", [(_capi.QR_C128, TOL_FUSED), (_capi.QR_C64, TOL_C64)])
def test_rotation_angles_gpu(prec, tol):
    """Rx and Ry run as three shears after the host folds out their magnitude and sign: the
    angles where that fold changes sign or degenerates (0, pi/2, pi, 3pi/2, 2pi and their
    neighbours), in compiled and interpreted stages, against the oracle."""
    import math
    angles = [0.0, 1e-12, 0.3, math.pi / 4, math.pi / 2, math.pi / 2 + 1e-9, 2.0, math.pi,
              math.pi + 1e-9, 4.2, 3 * math.pi / 2, 5.9, 2 * math.pi - 1e-9, 2 * math.pi]
    for n in (9, 14):
        gates = []
        for layer in range(6):
            for w in range(n):
                kind = ("Rx", "Ry", "H")[(w + layer) % 3]
                gates.append((kind, w, (), angles[(5 * w + layer) % len(angles)] if kind != "H" else 0.0, 0))
            for w in range(layer % 2, n - 1, 2):
                gates.append(("CNot", w + 1, (w,), 0.0, 0))
        p = H.to_program(gates)
        want = H.oracle_simulate(n, p)
        for jit in (1, 0):
            say("rotation angles n", n, "prec", prec, "jit", jit)
            got = run_gpu(n, p, fused=1, precision=prec, jit=jit)
            assert np.max(np.abs(got - want)) < tol, (n, jit)


def test_complex64_random():
    rng = np.random.default_rng(404)
    for trial in range(10):
        n = int(rng.integers(2, 15))
        gates, _ = H.random_gates(n, 50, rng)
        say("c64 trial", trial, "n", n)
        p = H.to_program(gates)
        got = run_gpu(n, p, fused=1, precision=_capi.QR_C64)
        want = H.oracle_simulate(n, p)
        assert np.max(np.abs(got - want)) < TOL_C64, trial


def test_qft10_state_and_bins():
    """QFT on 10 qubits: full state and measure_all(500) bins against the oracle, with the
    same 500 injected draws."""
    for seed in range(5):
        gates, _ = W.qft10(10, seed)
        p = W.to_program(gates)
        u = np.array(W.uniforms(500, seed + 1000))
        want = H.oracle_simulate(10, p)
        want_bins = H.oracle_measure(10, want, u)
        for fused in (0, 1):
            say("qft10 seed", seed, "fused", fused)
            st = _capi.State(10)
            st.set_options(fused=fused)
            st.apply(p)
            got = st.amplitudes()
            if fused == 0:
                assert np.array_equal(got, want)
            assert np.max(np.abs(got - want)) < TOL_FUSED
            bins = st.sample(u)
            st.close()
            # the sampler reproduces the left fold exactly on the state it holds ...
            assert np.array_equal(bins, H.oracle_measure(10, got, u))
            # ... and the fused state (within 1e-12 of the reference's) gives the reference's
            # bins for these draws: none lies within 1e-12 of a cumulative-probability boundary
            assert np.array_equal(bins, want_bins)


def test_sampler_exact_fold_bitwise():
    rng = np.random.default_rng(55)
    for n in (1, 4, 9, 13, 16):
        psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
        psi /= np.linalg.norm(psi)
        psi[rng.integers(1 << n)] = 0.0
        u = rng.random(1000)
        u[:4] = [0.0, 0.5, 1.0, 2.0]
        say("sampler n", n)
        st = _capi.State(n)
        st.set_amplitudes(psi)
        st.set_options(exact_sample=1)
        got = st.sample(u)
        want = H.oracle_measure(n, st.amplitudes(), u)
        assert np.array_equal(got, want)
        st.set_options(exact_sample=0)
        got2 = st.sample(u)
        assert np.array_equal(got2, want)
        st.close()


def test_sampler_none_for_non_unitary():
    images = np.array([[np.sqrt(2.0), 0.0], [0.0, 0.0]], dtype=np.complex128)
    p = _capi.Program()
    for w in range(3):
        p.add("H", w)
    p.add_custom(1, [], images, [1, 1])
    st = _capi.State(3)
    st.apply(p)
    out = st.sample([0.1, 0.99999, 0.999999999])
    assert list(out[:2]) == [0, 5] and out[2] == 5
    st.set_amplitudes(np.full(8, 0.25 + 0j))  # total probability 0.5
    out = st.sample([0.2, 0.6])
    assert out[0] == 3 and out[1] == np.iinfo(np.uint64).max
    st.close()


def test_brickwork_20q_vs_oracle():
    n = 20
    gates = W.brickwork(n, depth=40, seed=1234)
    p = W.to_program(gates)
    say("brickwork 20q oracle")
    want = H.oracle_simulate(n, p, threads=8)
    say("brickwork 20q gpu")
    st = _capi.State(n)
    st.apply(p)
    got = st.amplitudes()
    plan = st.last_plan()
    st.close()
    assert np.max(np.abs(got - want)) < TOL_FUSED
    assert plan["sweeps"] < 120


def test_balanced_plan_25q_vs_exact():
    """From 2^25 amplitudes, forward plans are stage-balanced (several ops-per-stage caps and
    tie rules are planned and the lowest estimated duration kept); a caller-set cap is
    honoured as given. Both match the reference-order mode."""
    n = 25
    gates = W.brickwork(n, depth=40, seed=77)
    p = W.to_program(gates)
    say("balanced 25q exact")
    want = run_gpu(n, p, fused=0)
    for jit in (0, 1):
        say("balanced 25q fused jit=%d" % jit)
        st = _capi.State(n)
        st.set_options(fused=1, jit=jit)
        st.apply(p)
        got = st.amplitudes()
        plan = st.last_plan()
        st.close()
        assert np.max(np.abs(got - want)) < TOL_FUSED, jit
        say("balanced 25q: largest stage %d gates, %d sweeps" % (plan["max_stage_gates"], plan["sweeps"]))
        assert 0 < plan["max_stage_gates"] and plan["sweeps"] <= len(gates) // 8
    say("balanced 25q stage_ops=7")
    st = _capi.State(n)
    st.set_options(fused=1, jit=0, stage_ops=7)
    st.apply(p)
    got = st.amplitudes()
    plan = st.last_plan()
    st.close()
    assert np.max(np.abs(got - want)) < TOL_FUSED
    assert plan["sweeps"] >= len(gates) // 7


def test_measured_plans_25q():
    """tune = 1: the first applications of a structure measure candidate plans and tile
    orders. Every one of them computes the circuit (within the fused tolerance of the
    reference-order result); once measured, applications repeat bit for bit."""
    n = 25
    gates = W.brickwork(n, depth=40, seed=78)
    p = W.to_program(gates)
    say("measured plans 25q exact")
    want = run_gpu(n, p, fused=0)
    st = _capi.State(n)
    st.set_options(fused=1, tune=1)
    measured = 0
    while True:
        st.reset()
        st.apply(p)
        plan = st.last_plan()
        got = st.amplitudes()
        err = float(np.max(np.abs(got - want)))
        say("measured plans 25q: tuning=%d sweeps=%d xcd=%d err=%.2e" % (plan["tuning"], plan["sweeps"],
                                                                         plan["xcd_order_sweeps"], err))
        assert err < TOL_FUSED
        assert plan["xcd_order_sweeps"] <= plan["tile_sweeps"]
        if not plan["tuning"]:
            break
        measured += 1
        assert measured <= 4  # at most 3 candidate plans, then one run with the other tile order
    assert measured >= 2
    first = got
    st.reset()
    st.apply(p)
    again = st.amplitudes()
    plan2 = st.last_plan()
    st.close()
    assert plan2["tuning"] == 0 and plan2["xcd_order_sweeps"] == plan["xcd_order_sweeps"]
    assert np.array_equal(first, again)


@pytest.mark.parametrize("precision", [_capi.QR_C128, _capi.QR_C64])
def test_sample_programs_batched(precision):
    """measure_all_without_cache's batched re-simulation (qr_sample_programs): every program's
    draws equal what qr_sample gives on that program's own state, bit for bit. Two structures
    interleaved (one with non-unitary custom tables, whose "None" rows leave states with
    norm below one), programs with no draws, and sizes from one tile to many."""
    rng = np.random.default_rng(880 + precision)
    for n in (6, 13, 17):
        base_a, cust_a = H.random_gates(n, 40, rng, with_custom=True)
        base_b, _ = H.random_gates(n, 25, rng)
        progs, draws, us = [], [], []
        for j in range(24):
            if j % 3 == 2:  # structure B: new angles
                gates = [(k, pos, ctl, float(rng.uniform(-3, 3)), ai) for (k, pos, ctl, _, ai) in base_b]
                p = H.to_program(gates)
            else:  # structure A: new custom tables of the same shape, some with None rows
                cust = {}
                for i, (imgs, some) in cust_a.items():
                    im = rng.standard_normal(imgs.shape) + 1j * rng.standard_normal(imgs.shape)
                    im /= np.linalg.norm(im, axis=1, keepdims=True)
                    cust[i] = (im, some)
                p = H.to_program(base_a, cust)
            k = int(rng.integers(0, 6))
            progs.append(p)
            draws.append(k)
            us.extend(rng.random(k).tolist())
        say("sample programs n=%d prec=%d: %d programs, %d draws" % (n, precision, len(progs), len(us)))
        got = _capi.sample_programs(n, progs, draws, us, precision)
        at = 0
        for p, k in zip(progs, draws):
            st = _capi.State(n, precision)
            st.set_options(fused=1, jit=1)  # the same compiled stages as the batched run
            st.apply(p)
            want = st.sample(np.array(us[at:at + k])) if k else np.zeros(0, dtype=np.uint64)
            st.close()
            assert np.array_equal(got[at:at + k], want), (n, at)
            at += k


_CACHE_SCRIPT = r"""
import hashlib, sys
sys.path.insert(0, sys.argv[1])
from quantr_amd import _capi, workloads as W
st = _capi.State(20)
st.set_options(fused=1, jit=1)
st.apply(W.to_program(W.brickwork(20, depth=12, seed=2020)))
print("digest", hashlib.sha256(st.amplitudes().tobytes()).hexdigest(), flush=True)
st.close()
big = _capi.State(25)  # large enough for the plan search, whose choice is cached too
big.set_options(fused=1, jit=1)
big.apply(W.to_program(W.brickwork(25, depth=8, seed=2525)))
print("digest25", hashlib.sha256(big.amplitudes().tobytes()).hexdigest(), flush=True)
"""


def test_stage_cache_on_disk(tmp_path):
    """Compiled stages and plan choices are kept on disk: a second process with the same cache
    directory loads every stage of the same circuits from it, compiles none, takes the plan
    search's choice without searching, and gets the same amplitudes bit for bit. A damaged code
    object (cut to half its size) fails its content hash and is never loaded: a third process
    compiles its kernels again, rewrites it, and gets the same amplitudes."""
    import glob
    import os
    import subprocess
    env = dict(os.environ, QR_JIT_CACHE=str(tmp_path / "jit"), QR_JIT_CACHE_LOG="1")
    runs = []
    for k in range(3):
        if k == 2:
            objs = sorted(glob.glob(str(tmp_path / "jit" / "*" / "g*.co")) + glob.glob(str(tmp_path / "jit" / "g*.co")))
            assert objs
            with open(objs[0], "r+b") as f:
                f.truncate(os.path.getsize(objs[0]) // 2)
        say("stage cache: process %d" % k)
        r = subprocess.run([sys.executable, "-c", _CACHE_SCRIPT, H.ROOT], env=env, capture_output=True, text=True,
                           timeout=240)
        assert r.returncode == 0, r.stderr[-2000:]
        logs = [l for l in r.stderr.splitlines() if l.startswith("[qr jit]")]
        plans = [l for l in r.stderr.splitlines() if l.startswith("[qr plan]")]
        say("stage cache: %s %s" % (logs, plans))
        digest = [l for l in r.stdout.splitlines() if l.startswith("digest")]
        runs.append((logs, digest, plans))
    first, second, third = runs
    assert first[0] and " 0 kernels from the disk cache" in first[0][0]
    assert second[0] and all(l.endswith(", 0 to compile") for l in second[0])
    assert "[qr plan] searching plans" in first[2] and "[qr plan] plan choice from the disk cache" in second[2]
    assert first[1] == second[1] and len(first[1]) == 2
    assert any(not l.endswith(", 0 to compile") for l in third[0])
    assert third[1] == first[1]


def test_norm_and_expectation():
    rng = np.random.default_rng(77)
    n = 12
    gates, _ = H.random_gates(n, 80, rng)
    p = H.to_program(gates)
    st = _capi.State(n)
    st.apply(p)
    psi = st.amplitudes()
    assert abs(st.norm_sqr() - 1.0) < 1e-12
    obs = _capi.Observable(W.hea_observable(n) + [(0.3, [(2, "Y"), (5, "Y")]), (-0.2, [(1, "X"), (7, "Y"), (9, "Z")])])
    got = st.expectation(obs)
    want = H.oracle_expectation(n, psi, obs)
    assert abs(got - want) < 1e-11
    st.close()


def _pass_observable(n, rng):
    """Terms that exercise every kind of expectation pass: Z strings (X mask 0), single and
    paired X/Y factors on low and high wires (grouped into tile passes), X masks spread over
    more wires than a tile holds (the direct pass), and repeated X masks with different Z
    parts (one group, several terms)."""
    terms = list(W.hea_observable(n))
    for _ in range(12):
        w = sorted(rng.choice(n, size=2, replace=False).tolist())
        terms.append((float(rng.normal()), [(w[0], str(rng.choice(["X", "Y"]))), (w[1], str(rng.choice(["X", "Y", "Z"])))]))
    wide = sorted(rng.choice(n, size=min(n, 13), replace=False).tolist())
    terms.append((0.37, [(w, "X") for w in wide]))
    terms.append((-0.21, [(w, "Y" if i % 3 == 0 else "X") for i, w in enumerate(wide[:11])]))
    for w in (0, n - 1, n // 2):
        terms.append((0.11, [(w, "X"), ((w + 1) % n, "Z")]))
        terms.append((-0.07, [(w, "X"), ((w + 2) % n, "Z"), ((w + 3) % n, "Z")]))
    return _capi.Observable(terms)


@pytest.mark.parametrize("precision,n,tol", [(_capi.QR_C128, 17, 1e-11), (_capi.QR_C64, 18, 2e-5)])
def test_expectation_passes(precision, n, tol):
    """Expectation values of states large enough for tiled passes (X-mask groups evaluated
    together from one read of the state, wide groups read directly) against the oracle, and
    the allocation-free entry point bitwise equal to qr_expectation."""
    rng = np.random.default_rng(700 + n)
    st = _capi.State(n, precision)
    st.apply(W.to_program(W.brickwork(n, depth=8, seed=n)))
    psi = st.amplitudes()  # the amplitudes the state holds (rounded for complex64)
    obs = _pass_observable(n, rng)
    got = st.expectation(obs)
    want = H.oracle_expectation(n, psi, obs)
    say("expectation passes n=%d prec=%d: got %.15f want %.15f" % (n, precision, got, want))
    scale = sum(abs(c) for c, _ in obs.terms)
    assert abs(got - want) < tol * max(1.0, scale)
    v = C.c_double()
    _capi.check(_capi.lib().qr_expectation_noalloc(st.handle, C.byref(obs.c_struct()), C.byref(v)))
    st.close()
    assert v.value == got


def _energy_oracle(n, gates, obs):
    psi = H.oracle_simulate(n, W.to_program(gates))
    return H.oracle_expectation(n, psi, obs)


@pytest.mark.parametrize("precision,tol", [(_capi.QR_C128, 1e-9), (_capi.QR_C64, 1e-4)])
def test_adjoint_vs_parameter_shift(precision, tol):
    n, layers = 6, 3
    gates, thetas = W.hea(n, layers, seed=5)
    gates = gates + [("Phase", 2, (), 0.37, 0), ("Rx", 4, (), 1.1, 0), ("H", 1, (), 0.0, 0),
                     ("Swap", 3, (0,), 0.0, 0), ("Rx", 0, (), -0.4, 0)]
    obs = _capi.Observable(W.hea_observable(n))
    g = _capi.Grad(n, precision)
    e = g.forward(W.to_program(gates), obs)
    grads = g.backward(1.0)
    e_ref = _energy_oracle(n, gates, obs)
    assert abs(e - e_ref) < (1e-12 if precision == _capi.QR_C128 else 1e-5)
    pidx = [i for i, x in enumerate(gates) if x[0] in ("Rx", "Ry", "Rz", "Phase")]
    assert len(grads) == len(pidx)
    for k, i in enumerate(pidx):
        kind, pos, ctl, arg, argi = gates[i]
        if kind == "Phase":
            ps = 0.0
        else:
            gp = list(gates); gp[i] = (kind, pos, ctl, arg + np.pi / 2, argi)
            gm = list(gates); gm[i] = (kind, pos, ctl, arg - np.pi / 2, argi)
            ps = 0.5 * (_energy_oracle(n, gp, obs) - _energy_oracle(n, gm, obs))
        h = 1e-6
        gp = list(gates); gp[i] = (kind, pos, ctl, arg + h, argi)
        gm = list(gates); gm[i] = (kind, pos, ctl, arg - h, argi)
        fd = (_energy_oracle(n, gp, obs) - _energy_oracle(n, gm, obs)) / (2 * h)
        assert abs(grads[k] - ps) < tol, (k, kind)
        assert abs(ps - fd) < 1e-7
    g.close()


def test_oom_is_reported():
    h = _capi.State.__new__(_capi.State)
    import ctypes as C
    hh = C.c_void_p()
    st = _capi.lib().qr_state_create(37, _capi.QR_C128, C.byref(hh))
    assert st == _capi.QR_ERR_OOM
    assert "allocate" in _capi.lib().qr_last_error().decode()


def test_register_dimension_error():
    st = _capi.State(3)
    with pytest.raises(_capi.QrError) as e:
        st.set_amplitudes(np.zeros(4, dtype=np.complex128))
    assert str(e.value) == ("The custom register has a product state dimension of 2, while the number of "
                            "qubits is 3. These must equal each other.")
    st.close()


def test_compiled_stages_random():
    """Tile stages compiled at run time (hipRTC) against the oracle."""
    rng = np.random.default_rng(606)
    for trial in range(8):
        n = int(rng.integers(6, 17))
        gates, customs = H.random_gates(n, 80, rng, with_custom=(trial % 2 == 0))
        say("jit trial", trial, "n", n)
        p = H.to_program(gates, customs)
        want = H.oracle_simulate(n, p)
        for prec, tol in ((_capi.QR_C128, TOL_FUSED), (_capi.QR_C64, TOL_C64)):
            got = run_gpu(n, p, fused=1, precision=prec, jit=1)
            assert np.max(np.abs(got - want)) < tol, (trial, prec)


def test_compiled_stages_brickwork_repeat():
    """Second application of the same structure switches to compiled stages (jit = 2)."""
    n = 18
    p = W.to_program(W.brickwork(n, depth=20, seed=77))
    want = H.oracle_simulate(n, p, threads=8)
    st = _capi.State(n)
    for rep in range(3):
        st.reset()
        st.apply(p)
        got = st.amplitudes()
        assert np.max(np.abs(got - want)) < TOL_FUSED, rep
    st.close()


def _unitary_customs(gates, customs, rng):
    for i in customs:
        k = 1 + len(gates[i][2])
        customs[i] = (W.random_unitary(k, int(rng.integers(1 << 30))).T.copy(), np.ones(1 << k, dtype=np.uint8))
    return customs


def test_adjoint_fused_vs_per_op():
    """The fused reverse sweep (psi and lam in one tile, gradient taps inside) against the
    one-launch-per-op reverse sweep, on random programs with every gate kind."""
    rng = np.random.default_rng(808)
    for trial in range(10):
        n = int(rng.integers(2, 15))
        gates, customs = H.random_gates(n, 60, rng, with_custom=(trial % 2 == 0))
        customs = _unitary_customs(gates, customs, rng)
        terms = []
        for _ in range(4):
            m = int(rng.integers(1, min(n, 3) + 1))
            wires = rng.permutation(n)[:m]
            terms.append((float(rng.uniform(-1, 1)), [(int(w), "XYZ"[int(rng.integers(3))]) for w in wires]))
        obs = _capi.Observable(terms)
        p = H.to_program(gates, customs)
        out = []
        for fused_backward in (True, False):
            say("adjoint trial", trial, "n", n, "fused backward", fused_backward)
            g = _capi.Grad(n)
            g.set_options(jit=2 if fused_backward else 0)
            e = g.forward(p, obs)
            out.append((e, g.backward(1.0)))
            g.close()
        assert abs(out[0][0] - out[1][0]) < 1e-12
        assert np.max(np.abs(out[0][1] - out[1][1])) < 1e-11, trial


@pytest.mark.parametrize("precision,tol", [(_capi.QR_C128, 1e-9), (_capi.QR_C64, 1e-5)])
def test_adjoint_hea_20q(precision, tol):
    """Config of the adjoint benchmark (20 qubits, 10 layers, 400 parameters) against the
    parameter-shift rule on the plain complex128 path, for a sample of parameters."""
    n, L = 20, 10
    gates, thetas = W.hea(n, L, seed=2020)
    obs = _capi.Observable(W.hea_observable(n))
    say("hea 20q adjoint precision", precision)
    g = _capi.Grad(n, precision)
    g.forward(W.to_program(gates), obs)
    grads = g.backward()
    g.close()
    assert len(grads) == 400
    st = _capi.State(n)
    pidx = [i for i, x in enumerate(gates) if x[0] in ("Rx", "Ry", "Rz")]
    sample = list(range(0, 400, 37)) + [399]
    scale = 0.0
    errs = []
    for j in sample:
        i = pidx[j]
        vals = []
        for s in (+1, -1):
            kind, pos, ctl, arg, argi = gates[i]
            gg = list(gates)
            gg[i] = (kind, pos, ctl, arg + s * np.pi / 2, argi)
            st.reset()
            st.apply(W.to_program(gg))
            vals.append(st.expectation(obs))
        ps = 0.5 * (vals[0] - vals[1])
        scale = max(scale, abs(ps))
        errs.append(abs(grads[j] - ps))
    st.close()
    assert max(errs) / scale < tol, (max(errs), scale)


def _random_instance(rng, n, with_custom):
    gates, customs = H.random_gates(n, 40, rng, with_custom=with_custom)
    customs = _unitary_customs(gates, customs, rng)
    terms = []
    for _ in range(3):
        m = int(rng.integers(1, min(n, 3) + 1))
        wires = rng.permutation(n)[:m]
        terms.append((float(rng.uniform(-1, 1)), [(int(w), "XYZ"[int(rng.integers(3))]) for w in wires]))
    return gates, customs, terms


def test_batch_matches_single_instances():
    """One captured graph over uneven instances against one gradient context per instance."""
    rng = np.random.default_rng(909)
    inst, ref = [], []
    for i in range(12):
        n = int(rng.integers(1, 15))
        gates, customs, terms = _random_instance(rng, n, with_custom=(i % 3 == 0))
        p = H.to_program(gates, customs)
        obs = _capi.Observable(terms)
        inst.append((n, p, obs))
        g = _capi.Grad(n)
        e = g.forward(p, obs)
        ref.append((e, g.backward(1.0)))
        g.close()
    say("batch of", len(inst))
    b = _capi.Batch(inst, _capi.QR_C128)
    ev, grads, lat = b.run()
    for i, (e, gr) in enumerate(ref):
        assert abs(ev[i] - e) < 1e-11, i
        assert len(grads[i]) == len(gr)
        assert np.max(np.abs(grads[i] - gr), initial=0.0) < 1e-11, i
    assert np.all(lat > 0)
    b.close()


def test_batch_rerun_new_angles_and_c64():
    """A second run of the captured graph with new angles, and the complex64 batch against the
    complex128 one."""
    rng = np.random.default_rng(910)
    sizes = [3, 7, 10, 12, 13, 16]
    insts = []
    for n in sizes:
        gates, thetas = W.hea(n, 3, seed=int(rng.integers(1 << 30)))
        insts.append((n, gates))
    obs = [_capi.Observable(W.hea_observable(n)) for n in sizes]
    b128 = _capi.Batch([(n, W.to_program(g), o) for (n, g), o in zip(insts, obs)], _capi.QR_C128)
    b64 = _capi.Batch([(n, W.to_program(g), o) for (n, g), o in zip(insts, obs)], _capi.QR_C64)
    for rep in range(2):
        new = []
        for n, gates in insts:
            gg = [(k, p, c, (a + 0.37 * (rep + 1)) if k in ("Ry", "Rz") else a, ai) for (k, p, c, a, ai) in gates]
            new.append((n, gg))
        progs = [W.to_program(g) for _, g in new]
        say("batch rerun", rep)
        ev, grads, _ = b128.run(progs)
        ev64, grads64, _ = b64.run(progs)
        for i, (n, g) in enumerate(new):
            gc = _capi.Grad(n)
            e = gc.forward(W.to_program(g), obs[i])
            gr = gc.backward()
            gc.close()
            assert abs(ev[i] - e) < 1e-11
            assert np.max(np.abs(grads[i] - gr)) < 1e-11
            scale = max(1.0, float(np.max(np.abs(gr))))
            assert np.max(np.abs(grads64[i] - gr)) / scale < 1e-4, (i, n)
    b128.close()
    b64.close()


def test_batch_rejects_structure_change():
    n = 4
    gates, _ = W.hea(n, 2, seed=1)
    b = _capi.Batch([(n, W.to_program(gates), _capi.Observable(W.hea_observable(n)))], _capi.QR_C128)
    with pytest.raises(_capi.QrError):
        b.run([W.to_program(gates[:-1])])
    b.close()


def test_batch_autotuned_widths_match():
    """Tuned reverse-sweep tile widths change the plan, not the results."""
    rng = np.random.default_rng(911)
    items = []
    for n in (16, 17, 12):
        gates, _ = W.hea(n, 2, seed=int(rng.integers(1 << 30)))
        items.append((n, W.to_program(gates), _capi.Observable(W.hea_observable(n))))
    say("batch autotune")
    b0 = _capi.Batch(items, _capi.QR_C128)
    b1 = _capi.Batch(items, _capi.QR_C128, autotune=True)
    b2 = _capi.Batch(items, _capi.QR_C128, dual_tile_bits=9, fwd_tile_bits=11)
    e0, g0, _ = b0.run()
    for b in (b1, b2):
        e, g, _ = b.run()
        assert np.max(np.abs(e - e0)) < 1e-11
        for x, y in zip(g, g0):
            assert np.max(np.abs(x - y)) < 1e-11
        b.close()
    b0.close()
