"""GPU parity at the benchmark sizes: the headline brickwork at 28 and 30 qubits, the capacity
circuits at 33 qubits (complex128) and 34 qubits (complex64), the complex64 drift at 30 qubits,
and the 20-qubit adjoint configuration against the oracle.

The chain of evidence at full size:
* the reference-order mode (fused = 0) is bitwise equal to the oracle's restatement of
  Circuit::apply_gate (simulation.rs:64-135) on a dense 30-qubit register;
* the default path (balanced fused plan, compiled stages) matches the reference-order mode on
  the whole depth-40 brickwork at 28 and 30 qubits: max |d amp| <= 1e-12 and
  ||d psi||_2 / ||psi||_2 <= 1e-12;
* circuit followed by its inverse returns |0...0> (norm and amplitude 0 within 1e-12).
"""
import math
import sys

import numpy as np
import pytest

import helpers as H
from quantr_amd import _capi
from quantr_amd import workloads as W

pytestmark = pytest.mark.gpu

CHUNK = 1 << 24


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if _capi.device_count() < 1:
        pytest.fail("no GPU visible: the HIP engine cannot run")


def compare_states(a, b, dim):
    """max |a - b|, ||a - b||_2 and ||b||_2 over two device states, read back in chunks."""
    dmax = 0.0
    d2 = 0.0
    n2 = 0.0
    for first in range(0, dim, CHUNK):
        cnt = min(CHUNK, dim - first)
        x = a.amplitudes_range(first, cnt)
        y = b.amplitudes_range(first, cnt)
        d = np.abs(x - y)
        dmax = max(dmax, float(d.max()))
        d2 += float(np.dot(d, d))
        n2 += float(np.vdot(y, y).real)
    return dmax, math.sqrt(d2), math.sqrt(n2)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("n", [28, 30])
def test_headline_brickwork_fused_vs_reference_order(n):
    """Depth-40 brickwork (seed 1234), complex128, from |0...0>: the benchmark's own plan
    (balanced stages, compiled at once from 28 qubits) against one sweep per gate in the
    reference's accumulation order (simulation.rs:64-135)."""
    gates = W.brickwork(n, depth=40, seed=1234)
    assert len(gates) == {28: 1668, 30: 1788}[n]
    prog = W.to_program(gates).c_struct()
    say("headline %dq: fused default path" % n)
    fused = _capi.State(n)
    fused.apply(prog)
    plan = fused.last_plan()
    fused.synchronize()
    say("headline %dq: %d sweeps (%d tile); reference-order mode" % (n, plan["sweeps"], plan["tile_sweeps"]))
    assert plan["tile_sweeps"] >= 20 and plan["sweeps"] < 60
    exact = _capi.State(n)
    exact.set_options(fused=0)
    exact.apply(prog)
    exact.synchronize()
    say("headline %dq: comparing" % n)
    dmax, dl2, nrm = compare_states(fused, exact, 1 << n)
    say("headline %dq: max|d amp| %.3e  ||d psi||_2 %.3e  ||psi||_2 %.15f" % (n, dmax, dl2, nrm))
    fused.close()
    exact.close()
    assert abs(nrm - 1.0) < 1e-12
    assert dmax <= 1e-12
    assert dl2 / nrm <= 1e-12


@pytest.mark.timeout(600)
def test_reference_order_vs_oracle_30q_dense_register():
    """The reference-order kernels against the oracle, bitwise, on a dense 30-qubit complex128
    register (every amplitude nonzero), for gates of every kind the headline uses on the
    lowest, middle and highest wires, plus a 3-qubit custom gate (the capacity config's)."""
    n = 30
    dim = 1 << n
    rng = np.random.default_rng(3030)
    say("dense 30q register: generating")
    psi = np.empty(dim, dtype=np.complex128)
    for first in range(0, dim, CHUNK):
        cnt = min(CHUNK, dim - first)
        psi.real[first:first + cnt] = (rng.random(cnt) - 0.5) * 2.0 ** -14
        psi.imag[first:first + cnt] = (rng.random(cnt) - 0.5) * 2.0 ** -14
    u3 = W.random_unitary(3, 31)
    gates = [("H", 0, (), 0.0, 0), ("Rx", 29, (), 1.234, 0), ("Ry", 14, (), 4.321, 0), ("Rz", 1, (), 2.5, 0),
             ("T", 28, (), 0.0, 0), ("S", 15, (), 0.0, 0), ("Y", 2, (), 0.0, 0), ("X", 27, (), 0.0, 0),
             ("Z", 16, (), 0.0, 0), ("CNot", 1, (0,), 0.0, 0), ("CNot", 29, (28,), 0.0, 0),
             ("Toffoli", 13, (2, 27), 0.0, 0), ("Custom", 20, (5, 25), 0.0, 0)]
    customs = {len(gates) - 1: (np.ascontiguousarray(u3.T), np.ones(8, dtype=np.uint8))}
    p = W.to_program(gates, customs)
    st = _capi.State(n)
    st.set_options(fused=0)
    say("dense 30q register: upload and GPU reference-order run")
    st.set_amplitudes(psi)
    st.apply(p)
    st.synchronize()
    say("dense 30q register: oracle (state vector form, 16 threads)")
    H.oracle_apply_inplace(n, psi, p, threads=16)
    say("dense 30q register: comparing bitwise")
    for first in range(0, dim, CHUNK):
        cnt = min(CHUNK, dim - first)
        got = st.amplitudes_range(first, cnt)
        assert np.array_equal(got.view(np.float64), psi[first:first + cnt].view(np.float64)), first
    st.close()


@pytest.mark.timeout(600)
def test_headline_round_trip_30q():
    """The 30-qubit headline circuit followed by its inverse returns |0...0> on the default path."""
    n = 30
    gates = W.brickwork(n, depth=40, seed=1234)
    inv, _ = W.inverse(gates)
    st = _capi.State(n)
    say("round trip 30q: forward")
    st.apply(W.to_program(gates).c_struct())
    say("round trip 30q: inverse")
    st.apply(W.to_program(inv).c_struct())
    a = st.amplitudes_range(0, 4)
    norm = st.norm_sqr()
    st.close()
    say("round trip 30q: |a0 - 1| %.3e, |norm - 1| %.3e, |a1| %.3e" % (abs(a[0] - 1), abs(norm - 1), abs(a[1])))
    assert abs(a[0] - 1.0) < 1e-12
    assert abs(norm - 1.0) < 1e-12
    assert np.max(np.abs(a[1:])) < 1e-12


@pytest.mark.timeout(900)
@pytest.mark.parametrize("n,prec,tol", [(33, _capi.QR_C128, 1e-12), (34, _capi.QR_C64, 2e-5)])
def test_capacity_round_trip(n, prec, tol):
    """The largest states one GPU holds (137 GB): brickwork depth 40 plus a seeded 3-qubit Haar
    custom gate, then the inverse program, from |0...0> (super_positions_unchecked.rs:39-46)."""
    gates, cust = W.capacity_circuit(n, depth=40)
    inv, icust = W.inverse(gates, cust)
    say("capacity %dq: forward (%d gates)" % (n, len(gates)))
    st = _capi.State(n, prec)
    st.apply(W.to_program(gates, cust).c_struct())
    plan = st.last_plan()
    norm = st.norm_sqr()
    say("capacity %dq: %d sweeps, |norm - 1| %.3e; inverse" % (n, plan["sweeps"], abs(norm - 1)))
    st.apply(W.to_program(inv, icust).c_struct())
    a = st.amplitudes_range(0, 2)
    back = st.norm_sqr()
    st.close()
    say("capacity %dq: round trip |a0 - 1| %.3e, |norm - 1| %.3e" % (n, abs(a[0] - 1), abs(back - 1)))
    assert abs(norm - 1.0) < tol
    assert abs(a[0] - 1.0) < tol
    assert abs(back - 1.0) < tol


@pytest.mark.timeout(600)
def test_complex64_drift_30q():
    """complex64 against complex128 on the 30-qubit headline circuit. Bounds hold the measured
    drift (max |d amp| 2.0e-10, max |d p| 4.9e-14, |norm - 1| 2.3e-6) with a margin of 10-40x."""
    n = 30
    prog = W.to_program(W.brickwork(n, depth=40, seed=1234)).c_struct()
    say("drift 30q: complex128 and complex64 runs")
    s128 = _capi.State(n, _capi.QR_C128)
    s64 = _capi.State(n, _capi.QR_C64)
    s128.apply(prog)
    s64.apply(prog)
    norm64 = s64.norm_sqr()
    da = dp = 0.0
    for first in range(0, 1 << n, CHUNK):
        a = s128.amplitudes_range(first, CHUNK)
        b = s64.amplitudes_range(first, CHUNK)
        da = max(da, float(np.max(np.abs(a - b))))
        dp = max(dp, float(np.max(np.abs(np.abs(a) ** 2 - np.abs(b) ** 2))))
    s128.close()
    s64.close()
    say("drift 30q: max|d amp| %.3e, max|d p| %.3e, |norm64 - 1| %.3e" % (da, dp, abs(norm64 - 1)))
    assert da < 5e-9
    assert dp < 1e-12
    assert abs(norm64 - 1.0) < 1e-4


@pytest.fixture(scope="module")
def hea20_reference():
    """The 20-qubit, 10-layer HEA (400 parameters, seed 2020) and, for a sample of parameters,
    the oracle's parameter-shift and central finite-difference (h = 1e-6, binary64) gradients."""
    n, L = 20, 10
    gates, thetas = W.hea(n, L, seed=2020)
    obs = _capi.Observable(W.hea_observable(n))
    pidx = [i for i, x in enumerate(gates) if x[0] in ("Rx", "Ry", "Rz")]
    sample = [0, 37, 133, 200, 281, 399]

    def energy(gg):
        psi = H.oracle_simulate(n, W.to_program(gg), threads=16)
        return H.oracle_expectation(n, psi, obs)

    ps, fd = [], []
    h = 1e-6
    for j in sample:
        i = pidx[j]
        kind, pos, ctl, arg, argi = gates[i]
        vals = []
        for d in (math.pi / 2, -math.pi / 2, h, -h):
            gg = list(gates)
            gg[i] = (kind, pos, ctl, arg + d, argi)
            vals.append(energy(gg))
        ps.append(0.5 * (vals[0] - vals[1]))
        fd.append((vals[2] - vals[3]) / (2 * h))
        say("hea 20q oracle: param %d shift %.15f fd %.15f" % (j, ps[-1], fd[-1]))
    return n, gates, obs, sample, np.array(ps), np.array(fd)


@pytest.mark.timeout(900)
def test_adjoint_hea_20q_vs_oracle(hea20_reference):
    """The adjoint gradient of the 20-qubit HEA against the oracle.

    complex128: the sampled gradients against the oracle's parameter-shift rule within 1e-9 of
    the largest sampled gradient, and against central finite differences (h = 1e-6 in binary64,
    whose own rounding error is about 1e-8) within 1e-7.
    complex64: all 400 gradients within 1e-5 of the largest complex128 gradient (the relative
    gradient error of the benchmark), and the sampled ones within the same bound of the oracle's
    parameter-shift values."""
    n, gates, obs, sample, ps, fd = hea20_reference
    out = {}
    for prec in (_capi.QR_C128, _capi.QR_C64):
        g = _capi.Grad(n, prec)
        g.forward(W.to_program(gates), obs)
        out[prec] = g.backward()
        g.close()
    g128, g64 = out[_capi.QR_C128], out[_capi.QR_C64]
    scale_s = float(np.max(np.abs(ps)))
    e_ps = float(np.max(np.abs(g128[sample] - ps))) / scale_s
    e_fd = float(np.max(np.abs(g128[sample] - fd))) / scale_s
    scale = float(np.max(np.abs(g128)))
    e64 = float(np.max(np.abs(g64 - g128))) / scale
    e64_ps = float(np.max(np.abs(g64[sample] - ps))) / scale
    say("hea 20q adjoint c128: rel err vs shift %.3e, vs fd %.3e; c64: rel err vs c128 %.3e, vs shift %.3e "
        "(max |grad| %.4f)" % (e_ps, e_fd, e64, e64_ps, scale))
    assert float(np.max(np.abs(ps - fd))) / scale_s < 1e-7
    assert e_ps < 1e-9
    assert e_fd < 1e-7
    assert e64 < 1e-5
    assert e64_ps < 1e-5


@pytest.mark.timeout(1200)
def test_batch_512_instances():
    """The batched-serving configuration: 512 HEA instances, n in 10..24, L in 2..12, complex64,
    one graph launch. Every instance against the single-instance adjoint path (complex64), and a
    seeded sample against the oracle (complex128): the expectation value of the shallowest
    24-qubit instance, and the parameter-shift gradient of two parameters each for three small
    instances."""
    inst = W.batch_instances(512)
    obs = {n: _capi.Observable(W.hea_observable(n)) for n in range(10, 25)}
    items = [(n, W.to_program(g), obs[n]) for n, _, g, _ in inst]
    say("batch 512: creating (plans, stage compilation, capture)")
    b = _capi.Batch(items, _capi.QR_C64)
    say("batch 512: running")
    ev, grads, lat = b.run()
    b.close()
    assert len(ev) == 512 and np.all(lat > 0)
    worst_e = worst_g = 0.0
    for i, (n, L, gates, thetas) in enumerate(inst):
        g = _capi.Grad(n, _capi.QR_C64)
        e = g.forward(W.to_program(gates), obs[n])
        gr = g.backward()
        g.close()
        assert len(grads[i]) == len(thetas) == len(gr)
        scale = max(1.0, float(np.max(np.abs(gr))))
        worst_e = max(worst_e, abs(ev[i] - e) / max(1.0, abs(e)))
        worst_g = max(worst_g, float(np.max(np.abs(grads[i] - gr))) / scale)
        if i % 64 == 0:
            say("batch 512: instance %d (n %d, L %d) checked" % (i, n, L))
    say("batch 512: vs single-instance complex64: worst rel |d E| %.3e, worst rel |d grad| %.3e" % (worst_e, worst_g))
    assert worst_e < 1e-5 and worst_g < 1e-5
    # oracle sample (complex128 arithmetic)
    deep24 = [i for i, x in enumerate(inst) if x[0] == 24]
    j24 = min(deep24, key=lambda i: inst[i][1])
    n, L, gates, _ = inst[j24]
    say("batch 512: oracle expectation of instance %d (n %d, L %d)" % (j24, n, L))
    e_ref = H.oracle_expectation(n, H.oracle_simulate(n, W.to_program(gates), threads=16), obs[n])
    assert abs(ev[j24] - e_ref) < 1e-5 * max(1.0, abs(e_ref)), (ev[j24], e_ref)
    small = [i for i, x in enumerate(inst) if x[0] <= 13][:3]
    for i in small:
        n, L, gates, thetas = inst[i]
        pidx = [k for k, x in enumerate(gates) if x[0] in ("Ry", "Rz")]
        for j in (0, len(pidx) - 1):
            k = pidx[j]
            vals = []
            for s in (+1, -1):
                gg = list(gates)
                kind, pos, ctl, a, ai = gg[k]
                gg[k] = (kind, pos, ctl, a + s * math.pi / 2, ai)
                vals.append(H.oracle_expectation(n, H.oracle_simulate(n, W.to_program(gg), threads=16), obs[n]))
            ps = 0.5 * (vals[0] - vals[1])
            scale = max(1.0, float(np.max(np.abs(grads[i]))))
            say("batch 512: instance %d param %d: batch %.9f oracle shift %.9f" % (i, j, grads[i][j], ps))
            assert abs(grads[i][j] - ps) < 1e-5 * scale
