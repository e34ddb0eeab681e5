"""measure_all's sampler against the reference's left fold (super_positions.rs:332-342).

The default sampler rebuilds the sequential fold S_i = fl(S_{i-1} + re_i^2 + im_i^2) exactly from
block-parallel integer sums; every test requires identical indices (not a high agreement rate)
against the oracle's fold (orc_measure) and the engine's single-lane fold (exact_sample = 1), with
draws placed exactly on, just below and just above fold values S_i, on states built to hit
every path: binade crossings, halfway cases, zero blocks, subnormal and huge probabilities.
"""
import sys

import numpy as np
import pytest

import helpers as H
from quantr_amd import _capi
from quantr_amd import workloads as W

pytestmark = pytest.mark.gpu
NONE = np.iinfo(np.uint64).max


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if _capi.device_count() < 1:
        pytest.fail("no GPU visible: the HIP engine cannot run")


def left_fold(psi):
    """The reference's cumulative probabilities: re*re + im*im (no fused multiply-add) folded in
    index order (np.add.accumulate is a sequential fold)."""
    p = psi.real * psi.real + psi.imag * psi.imag
    return np.add.accumulate(p)


def boundary_draws(S, rng, count):
    """Draws on, just below and just above fold values, plus uniforms, 0, the total and beyond."""
    idx = rng.integers(0, S.size, count)
    v = S[idx]
    u = np.concatenate([v, np.nextafter(v, -np.inf), np.nextafter(v, np.inf), rng.random(count),
                        [0.0, S[-1], np.nextafter(S[-1], -np.inf), np.nextafter(S[-1], np.inf), 0.999999, -0.5,
                         np.inf]])
    return u[~np.isnan(u)]


def states(n, rng):
    dim = 1 << n
    out = {}
    x = rng.standard_normal(dim) + 1j * rng.standard_normal(dim)
    out["random"] = x / np.linalg.norm(x)
    g = np.zeros(dim, dtype=np.complex128)
    g[0] = g[-1] = np.sqrt(0.5)
    out["ghz"] = g
    b = np.zeros(dim, dtype=np.complex128)
    b[-1] = 1.0
    out["last_basis"] = b
    out["uniform"] = np.full(dim, 2.0 ** (-n / 2.0), dtype=np.complex128)
    t = x.copy()
    t[: dim // 2] *= 1e-150  # the first half's fold runs through very small binades
    out["tiny_then_normal"] = t / np.linalg.norm(t)
    s = x.copy()
    s[: dim // 3] *= 1e-162  # probabilities below the normal range (subnormal)
    out["subnormal_head"] = s / np.linalg.norm(s)
    out["norm_1000"] = x / np.linalg.norm(x) * np.sqrt(1000.0)
    q = np.zeros(dim, dtype=np.complex128)
    q[rng.integers(0, dim, 64)] = rng.standard_normal(64) * (2.0 ** -3)  # sparse, dyadic-ish values
    out["sparse"] = q
    h = np.zeros(dim, dtype=np.complex128)
    h[::3] = 2.0 ** -27  # p = 2^-54: halfway cases once the fold passes 0.5... of its own scale
    h[1::3] = 2.0 ** -(n // 2)
    out["halfway"] = h
    return out


@pytest.mark.parametrize("n", [3, 11, 12, 13, 14, 15, 17, 20])
@pytest.mark.parametrize("precision", [_capi.QR_C128, _capi.QR_C64])
def test_sampler_matches_left_fold(n, precision):
    rng = np.random.default_rng(1000 + n)
    for name, psi in states(n, rng).items():
        st = _capi.State(n, precision)
        st.set_amplitudes(psi)
        held = st.amplitudes()  # complex64 states hold rounded values; the fold runs on those
        S = left_fold(held)
        u = boundary_draws(S, rng, 300)
        say("sampler n %d prec %d state %s: %d draws" % (n, precision, name, u.size))
        got = st.sample(u)
        want = H.oracle_measure(n, held, u)
        assert np.array_equal(got, want), (n, name)
        ref = np.searchsorted(S, u, side="right").astype(np.uint64)  # first i with u < S_i
        ref[ref >= S.size] = NONE
        assert np.array_equal(got, ref), (n, name)
        st.set_options(exact_sample=1)
        assert np.array_equal(st.sample(u), want), (n, name)
        st.close()


@pytest.mark.timeout(600)
def test_sampler_28q_boundaries():
    """A 28-qubit brickwork state: draws on, below and above 2000 fold values computed from
    the state itself, plus 5000 uniforms, against the single-lane fold and the host fold."""
    n = 28
    st = _capi.State(n)
    st.apply(W.to_program(W.brickwork(n, depth=10, seed=2828)).c_struct())
    say("sampler 28q: reading the state back")
    psi = st.amplitudes()
    S = left_fold(psi)
    del psi
    rng = np.random.default_rng(28)
    u = boundary_draws(S, rng, 2000)
    say("sampler 28q: %d draws, default sampler" % u.size)
    got = st.sample(u)
    ref = np.searchsorted(S, u, side="right").astype(np.uint64)
    ref[ref >= S.size] = NONE
    assert np.array_equal(got, ref)
    say("sampler 28q: single-lane fold")
    st.set_options(exact_sample=1)
    assert np.array_equal(st.sample(u), ref)
    st.close()
