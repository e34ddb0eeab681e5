"""qr_expectation_noalloc through the C ABI: the allocation-free expectation value equals
qr_expectation's bit for bit, on two calls in a row (the second with a smaller state and a
smaller observable), and sizes beyond its fixed arrays are rejected before any GPU work."""
import ctypes as C
import sys

import numpy as np
import pytest

import helpers as H
from quantr_amd import _capi
from quantr_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc(st, obs):
    o = obs.c_struct()
    v = C.c_double(0.0)
    _capi.check(_capi.lib().qr_expectation_noalloc(st.handle, C.byref(o), C.byref(v)))
    return v.value


def random_observable(n, terms, rng):
    out = []
    for _ in range(terms):
        m = int(rng.integers(1, min(n, 4) + 1))
        wires = rng.permutation(n)[:m]
        out.append((float(rng.uniform(-1, 1)), [(int(w), "XYZ"[int(rng.integers(3))]) for w in wires]))
    return _capi.Observable(out)


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if _capi.device_count() < 1:
        pytest.fail("no GPU visible: the HIP engine cannot run")


def test_noalloc_expectation_matches_two_calls():
    rng = np.random.default_rng(7070)
    cases = [(20, 300, _capi.QR_C128), (12, 5, _capi.QR_C128), (16, 40, _capi.QR_C64), (10, 3, _capi.QR_C64)]
    for n, terms, prec in cases:
        say("noalloc: n %d, %d terms, precision %d" % (n, terms, prec))
        st = _capi.State(n, prec)
        st.apply(W.to_program(W.brickwork(n, depth=6, seed=n)).c_struct())
        obs = random_observable(n, terms, rng)
        want = st.expectation(obs)
        got = noalloc(st, obs)
        assert got == want, (n, terms, got, want)
        again = noalloc(st, obs)  # the array keeps the last call's contents; results do not depend on them
        assert again == want
        psi = st.amplitudes()
        assert abs(got - H.oracle_expectation(n, psi, obs)) < 1e-11 * max(1.0, terms)
        st.close()


def test_noalloc_rejects_oversize_and_bad_terms():
    st = _capi.State(4)
    big = _capi.Observable([(1.0, [(0, "Z")])] * 16385)
    with pytest.raises(_capi.QrError) as e:
        noalloc(st, big)
    assert "at most 16384" in str(e.value)
    with pytest.raises(_capi.QrError) as e:
        noalloc(st, _capi.Observable([(1.0, [(4, "Z")])]))
    assert str(e.value) == "observable wire 4 is out of range"
    assert noalloc(st, _capi.Observable([])) == 0.0
    st.close()
