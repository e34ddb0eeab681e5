"""Test helpers: the oracle binding (tests only), fixture conversion, numpy brute force."""
import ctypes as C
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_vectors.json")

from quantr_amd import _capi  # noqa: E402
from quantr_amd._capi import OP_DTYPE, TERM_DTYPE, KIND, Program  # noqa: E402

_orc = None


def oracle():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            import subprocess
            subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "_build/liboracle.so"])
        l = C.CDLL(ORACLE_SO)
        sig = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p,
               C.c_uint64, C.c_void_p, C.c_uint64]
        l.orc_simulate_hashmap.argtypes = sig
        l.orc_simulate_sv.argtypes = sig + [C.c_int]
        l.orc_measure.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        l.orc_measure_literal.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        l.orc_expectation.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                      C.c_void_p]
        l.orc_last_error.restype = C.c_char_p
        assert l.orc_op_size() == OP_DTYPE.itemsize
        _orc = l
    return _orc


def _arr(a):
    return a.ctypes.data if a.size else None


def oracle_simulate(n, program, register=None, method="sv", threads=0):
    """Runs the oracle on `program` (a _capi.Program) from |0..0> or `register`."""
    ops, ctl, img, some = program.arrays()
    state = np.zeros(1 << n, dtype=np.complex128)
    if register is None:
        state[0] = 1.0
    else:
        state[:] = register
    o = oracle()
    if method == "hashmap":
        rc = o.orc_simulate_hashmap(n, state.ctypes.data, _arr(ops), len(ops), _arr(ctl), len(ctl), _arr(img),
                                    len(img), _arr(some), len(some))
    else:
        rc = o.orc_simulate_sv(n, state.ctypes.data, _arr(ops), len(ops), _arr(ctl), len(ctl), _arr(img),
                               len(img), _arr(some), len(some), threads)
    if rc != 0:
        raise RuntimeError(o.orc_last_error().decode())
    return state


def oracle_apply_inplace(n, state, program, threads=0):
    """Applies `program` to `state` (complex128, 2^n) in place with the oracle state vector."""
    ops, ctl, img, some = program.arrays()
    rc = oracle().orc_simulate_sv(n, state.ctypes.data, _arr(ops), len(ops), _arr(ctl), len(ctl), _arr(img),
                                  len(img), _arr(some), len(some), threads)
    if rc != 0:
        raise RuntimeError(oracle().orc_last_error().decode())


def oracle_measure(n, state, uniforms, literal=False):
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    out = np.empty(u.size, dtype=np.uint64)
    st = np.ascontiguousarray(state, dtype=np.complex128)
    f = oracle().orc_measure_literal if literal else oracle().orc_measure
    f(n, st.ctypes.data, u.ctypes.data, u.size, out.ctypes.data)
    return out


def oracle_expectation(n, state, observable):
    o = observable.c_struct()
    v = C.c_double()
    st = np.ascontiguousarray(state, dtype=np.complex128)
    t, w, p = o._keep
    rc = oracle().orc_expectation(n, st.ctypes.data, _arr(t), len(t), _arr(w), _arr(p), C.byref(v))
    assert rc == 0
    return v.value


# ---- fixtures -----------------------------------------------------------------------------

def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def table_arrays(t, simulate_nested):
    """Custom-gate table of the fixture -> (images[j][r], some[j])."""
    k = t["k"]
    dim = 1 << k
    images = np.zeros((dim, dim), dtype=np.complex128)
    some = np.array(t["some"], dtype=np.uint8) if "some" in t else np.ones(dim, dtype=np.uint8)
    if "images" in t:
        for j in range(dim):
            for r in range(dim):
                images[j, r] = complex(*t["images"][j][r])
    elif "sparse" in t:
        for j, r, re, im in t["sparse"]:
            images[int(j), int(r)] = complex(re, im)
    else:  # nested circuit simulated on each basis register (tests/qft.rs:51-67)
        for j in range(dim):
            reg = np.zeros(dim, dtype=np.complex128)
            reg[j] = 1.0
            images[j] = simulate_nested(k, program_from_fixture(t["program"], None), reg)
    return images, some


def program_from_fixture(prog, tables, simulate_nested=None):
    p = Program()
    for gte in prog:
        if gte["kind"] == "Custom":
            images, some = table_arrays(tables[gte["table"]], simulate_nested)
            p.add_custom(gte["pos"], gte["controls"], images, some)
        else:
            p.add(gte["kind"], gte["pos"], gte["controls"], gte["arg"], gte["arg_int"])
    return p


def cvec(lst):
    return np.array([complex(a, b) for a, b in lst], dtype=np.complex128)


# ---- numpy brute force (dense Kronecker products, small n) ---------------------------------

S2 = 1.0 / np.sqrt(2.0)


def textbook_matrix(kind, arg=0.0, arg_int=0):
    """Textbook matrices (numpy), independent of the engine's and the oracle's tables.
    Acting wires ordered [controls..., target]."""
    c, s = np.cos(arg / 2), np.sin(arg / 2)
    one = {
        "H": np.array([[1, 1], [1, -1]]) * S2,
        "X": np.array([[0, 1], [1, 0]]),
        "Y": np.array([[0, -1j], [1j, 0]]),
        "Z": np.diag([1, -1]),
        "S": np.diag([1, 1j]),
        "Sdag": np.diag([1, -1j]),
        "T": np.diag([1, np.exp(1j * np.pi / 4)]),
        "Tdag": np.diag([1, np.exp(-1j * np.pi / 4)]),
        "Rx": np.array([[c, -1j * s], [-1j * s, c]]),
        "Ry": np.array([[c, -s], [s, c]]),
        "Rz": np.diag([np.exp(-1j * arg / 2), np.exp(1j * arg / 2)]),
        "Phase": np.exp(1j * arg / 2) * np.eye(2),
        # quantr's X90/Y90 family as implemented (standard_gate_ops.rs:97-127):
        "X90": -1j * np.array([[0, 1], [1, 0]]),
        "MX90": 1j * np.array([[0, 1], [1, 0]]),
        "Y90": np.array([[0, 1], [-1, 0]]),
        "MY90": np.array([[0, -1], [1, 0]]),
    }
    if kind in one:
        return np.asarray(one[kind], dtype=np.complex128)

    def controlled(u, nc):
        d = 2 ** (nc + 1)
        m = np.eye(d, dtype=np.complex128)
        m[d - 2:, d - 2:] = u
        return m

    if kind == "CNot":
        return controlled(one["X"], 1)
    if kind == "CY":
        return controlled(one["Y"], 1)
    if kind == "CZ":
        return controlled(one["Z"], 1)
    if kind == "CR":
        return controlled(np.diag([1, np.exp(1j * arg)]), 1)
    if kind == "CRk":
        return controlled(np.diag([1, np.exp(2j * np.pi / 2 ** arg_int)]), 1)
    if kind == "Toffoli":
        return controlled(one["X"], 2)
    if kind == "Swap":
        return np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)
    raise ValueError(kind)


def custom_matrix(images, some):
    """Dense matrix of a custom table with the reference's None-overwrite rule."""
    dim = images.shape[0]
    m = np.zeros((dim, dim), dtype=np.complex128)
    for r in range(dim):
        if not some[r]:
            m[r, r] = 1.0
            continue
        for j in range(dim):
            if some[j]:
                m[r, j] = images[j, r]
    return m


def embed(n, wires, u):
    """Full 2^n x 2^n operator of u acting on `wires` (MSB first), wire 0 = MSB of the index."""
    k = len(wires)
    dim = 1 << n
    full = np.zeros((dim, dim), dtype=np.complex128)
    bits = [n - 1 - w for w in wires]
    for i in range(dim):
        li = 0
        for b in bits:
            li = (li << 1) | ((i >> b) & 1)
        base = i
        for b in bits:
            base &= ~(1 << b)
        for lo in range(1 << k):
            o = base
            for q, b in enumerate(bits):
                if (lo >> (k - 1 - q)) & 1:
                    o |= 1 << b
            full[o, i] += u[lo, li]
    return full


def brute_force(n, gates, customs=None, register=None):
    """gates: list of (kind, pos, controls, arg, arg_int); customs: index -> (images, some)."""
    psi = np.zeros(1 << n, dtype=np.complex128)
    if register is None:
        psi[0] = 1.0
    else:
        psi[:] = register
    for i, (kind, pos, ctl, arg, argi) in enumerate(gates):
        if kind == "Custom":
            u = custom_matrix(*customs[i])
        else:
            u = textbook_matrix(kind, arg, argi)
        psi = embed(n, list(ctl) + [pos], u) @ psi
    return psi


def random_gates(n, count, rng, kinds=None, with_custom=False):
    """Random gate list over every standard kind (and optionally custom tables)."""
    singles = ["H", "X", "Y", "Z", "S", "Sdag", "T", "Tdag", "Rx", "Ry", "Rz", "X90", "Y90", "MX90",
               "MY90", "Phase"]
    doubles = ["CR", "CRk", "CZ", "CY", "CNot", "Swap"]
    kinds = kinds or (singles + (doubles if n >= 2 else []) + (["Toffoli"] if n >= 3 else []))
    gates, customs = [], {}
    for i in range(count):
        if with_custom and n >= 2 and rng.random() < 0.15:
            k = int(rng.integers(1, min(n, 3) + 1))
            wires = [int(x) for x in rng.permutation(n)[:k]]
            dim = 1 << k
            images = rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))
            images /= np.linalg.norm(images, axis=1, keepdims=True)
            some = (rng.random(dim) < 0.75).astype(np.uint8)
            customs[i] = (images, some)
            gates.append(("Custom", wires[-1], tuple(wires[:-1]), 0.0, 0))
            continue
        kind = kinds[int(rng.integers(len(kinds)))]
        nw = 3 if kind == "Toffoli" else (2 if kind in doubles else 1)
        wires = [int(x) for x in rng.permutation(n)[:nw]]
        arg = float(rng.uniform(-2 * np.pi, 2 * np.pi))
        argi = int(rng.integers(-3, 9))
        gates.append((kind, wires[-1], tuple(wires[:-1]), arg, argi))
    return gates, customs


def to_program(gates, customs=None):
    p = Program()
    for i, (kind, pos, ctl, arg, argi) in enumerate(gates):
        if kind == "Custom":
            p.add_custom(pos, ctl, *customs[i])
        else:
            p.add(kind, pos, ctl, arg, argi)
    return p
