"""The C ABI library: exports, validation errors and behaviour without a GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import helpers as H
from quantr_amd import _capi

HEADER = os.path.join(H.ROOT, "include", "quantr_hip.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qr_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = _capi.lib()
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(_capi.EXPORTED) == syms


def test_version():
    lib = _capi.lib()
    assert lib.qr_abi_version() == _capi.ABI_VERSION == 4
    assert b"gfx950" in lib.qr_version()


def validate(n, prog):
    p = prog.c_struct()
    st = _capi.lib().qr_validate_program(n, C.byref(p))
    return st, _capi.lib().qr_last_error().decode()


def test_validation_messages():
    p = _capi.Program().add("X", 3)
    st, msg = validate(3, p)
    assert st == _capi.QR_ERR_ARG
    assert msg == "The position, 3, is out of bounds for the circuit with 3 qubits."
    st, msg = validate(3, _capi.Program().add("CNot", 0, [0]))
    assert st == _capi.QR_ERR_ARG
    assert msg == "The gate, CNot(0), has a control node that equals the gate's position 0."
    st, msg = validate(3, _capi.Program().add("CNot", 1, [3]))
    assert msg == "The control node at position 3, is greater than the umnber of qubits 3."
    st, msg = validate(3, _capi.Program().add("Toffoli", 2, [0, 0]))
    assert msg == "The gate, Toffoli(0, 0), has overlapping control nodes."
    st, msg = validate(3, _capi.Program().add("CNot", 1, []))
    assert st == _capi.QR_ERR_ARG and "control" in msg
    st, msg = validate(0, _capi.Program())
    assert msg == "The initialised circuit must have at least one wire."
    assert validate(3, _capi.Program().add("H", 0).add("Toffoli", 2, [0, 1]))[0] == _capi.QR_OK


def test_validation_custom_tables():
    p = _capi.Program().add_custom(1, [0], np.eye(4), [1, 1, 1, 1])
    assert validate(2, p)[0] == _capi.QR_OK
    ops, ctl, img, some = p.arrays()
    bad = _capi.program_struct(ops, ctl, img[:-2], some)
    assert _capi.lib().qr_validate_program(2, C.byref(bad)) == _capi.QR_ERR_ARG
    assert "image table" in _capi.lib().qr_last_error().decode()
    bad = _capi.program_struct(ops, ctl, img, some[:-1])
    assert _capi.lib().qr_validate_program(2, C.byref(bad)) == _capi.QR_ERR_ARG


def test_identity_entries_are_skipped():
    p = _capi.Program().add("Id", 0).add("H", 1)
    assert validate(2, p)[0] == _capi.QR_OK


def test_no_device_is_reported():
    if _capi.device_count() > 0:
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    st = _capi.lib().qr_state_create(4, _capi.QR_C128, C.byref(h))
    assert st == _capi.QR_ERR_NO_DEVICE
    assert not h.value


def test_null_arguments():
    lib = _capi.lib()
    assert lib.qr_state_create(3, 0, None) == _capi.QR_ERR_ARG
    assert lib.qr_apply_program(None, None) == _capi.QR_ERR_ARG
    assert lib.qr_sample(None, None, 0, None) == _capi.QR_ERR_ARG
