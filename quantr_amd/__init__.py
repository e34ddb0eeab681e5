"""quantr_amd — MI355X (gfx950) state-vector engine behind quantr's Circuit/Gate API.

Layers: include/quantr_hip.h (C ABI, libquantr_hip.so), include/quantr.hpp (C++ mirror of
quantr's Rust API, exposed to Python as quantr_amd.quantr), quantr_amd._capi (ctypes binding
of the C ABI).
"""
from . import _capi  # noqa: F401

__all__ = ["_capi"]
