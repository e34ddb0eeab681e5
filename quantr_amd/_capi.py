"""ctypes binding of the C ABI declared in include/quantr_hip.h.

The library is loaded from quantr_amd/lib/libquantr_hip.so (built in-tree). There is no
fallback: if the library is missing or no GPU is visible, calls fail loudly.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libquantr_hip.so")

QR_OK, QR_ERR_ARG, QR_ERR_OOM, QR_ERR_HIP, QR_ERR_UNSUPPORTED, QR_ERR_NO_DEVICE = range(6)
QR_C128, QR_C64 = 0, 1

# qr_gate_kind (one per quantr::Gate variant, gate.rs:19-106)
GATE_KINDS = ["Id", "H", "X", "Y", "Z", "S", "Sdag", "T", "Tdag", "Rx", "Ry", "Rz", "X90", "Y90",
              "MX90", "MY90", "Phase", "CR", "CRk", "CZ", "CY", "CNot", "Swap", "Toffoli", "Custom"]
KIND = {name: i for i, name in enumerate(GATE_KINDS)}

OP_DTYPE = np.dtype([
    ("kind", "<u4"), ("position", "<u4"), ("ncontrols", "<u4"), ("reserved0", "<u4"),
    ("controls_offset", "<u8"), ("arg", "<f8"), ("arg_int", "<i4"), ("reserved1", "<u4"),
    ("image_offset", "<u8"), ("some_offset", "<u8"),
])
assert OP_DTYPE.itemsize == 56

TERM_DTYPE = np.dtype([("coeff", "<f8"), ("nfactors", "<u4"), ("reserved", "<u4"), ("offset", "<u8")])
assert TERM_DTYPE.itemsize == 24


class QrProgram(C.Structure):
    _fields_ = [("ops", C.c_void_p), ("nops", C.c_uint64), ("controls", C.c_void_p),
                ("ncontrols", C.c_uint64), ("images", C.c_void_p), ("nimages", C.c_uint64),
                ("some", C.c_void_p), ("nsome", C.c_uint64)]


class QrObservable(C.Structure):
    _fields_ = [("terms", C.c_void_p), ("nterms", C.c_uint64), ("wires", C.c_void_p),
                ("paulis", C.c_void_p), ("nfactors", C.c_uint64)]


class QrOptions(C.Structure):
    _fields_ = [("fused", C.c_uint32), ("tile_bits", C.c_uint32), ("low_bits", C.c_uint32),
                ("exact_sample", C.c_uint32), ("tile_variant", C.c_uint32), ("reg_bits", C.c_uint32),
                ("jit", C.c_uint32), ("stage_ops", C.c_uint32), ("tune", C.c_uint32)]


class QrBatchOptions(C.Structure):
    _fields_ = [("autotune", C.c_uint32), ("fwd_tile_bits", C.c_uint32), ("dual_tile_bits", C.c_uint32),
                ("streams", C.c_uint32)]


class QrPlanStats(C.Structure):
    _fields_ = [("gates", C.c_uint64), ("sweeps", C.c_uint64), ("tile_sweeps", C.c_uint64),
                ("bytes", C.c_double), ("max_stage_gates", C.c_uint64), ("restore_sweeps", C.c_uint64),
                ("tuning", C.c_uint64), ("xcd_order_sweeps", C.c_uint64)]


class QrProfileEntry(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("total_ms", C.c_double),
                ("bytes", C.c_double)]


# (name, restype, argtypes) of every exported entry point; tests check the export table.
_SIGNATURES = [
    ("qr_abi_version", C.c_uint32, []),
    ("qr_version", C.c_char_p, []),
    ("qr_last_error", C.c_char_p, []),
    ("qr_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("qr_state_create", C.c_int, [C.c_uint32, C.c_int, C.POINTER(C.c_void_p)]),
    ("qr_state_destroy", C.c_int, [C.c_void_p]),
    ("qr_state_num_qubits", C.c_int, [C.c_void_p, C.POINTER(C.c_uint32)]),
    ("qr_state_precision", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("qr_state_reset", C.c_int, [C.c_void_p]),
    ("qr_state_set_amplitudes", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("qr_state_get_amplitudes", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    ("qr_state_get_amplitudes_range", C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p]),
    ("qr_state_set_options", C.c_int, [C.c_void_p, C.POINTER(QrOptions)]),
    ("qr_state_get_options", C.c_int, [C.c_void_p, C.POINTER(QrOptions)]),
    ("qr_state_synchronize", C.c_int, [C.c_void_p]),
    ("qr_validate_program", C.c_int, [C.c_uint32, C.POINTER(QrProgram)]),
    ("qr_apply_program", C.c_int, [C.c_void_p, C.POINTER(QrProgram)]),
    ("qr_sample", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    ("qr_sample_programs", C.c_int, [C.c_uint32, C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                                     C.c_void_p]),
    ("qr_state_norm_sqr", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("qr_expectation", C.c_int, [C.c_void_p, C.POINTER(QrObservable), C.POINTER(C.c_double)]),
    ("qr_expectation_noalloc", C.c_int, [C.c_void_p, C.POINTER(QrObservable), C.POINTER(C.c_double)]),
    ("qr_grad_create", C.c_int, [C.c_uint32, C.c_int, C.POINTER(C.c_void_p)]),
    ("qr_grad_destroy", C.c_int, [C.c_void_p]),
    ("qr_grad_forward", C.c_int, [C.c_void_p, C.POINTER(QrProgram), C.POINTER(QrObservable),
                                  C.POINTER(C.c_double)]),
    ("qr_grad_num_params", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    ("qr_grad_backward", C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_uint64]),
    ("qr_grad_set_options", C.c_int, [C.c_void_p, C.c_void_p]),
    ("qr_batch_create", C.c_int, [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                  C.c_void_p, C.POINTER(C.c_void_p)]),
    ("qr_batch_destroy", C.c_int, [C.c_void_p]),
    ("qr_batch_param_counts", C.c_int, [C.c_void_p, C.c_void_p]),
    ("qr_batch_run", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("qr_batch_grad", C.c_int, [C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("qr_profile_enable", C.c_int, [C.c_void_p, C.c_int]),
    ("qr_profile_read", C.c_int, [C.c_void_p, C.POINTER(QrProfileEntry), C.c_uint64,
                                  C.POINTER(C.c_uint64)]),
    ("qr_state_last_plan", C.c_int, [C.c_void_p, C.POINTER(QrPlanStats)]),
    ("qr_measure_stream", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_double)]),
]
EXPORTED = [s[0] for s in _SIGNATURES]

_lib = None
ABI_VERSION = 4  # QR_ABI_VERSION of include/quantr_hip.h; the structures above follow it


class QrError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def lib():
    """Loads libquantr_hip.so (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libquantr_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        l = C.CDLL(LIB_PATH)
        for name, res, args in _SIGNATURES:
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        if l.qr_abi_version() != ABI_VERSION:
            raise RuntimeError("%s has C ABI version %d; this binding expects %d (rebuild with "
                               "__graft_entry__.build())" % (LIB_PATH, l.qr_abi_version(), ABI_VERSION))
        _lib = l
    return _lib


def check(status):
    if status != QR_OK:
        msg = lib().qr_last_error().decode("utf-8", "replace")
        raise QrError(status, msg)


def device_count():
    c = C.c_int(0)
    check(lib().qr_device_count(C.byref(c)))
    return c.value


class Program:
    """A gate program in quantr terms (one entry per non-identity gate, in order)."""

    def __init__(self):
        self.ops = []
        self.controls = []
        self.images = []
        self.some = []

    def add(self, kind, position, controls=(), arg=0.0, arg_int=0):
        k = KIND[kind] if isinstance(kind, str) else int(kind)
        self.ops.append((k, int(position), len(controls), 0, len(self.controls), float(arg), int(arg_int), 0, 0, 0))
        self.controls.extend(int(c) for c in controls)
        return self

    def add_custom(self, position, controls, images, some):
        """images: (2^k, 2^k) complex array, row j = image of local basis state j."""
        k = len(controls) + 1
        dim = 1 << k
        images = np.asarray(images, dtype=np.complex128).reshape(dim, dim)
        some = np.asarray(some, dtype=np.uint8).reshape(dim)
        img_off = len(self.images)
        some_off = len(self.some)
        self.images.extend(np.ascontiguousarray(images).view(np.float64).ravel().tolist())
        self.some.extend(some.tolist())
        self.ops.append((KIND["Custom"], int(position), len(controls), 0, len(self.controls), 0.0, 0, 0,
                         img_off, some_off))
        self.controls.extend(int(c) for c in controls)
        return self

    def arrays(self):
        ops = np.array(self.ops, dtype=OP_DTYPE) if self.ops else np.zeros(0, dtype=OP_DTYPE)
        ctl = np.array(self.controls, dtype=np.uint32)
        img = np.array(self.images, dtype=np.float64)
        some = np.array(self.some, dtype=np.uint8)
        return ops, ctl, img, some

    def c_struct(self):
        ops, ctl, img, some = self.arrays()
        return program_struct(ops, ctl, img, some)

    def __len__(self):
        return len(self.ops)


def program_struct(ops, ctl, img, some):
    """Builds a QrProgram from numpy arrays; the arrays are kept alive on the struct."""
    p = QrProgram(ops.ctypes.data if len(ops) else None, len(ops),
                  ctl.ctypes.data if len(ctl) else None, len(ctl),
                  img.ctypes.data if len(img) else None, len(img),
                  some.ctypes.data if len(some) else None, len(some))
    p._keep = (ops, ctl, img, some)
    return p


class Observable:
    """Pauli sum: list of (coeff, [(wire, 'X'|'Y'|'Z'), ...])."""

    def __init__(self, terms=()):
        self.terms = list(terms)

    def add(self, coeff, factors):
        self.terms.append((float(coeff), list(factors)))
        return self

    def c_struct(self):
        t = np.zeros(len(self.terms), dtype=TERM_DTYPE)
        wires, paulis = [], []
        for i, (c, fac) in enumerate(self.terms):
            t[i]["coeff"] = c
            t[i]["nfactors"] = len(fac)
            t[i]["offset"] = len(wires)
            for w, p in fac:
                wires.append(int(w))
                paulis.append(ord(p))
        w = np.array(wires, dtype=np.uint32)
        p = np.array(paulis, dtype=np.uint8)
        o = QrObservable(t.ctypes.data if len(t) else None, len(t), w.ctypes.data if len(w) else None,
                         p.ctypes.data if len(p) else None, len(w))
        o._keep = (t, w, p)
        return o


def sample_programs(n, programs, draws, uniforms, precision=QR_C128):
    """qr_sample_programs: programs[j] simulated from |0...0> and measured with draws[j]
    consecutive uniforms; returns the indices (UINT64_MAX = None), draws in order."""
    structs = [p.c_struct() if isinstance(p, Program) else p for p in programs]
    arr = (QrProgram * max(1, len(structs)))(*structs)
    d = np.ascontiguousarray(np.asarray(draws, dtype=np.uint64))
    u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
    if int(d.sum()) != u.size:
        raise ValueError("draws must add up to the number of uniforms")
    out = np.empty(u.size, dtype=np.uint64)
    check(lib().qr_sample_programs(n, precision, C.cast(arr, C.c_void_p), len(structs), d.ctypes.data,
                                   u.ctypes.data, out.ctypes.data))
    return out


class State:
    """A device-resident state (qr_state)."""

    def __init__(self, num_qubits, precision=QR_C128):
        self._h = C.c_void_p()
        check(lib().qr_state_create(num_qubits, precision, C.byref(self._h)))
        self.n = num_qubits
        self.precision = precision

    def close(self):
        if self._h:
            lib().qr_state_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_options(self, fused=1, tile_bits=0, low_bits=0, exact_sample=2, tile_variant=0, reg_bits=0, jit=2,
                    stage_ops=0, tune=0):
        o = QrOptions(fused, tile_bits, low_bits, exact_sample, tile_variant, reg_bits, jit, stage_ops, tune)
        check(lib().qr_state_set_options(self._h, C.byref(o)))

    def reset(self):
        check(lib().qr_state_reset(self._h))

    def set_amplitudes(self, amps):
        a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex128))
        check(lib().qr_state_set_amplitudes(self._h, a.ctypes.data, a.size))

    def amplitudes(self):
        out = np.empty(1 << self.n, dtype=np.complex128)
        check(lib().qr_state_get_amplitudes(self._h, out.ctypes.data, out.size))
        return out

    def amplitudes_range(self, first, count):
        out = np.empty(count, dtype=np.complex128)
        check(lib().qr_state_get_amplitudes_range(self._h, first, count, out.ctypes.data))
        return out

    def apply(self, program):
        p = program.c_struct() if isinstance(program, Program) else program
        check(lib().qr_apply_program(self._h, C.byref(p)))

    def synchronize(self):
        check(lib().qr_state_synchronize(self._h))

    def sample(self, uniforms):
        u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
        out = np.empty(u.size, dtype=np.uint64)
        check(lib().qr_sample(self._h, u.ctypes.data, u.size, out.ctypes.data))
        return out

    def norm_sqr(self):
        v = C.c_double()
        check(lib().qr_state_norm_sqr(self._h, C.byref(v)))
        return v.value

    def expectation(self, observable):
        o = observable.c_struct()
        v = C.c_double()
        check(lib().qr_expectation(self._h, C.byref(o), C.byref(v)))
        return v.value

    def last_plan(self):
        s = QrPlanStats()
        check(lib().qr_state_last_plan(self._h, C.byref(s)))
        return {"gates": s.gates, "sweeps": s.sweeps, "tile_sweeps": s.tile_sweeps, "bytes": s.bytes,
                "max_stage_gates": s.max_stage_gates, "restore_sweeps": s.restore_sweeps, "tuning": s.tuning,
                "xcd_order_sweeps": s.xcd_order_sweeps}

    def measure_stream(self, reps=5):
        """Average ms of an in-place read+write sweep of this state (the streaming ceiling)."""
        v = C.c_double()
        check(lib().qr_measure_stream(self._h, reps, C.byref(v)))
        return v.value

    def profile_enable(self, on=True):
        check(lib().qr_profile_enable(self._h, 1 if on else 0))

    def profile_read(self):
        cnt = C.c_uint64()
        check(lib().qr_profile_read(self._h, None, 0, C.byref(cnt)))
        arr = (QrProfileEntry * max(cnt.value, 1))()
        check(lib().qr_profile_read(self._h, arr, cnt.value, C.byref(cnt)))
        return {arr[i].name.decode(): {"launches": arr[i].launches, "total_ms": arr[i].total_ms,
                                       "bytes": arr[i].bytes} for i in range(cnt.value)}


class Grad:
    """Adjoint differentiation context (qr_grad)."""

    def __init__(self, num_qubits, precision=QR_C128):
        self._h = C.c_void_p()
        check(lib().qr_grad_create(num_qubits, precision, C.byref(self._h)))
        self.n = num_qubits

    def close(self):
        if self._h:
            lib().qr_grad_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, program, observable):
        p = program.c_struct() if isinstance(program, Program) else program
        o = observable.c_struct()
        v = C.c_double()
        check(lib().qr_grad_forward(self._h, C.byref(p), C.byref(o), C.byref(v)))
        return v.value

    def set_options(self, fused=1, tile_bits=0, low_bits=0, exact_sample=2, tile_variant=0, reg_bits=0, jit=2,
                    stage_ops=0, tune=0):
        o = QrOptions(fused, tile_bits, low_bits, exact_sample, tile_variant, reg_bits, jit, stage_ops, tune)
        check(lib().qr_grad_set_options(self._h, C.byref(o)))

    def num_params(self):
        v = C.c_uint64()
        check(lib().qr_grad_num_params(self._h, C.byref(v)))
        return v.value

    def backward(self, grad_out=1.0):
        n = self.num_params()
        out = np.zeros(n, dtype=np.float64)
        check(lib().qr_grad_backward(self._h, float(grad_out), out.ctypes.data if n else None, n))
        return out


def batch_grad(instances, precision=QR_C64):
    """instances: list of (num_qubits, Program, Observable). Returns (expvals, grads list, latency_us)."""
    m = len(instances)
    nq = np.array([i[0] for i in instances], dtype=np.uint32)
    progs = (QrProgram * m)()
    obs = (QrObservable * m)()
    keep = []
    counts = []
    for i, (n, p, o) in enumerate(instances):
        ps, os_ = p.c_struct(), o.c_struct()
        keep.append((ps, os_))
        progs[i] = ps
        obs[i] = os_
        ops = p.arrays()[0]
        counts.append(int(np.isin(ops["kind"], [KIND["Rx"], KIND["Ry"], KIND["Rz"], KIND["Phase"]]).sum()))
    offs = np.zeros(m, dtype=np.uint64)
    if m:
        offs[1:] = np.cumsum(counts)[:-1]
    grads = np.zeros(int(sum(counts)), dtype=np.float64)
    ev = np.zeros(m, dtype=np.float64)
    lat = np.zeros(m, dtype=np.float64)
    check(lib().qr_batch_grad(m, nq.ctypes.data, C.addressof(progs), C.addressof(obs), precision,
                              ev.ctypes.data, grads.ctypes.data if grads.size else None, offs.ctypes.data,
                              lat.ctypes.data))
    out = [grads[int(offs[i]):int(offs[i]) + counts[i]] for i in range(m)]
    return ev, out, lat


class Batch:
    """Many (num_qubits, program, observable) instances, each with expectation value and full
    adjoint gradient, executed as one captured graph launch per run (qr_batch)."""

    def __init__(self, instances, precision=QR_C64, autotune=False, fwd_tile_bits=0, dual_tile_bits=0, streams=0):
        self.m = len(instances)
        self.n = [int(i[0]) for i in instances]
        nq = np.array(self.n, dtype=np.uint32)
        obs = (QrObservable * self.m)()
        self._keep_obs = []
        for i, (_, _, o) in enumerate(instances):
            os_ = o.c_struct()
            self._keep_obs.append(os_)
            obs[i] = os_
        progs, keep = self._programs([i[1] for i in instances])
        self._h = C.c_void_p()
        opts = QrBatchOptions(1 if autotune else 0, fwd_tile_bits, dual_tile_bits, streams)
        check(lib().qr_batch_create(self.m, nq.ctypes.data, C.addressof(progs), C.addressof(obs), precision,
                                    C.byref(opts), C.byref(self._h)))
        self.counts = np.zeros(self.m, dtype=np.uint64)
        check(lib().qr_batch_param_counts(self._h, self.counts.ctypes.data))
        self.offsets = np.zeros(self.m, dtype=np.uint64)
        if self.m:
            self.offsets[1:] = np.cumsum(self.counts)[:-1]

    def _programs(self, programs):
        arr = (QrProgram * self.m)()
        keep = []
        for i, p in enumerate(programs):
            ps = p.c_struct() if isinstance(p, Program) else p
            keep.append(ps)
            arr[i] = ps
        return arr, keep

    def run(self, programs=None):
        """Returns (expvals, list of gradient arrays, latency_us per instance)."""
        arr = None
        keep = None
        if programs is not None:
            arr, keep = self._programs(programs)
        grads = np.zeros(int(self.counts.sum()), dtype=np.float64)
        ev = np.zeros(self.m, dtype=np.float64)
        lat = np.zeros(self.m, dtype=np.float64)
        check(lib().qr_batch_run(self._h, C.addressof(arr) if arr is not None else None, ev.ctypes.data,
                                 grads.ctypes.data if grads.size else None, self.offsets.ctypes.data, lat.ctypes.data))
        out = [grads[int(self.offsets[i]):int(self.offsets[i] + self.counts[i])] for i in range(self.m)]
        return ev, out, lat

    def close(self):
        if self._h:
            lib().qr_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
