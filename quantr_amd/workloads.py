"""Seeded synthetic workloads, expressed as quantr gate programs.

Each generator returns a list of gates in quantr's terms — (kind, position, controls, arg,
arg_int) tuples — in application order, i.e. the Id-free flattening of a Circuit
(simulation.rs:37-56). Random numbers come from SplitMix64 so the same lists can be
regenerated anywhere without numpy's generator versions mattering.
"""
import math

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & MASK64

    def next_u64(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        """53-bit uniform in [0, 1)."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def below(self, n):
        return self.next_u64() % n


def qft10(n=10, seed=0):
    """QFT on |x>: X on wires flipped with p = 1/2, H + CRk ladders (tests/qft.rs:55-62), final
    swaps. Returns (gates, x_mask_wires)."""
    rng = SplitMix64(seed)
    gates = []
    flipped = []
    for w in range(n):
        if rng.next_u64() & 1:
            gates.append(("X", w, (), 0.0, 0))
            flipped.append(w)
    for pos in range(n):
        gates.append(("H", pos, (), 0.0, 0))
        for k in range(2, n - pos + 1):
            gates.append(("CRk", pos, (pos + k - 1,), 0.0, k))
    for i in range(n // 2):
        gates.append(("Swap", i, (n - 1 - i,), 0.0, 0))
    return gates, flipped


SINGLE_KINDS = ["H", "X", "Y", "Z", "S", "T", "Rx", "Ry", "Rz"]


def brickwork(n, depth=40, seed=1234):
    """Random brickwork: per layer one gate per wire from {H,X,Y,Z,S,T,Rx,Ry,Rz} (angles
    U[0, 2pi)), then CNot(control w, target w+1) on even or odd pairs, and every 5th layer a
    Toffoli on 3 distinct seeded wires."""
    rng = SplitMix64(seed)
    gates = []
    for layer in range(depth):
        for w in range(n):
            kind = SINGLE_KINDS[rng.below(len(SINGLE_KINDS))]
            arg = 2.0 * math.pi * rng.uniform() if kind in ("Rx", "Ry", "Rz") else 0.0
            gates.append((kind, w, (), arg, 0))
        for w in range(layer % 2, n - 1, 2):
            gates.append(("CNot", w + 1, (w,), 0.0, 0))
        if (layer + 1) % 5 == 0:
            ws = []
            while len(ws) < 3:
                x = rng.below(n)
                if x not in ws:
                    ws.append(x)
            gates.append(("Toffoli", ws[2], (ws[0], ws[1]), 0.0, 0))
    return gates


def hea(n, layers, seed=2020, angle_lo=-math.pi, angle_hi=math.pi):
    """Hardware-efficient ansatz: per layer Ry(theta) on every wire, Rz(theta) on every wire,
    then a ring of CNot(i -> i+1 mod n). Returns (gates, thetas)."""
    rng = SplitMix64(seed)
    gates = []
    thetas = []
    for _ in range(layers):
        for kind in ("Ry", "Rz"):
            for w in range(n):
                t = angle_lo + (angle_hi - angle_lo) * rng.uniform()
                thetas.append(t)
                gates.append((kind, w, (), t, 0))
        for i in range(n):
            if n > 1:
                gates.append(("CNot", (i + 1) % n, (i,), 0.0, 0))
    return gates, thetas


def hea_observable(n):
    """sum_i Z_i Z_{i+1} (open chain) + 0.5 sum_i X_i."""
    terms = [(1.0, [(i, "Z"), (i + 1, "Z")]) for i in range(n - 1)]
    terms += [(0.5, [(i, "X")]) for i in range(n)]
    return terms


def random_unitary(k, seed):
    """Haar-random 2^k x 2^k unitary: QR of a complex Gaussian matrix with the phases of R's
    diagonal divided out (Mezzadri 2007). Uses numpy's seeded generator."""
    import numpy as np

    d = 1 << k
    g = np.random.default_rng(seed)
    z = (g.standard_normal((d, d)) + 1j * g.standard_normal((d, d))) / math.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph  # columns scaled


def capacity_circuit(n, depth=40, seed=1234, custom_seed=3):
    """The brickwork family with one seeded 3-qubit custom gate after the middle layer: a
    Haar-random 8x8 unitary (QR of a complex Gaussian matrix, Mezzadri 2007) tabulated from its
    product-state map (input j -> column j, always Some) on 3 seeded distinct wires.
    Returns (gates, customs) with customs = {index: (images, some)}, images[j][r] = U[r][j]."""
    import numpy as np

    gates = brickwork(n, depth=depth, seed=seed)
    per_layer = n + n // 2  # singles + CNots of an even layer; only a cut position is needed
    cut = min(len(gates), (depth // 2) * per_layer)
    rng = SplitMix64(custom_seed)
    ws = []
    while len(ws) < 3:
        x = rng.below(n)
        if x not in ws:
            ws.append(x)
    u = random_unitary(3, custom_seed)
    images = np.ascontiguousarray(u.T)
    some = np.ones(8, dtype=np.uint8)
    gates = gates[:cut] + [("Custom", ws[2], (ws[0], ws[1]), 0.0, 0)] + gates[cut:]
    return gates, {cut: (images, some)}


INVERSE_KIND = {"S": "Sdag", "Sdag": "S", "T": "Tdag", "Tdag": "T", "X90": "MX90", "MX90": "X90",
                "Y90": "MY90", "MY90": "Y90"}


def inverse(gates, customs=None):
    """The inverse program: reversed order, each gate replaced by its inverse (angles negated,
    S/T and the 90-degree gates by their partners, custom tables by the conjugate transpose).
    CRk has no inverse among quantr's gates and is rejected."""
    import numpy as np

    out, out_custom = [], {}
    for i in range(len(gates) - 1, -1, -1):
        kind, pos, ctl, arg, argi = gates[i]
        if kind == "CRk":
            raise ValueError("CRk has no inverse gate in quantr")
        if kind == "Custom":
            images, some = customs[i]
            out_custom[len(out)] = (np.ascontiguousarray(np.conj(images).T), some)
            out.append((kind, pos, ctl, 0.0, 0))
        elif kind in ("Rx", "Ry", "Rz", "Phase", "CR"):
            out.append((kind, pos, ctl, -arg, argi))
        else:
            out.append((INVERSE_KIND.get(kind, kind), pos, ctl, arg, argi))
    return out, out_custom


def batch_instances(count, seed=512):
    """The batched-serving workload: `count` HEA instances with n ~ U{10..24} qubits and
    L ~ U{2..12} layers, each with its own angles. Returns [(n, L, gates, thetas), ...]."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = 10 + rng.below(15)
        L = 2 + rng.below(11)
        gates, thetas = hea(n, L, seed=rng.next_u64())
        out.append((n, L, gates, thetas))
    return out


def uniforms(count, seed):
    rng = SplitMix64(seed)
    return [rng.uniform() for _ in range(count)]


def to_program(gates, custom=None):
    """gates -> _capi.Program. custom: dict index -> (images, some) for 'Custom' entries."""
    from ._capi import Program

    p = Program()
    for i, g in enumerate(gates):
        kind, pos, ctl, arg, argi = g
        if kind == "Custom":
            images, some = custom[i]
            p.add_custom(pos, ctl, images, some)
        else:
            p.add(kind, pos, ctl, arg, argi)
    return p
