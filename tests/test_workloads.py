"""Workload generators: gate counts of the benchmark circuits, and the inverse program used by
the capacity round trip (circuit then inverse = identity), through the oracle."""
import numpy as np

import helpers as H
from quantr_amd import workloads as W


def test_brickwork_gate_counts():
    # 30 qubits: 40 x 30 singles + 20 x 15 + 20 x 14 CNots + 8 Toffolis
    assert len(W.brickwork(30, depth=40, seed=1234)) == 1788
    assert len(W.brickwork(28, depth=40, seed=1234)) == 1668


def test_hea_parameter_count():
    gates, thetas = W.hea(20, 10, seed=2020)
    assert len(thetas) == 400 and len(gates) == 600


def test_capacity_circuit_round_trip():
    for n in (5, 8, 9):
        gates, cust = W.capacity_circuit(n, depth=8)
        assert sum(1 for g in gates if g[0] == "Custom") == 1
        images, _ = next(iter(cust.values()))
        u = images.T
        assert np.max(np.abs(u.conj().T @ u - np.eye(8))) < 1e-12
        inv, icust = W.inverse(gates, cust)
        psi = H.oracle_simulate(n, W.to_program(gates, cust))
        back = H.oracle_simulate(n, W.to_program(inv, icust), register=psi)
        e0 = np.zeros(1 << n, dtype=np.complex128)
        e0[0] = 1.0
        assert np.max(np.abs(back - e0)) < 1e-12


def test_inverse_of_every_invertible_kind():
    rng = np.random.default_rng(9)
    for trial in range(5):
        n = int(rng.integers(3, 7))
        kinds = ["H", "X", "Y", "Z", "S", "Sdag", "T", "Tdag", "Rx", "Ry", "Rz", "X90", "Y90", "MX90", "MY90",
                 "Phase", "CR", "CZ", "CY", "CNot", "Swap", "Toffoli"]
        gates, _ = H.random_gates(n, 40, rng, kinds=kinds)
        inv, _ = W.inverse(gates)
        psi = H.oracle_simulate(n, H.to_program(gates))
        back = H.oracle_simulate(n, H.to_program(inv), register=psi)
        assert abs(back[0] - 1.0) < 1e-12, trial
