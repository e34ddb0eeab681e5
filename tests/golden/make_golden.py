"""Writes tests/golden/reference_vectors.json: the reference's own golden vectors as data.

Every case below transcribes one test of quantr 0.6.0 — the circuit it builds (as builder
calls) and the state vector or bin condition it asserts — with its file:line. Amplitudes are
the hand-computed values of those tests. The Id-free gate program of each case is derived
with the column layout rule of Circuit::add_gates / push_multi_gates (circuit.rs:210-270),
which the layout cases (circuit.rs:535-575) pin.

Run: python tests/golden/make_golden.py
"""
import json
import math
import os

S = 1.0 / math.sqrt(2.0)  # std::f64::consts::FRAC_1_SQRT_2
PI = math.pi
HERE = os.path.dirname(os.path.abspath(__file__))

SINGLE = {"Id", "H", "X", "Y", "Z", "S", "Sdag", "T", "Tdag", "Rx", "Ry", "Rz", "X90", "Y90",
          "MX90", "MY90", "Phase"}


def g(kind, controls=(), arg=0.0, arg_int=0, table=None, name=None):
    d = {"kind": kind, "controls": list(controls), "arg": arg, "arg_int": arg_int}
    if table:
        d["table"] = table
        d["name"] = name or "X"
    return d


ID = g("Id")


def push_multi_gates(col):
    """circuit.rs:230-270: with two or more non-Id gates in a column, every multi-wire gate
    moves to its own new column (in wire order) after the single-wire gates."""
    non_id = [x for x in col if x["kind"] != "Id"]
    if len(non_id) < 2:
        return [col]
    extra = []
    out = [dict(x) for x in col]
    for pos, x in enumerate(col):
        if x["kind"] not in SINGLE:
            c = [ID] * len(col)
            c[pos] = x
            extra.append(c)
            out[pos] = ID
    return [out] + extra


def flatten(n, calls):
    """Builder calls -> (layout with Ids, Id-free program of (gate, position))."""
    cols = []
    for call in calls:
        op = call[0]
        if op == "add_gate":
            col = [ID] * n
            col[call[2]] = call[1]
        elif op == "add_gates":
            col = list(call[1])
        elif op == "add_gates_with_positions":
            col = [ID] * n
            for pos, gate in call[1]:
                col[pos] = gate
        elif op == "add_repeating_gate":
            col = [ID] * n
            for pos in call[2]:
                col[pos] = call[1]
        else:
            raise ValueError(op)
        cols.extend(push_multi_gates(col))
    layout = [x for c in cols for x in c]
    prog = [dict(x, pos=i % n) for i, x in enumerate(layout) if x["kind"] != "Id"]
    return layout, prog


def c(re=0.0, im=0.0):
    return [re, im]


def vec(n, entries):
    v = [c() for _ in range(1 << n)]
    for i, val in entries.items():
        v[i] = val
    return v


CASES = []


def case(name, source, n, calls, expected=None, tol=1e-6, register=None, bins=None, shots=None):
    layout, prog = flatten(n, calls)
    d = {"name": name, "source": source, "n": n, "calls": calls, "program": prog, "tol": tol}
    if expected is not None:
        d["expected"] = expected
    if register is not None:
        d["register"] = register
    if bins is not None:
        d["bins"] = bins
        d["shots"] = shots
    CASES.append(d)


H, X, Y, Z = g("H"), g("X"), g("Y"), g("Z")

case("swap_and_conjugate_gates", "src/circuit.rs:603-613", 2,
     [["add_gates", [H, H]], ["add_gates", [g("S"), g("Sdag")]]],
     [c(0.5), c(0, -0.5), c(0, 0.5), c(0.5)])
case("t_and_conjugate_gates", "src/circuit.rs:615-625", 2,
     [["add_gates", [H, H]], ["add_gates", [g("T"), g("Tdag")]]],
     [c(0.5), c(0.5 * S, -0.5 * S), c(0.5 * S, 0.5 * S), c(0.5)])
case("custom_gates", "src/circuit.rs:628-641 (example_cnot :504-512)", 3,
     [["add_gate", H, 2], ["add_gate", g("Custom", [2], table="example_cnot", name="cNot"), 1]],
     vec(3, {0: c(S), 3: c(S)}))
case("toffoli_gates", "src/circuit.rs:643-658", 4,
     [["add_gate", X, 0], ["add_gate", H, 3], ["add_gate", Y, 3], ["add_gate", g("Toffoli", [3, 0]), 1]],
     vec(4, {8: c(0, -S), 13: c(0, S)}))
case("runs_three_pauli_gates_with_hadamard", "src/circuit.rs:688-701", 4,
     [["add_gates", [Z, Y, H, X]]], vec(4, {5: c(0, S), 7: c(0, S)}))
case("hash_map_with_two_gates", "src/circuit.rs:703-713", 3,
     [["add_gates_with_positions", [[0, X], [2, H]]]], vec(3, {4: c(S), 5: c(S)}))
case("two_hadamard_gates_work", "src/circuit.rs:722-731", 2,
     [["add_gates", [H, H]]], [c(0.5), c(0.5), c(0.5), c(0.5)])
case("add_two_rows_single_gates", "src/circuit.rs:733-748", 4,
     [["add_gates_with_positions", [[0, X]]], ["add_gates_with_positions", [[3, X], [2, H]]]],
     vec(4, {9: c(S), 11: c(S)}))
case("cy_and_swap_gates_work", "src/circuit.rs:750-768", 4,
     [["add_repeating_gate", X, [1, 2]], ["add_gate", g("CY", [2]), 0], ["add_gate", g("Swap", [3]), 2],
      ["add_gate", g("CY", [0]), 3]], vec(4, {12: c(1.0)}))
case("cz_and_swap_gates_work", "src/circuit.rs:770-786", 3,
     [["add_repeating_gate", X, [0, 2]], ["add_gate", g("Swap", [1]), 2], ["add_gate", g("CZ", [1]), 0]],
     vec(3, {6: c(-1.0)}))
case("cnot_gate_simply_use_works", "src/circuit.rs:788-802", 2,
     [["add_gate", H, 0], ["add_gate", g("CNot", [1]), 0]], [c(S), c(), c(S), c()])
case("cnot_gate_simply_flipped", "src/circuit.rs:804-818", 2,
     [["add_gate", H, 0], ["add_gate", g("CNot", [0]), 1]], [c(S), c(), c(), c(S)])
case("cnot_gate_extended_control_works_asymmetric", "src/circuit.rs:820-837", 4,
     [["add_gate", H, 1], ["add_gate", g("CNot", [1]), 3], ["add_gate", Y, 1]],
     vec(4, {1: c(0, -S), 4: c(0, S)}))
case("rx_gate", "src/circuit.rs:847-860", 2,
     [["add_gates", [H, H]], ["add_gate", g("Rx", arg=PI), 0]], [c(0, -0.5)] * 4)
case("ry_gate", "src/circuit.rs:862-875", 2,
     [["add_gates", [H, H]], ["add_gate", g("Ry", arg=PI), 0]], [c(-0.5), c(-0.5), c(0.5), c(0.5)])
case("rz_gate", "src/circuit.rs:877-890", 2,
     [["add_gates", [H, H]], ["add_gate", g("Rz", arg=PI), 0]], [c(0, -0.5), c(0, -0.5), c(0, 0.5), c(0, 0.5)])
case("global_gate", "src/circuit.rs:892-905", 2,
     [["add_gates", [H, H]], ["add_gate", g("Phase", arg=PI), 0]], [c(0, 0.5)] * 4)
case("x90_and_mx90_gate", "src/circuit.rs:907-921", 2,
     [["add_gates", [H, H]], ["add_gate", g("MX90"), 0], ["add_gate", g("X90"), 1]], [c(0.5)] * 4)
case("y90_and_my90_gate", "src/circuit.rs:923-937", 2,
     [["add_gates", [H, H]], ["add_gate", g("MY90"), 0], ["add_gate", g("Y90"), 1]],
     [c(-0.5), c(0.5), c(0.5), c(-0.5)])
case("cr_gate", "src/circuit.rs:939-952", 3,
     [["add_gates", [X, X, X]], ["add_gate", g("CR", [2], arg=-PI * 0.5), 1]], vec(3, {7: c(0, -1.0)}))
case("crk_gate", "src/circuit.rs:954-967", 3,
     [["add_gates", [X, X, X]], ["add_gate", g("CRk", [2], arg_int=2), 1]], vec(3, {7: c(0, 1.0)}))
case("custom_register", "src/circuit.rs:969-982", 3,
     [["add_gate", X, 1]], vec(3, {7: c(1.0)}), register=vec(3, {5: c(1.0)}))
case("simple_qft", "tests/qft.rs:21-47 (qft :51-67)", 3,
     [["add_repeating_gate", X, [1, 2]], ["add_gate", g("Custom", [0, 1], table="qft3", name="QFT"), 2]],
     [c(S * 0.5), c(-S * 0.5), c(0, -S * 0.5), c(0, S * 0.5), c(-0.25, 0.25), c(0.25, -0.25),
      c(0.25, 0.25), c(-0.25, -0.25)], tol=1e-8)
case("grovers_3qubit", "tests/grovers.rs:21-72", 3,
     [["add_repeating_gate", H, [0, 1, 2]], ["add_gate", g("CZ", [1]), 2],
      ["add_repeating_gate", H, [0, 1, 2]], ["add_repeating_gate", X, [0, 1, 2]], ["add_gate", H, 2],
      ["add_gate", g("Toffoli", [0, 1]), 2], ["add_gate", H, 2], ["add_repeating_gate", X, [0, 1, 2]],
      ["add_repeating_gate", H, [0, 1, 2]]],
     [c(), c(), c(), c(-S), c(), c(), c(), c(-S)], tol=1e-8,
     bins={"allowed": ["011", "111"], "min_each": 201}, shots=500)

# tests/grovers.rs:75-155 (x3sudoko): 10 qubits, multi-controlled custom NOTs.
sud = [["add_repeating_gate", H, [0, 1, 2, 3, 4, 5]], ["add_gate", X, 8], ["add_gate", X, 9], ["add_gate", H, 9]]
oracle_part = []
for i in range(3):
    oracle_part.append(["add_gate", g("Toffoli", [i, i + 3]), 8])
oracle_part.append(["add_gate", g("Custom", [0, 1, 2], table="multicnot4"), 6])
for i in range(3):
    oracle_part.append(["add_gate", g("CNot", [i]), 6])
oracle_part.append(["add_gate", g("Custom", [3, 4, 5], table="multicnot4"), 7])
for i in range(3, 6):
    oracle_part.append(["add_gate", g("CNot", [i]), 7])
sud += oracle_part
sud.append(["add_gate", g("Custom", [6, 7, 8], table="multicnot4"), 9])
sud += oracle_part
sud += [["add_repeating_gate", H, [0, 1, 2, 3, 4, 5]], ["add_repeating_gate", X, [0, 1, 2, 3, 4, 5]],
        ["add_gate", H, 5], ["add_gate", g("Custom", [0, 1, 2, 3, 4], table="multicnot6"), 5],
        ["add_gate", H, 5], ["add_repeating_gate", X, [0, 1, 2, 3, 4, 5]],
        ["add_repeating_gate", H, [0, 1, 2, 3, 4, 5]]]
case("x3sudoko", "tests/grovers.rs:74-155 (multicnot :157-172)", 10, sud,
     bins={"prefix_len": 6, "solutions": ["001100", "001010", "010100", "010001", "100010", "100001"],
           "solution_min": 151, "other_max": 149}, shots=5000)


def table_example_cnot():
    """circuit.rs:504-512: |00>,|01> -> None; |10> -> |11>; |11> -> |10>."""
    some = [0, 0, 1, 1]
    images = [[c() for _ in range(4)] for _ in range(4)]
    images[2][3] = c(1.0)
    images[3][2] = c(1.0)
    return {"k": 2, "some": some, "images": images}


def table_multicnot(nc):
    """tests/grovers.rs:157-172: all ones -> last qubit cleared; all ones but the last ->
    last qubit set; every other input -> None."""
    dim = 1 << nc
    some = [0] * dim
    images = {}
    some[dim - 1] = 1
    some[dim - 2] = 1
    images[dim - 1] = dim - 2
    images[dim - 2] = dim - 1
    return {"k": nc, "some": some, "sparse": [[j, r, 1.0, 0.0] for j, r in images.items()]}


TABLES = {
    "example_cnot": table_example_cnot(),
    "multicnot4": table_multicnot(4),
    "multicnot6": table_multicnot(6),
    # tests/qft.rs:51-67: the image of |j> is a nested 3-qubit circuit (H and CRk ladder,
    # no swaps) simulated on the register |j>; tests build the table by simulating it.
    "qft3": {"k": 3, "circuit": [["add_gate", H, 0], ["add_gate", g("CRk", [1], arg_int=2), 0],
                                  ["add_gate", g("CRk", [2], arg_int=3), 0], ["add_gate", H, 1],
                                  ["add_gate", g("CRk", [2], arg_int=2), 1], ["add_gate", H, 2]]},
}
for t in TABLES.values():
    if "circuit" in t:
        t["program"] = flatten(t["k"], t["circuit"])[1]

LAYOUTS = [
    {"name": "pushes_multi_gates", "source": "src/circuit.rs:534-550", "n": 3,
     "calls": [["add_gates", [g("CNot", [2]), g("CNot", [0]), H]],
               ["add_gates", [g("Toffoli", [1, 2]), H, g("CNot", [0])]]],
     "layout": [ID, ID, H, g("CNot", [2]), ID, ID, ID, g("CNot", [0]), ID, ID, H, ID,
                g("Toffoli", [1, 2]), ID, ID, ID, ID, g("CNot", [0])]},
    {"name": "pushes_multi_gates_using_vec", "source": "src/circuit.rs:552-575", "n": 3,
     "calls": [["add_gates_with_positions", [[2, H], [0, g("CNot", [2])], [1, g("CNot", [0])]]],
               ["add_gates_with_positions", [[2, g("CNot", [0])], [0, g("Toffoli", [1, 2])], [1, H]]]],
     "layout": [ID, ID, H, g("CNot", [2]), ID, ID, ID, g("CNot", [0]), ID, ID, H, ID,
                g("Toffoli", [1, 2]), ID, ID, ID, ID, g("CNot", [0])]},
    {"name": "get_gates_doc", "source": "src/circuit.rs:94-101", "n": 3,
     "calls": [["add_gate", X, 2]], "layout": [ID, ID, X]},
]
for lay in LAYOUTS:
    got = flatten(lay["n"], lay["calls"])[0]
    assert got == lay["layout"], lay["name"]

# Builder validation errors (the reference's should_panic tests) and their messages.
ERRORS = [
    {"name": "catches_overlapping_nodes_custom_gate", "source": "src/circuit.rs:514-523", "n": 3,
     "calls": [["add_gates", [ID, g("Custom", [1], table="example_cnot"), ID]]],
     "message_prefix": "The gate, Custom("},
    {"name": "catches_overlapping_control_nodes", "source": "src/circuit.rs:525-532", "n": 3,
     "calls": [["add_gates", [g("CNot", [0]), ID, ID]]],
     "message": "The gate, CNot(0), has a control node that equals the gate's position 0."},
    {"name": "catches_overlapping_control_nodes_using_vec", "source": "src/circuit.rs:577-586", "n": 3,
     "calls": [["add_gates_with_positions", [[2, H], [0, g("CNot", [0])], [1, g("CNot", [0])]]]],
     "message": "The gate, CNot(0), has a control node that equals the gate's position 0."},
    {"name": "control_node_greater_than_circuit_size", "source": "src/circuit.rs:588-597", "n": 3,
     "calls": [["add_gates_with_positions", [[2, H], [0, g("CNot", [2])], [1, g("CNot", [3])]]]],
     "message": "The control node at position 3, is greater than the umnber of qubits 3."},
    {"name": "catches_repeating_positions", "source": "src/circuit.rs:715-720", "n": 4,
     "calls": [["add_repeating_gate", X, [0, 1, 1, 3]]],
     "message": "Attempted to add more than one gate onto a single wire. The positions in [0, 1, 1, 3] must all differ."},
    {"name": "custom_non_ascii_name", "source": "src/circuit.rs:839-845", "n": 3,
     "calls": [["add_gate", g("Custom", [0], table="example_cnot", name="NonAscii†"), 1]],
     "message_prefix": "The custom function name, NonAscii†, does not only use ASCII chars."},
    {"name": "custom_register_wrong_dimension", "source": "src/circuit.rs:984-991", "n": 3,
     "calls": [["add_gate", X, 1]], "register_dim": 2,
     "message": "The custom register has a product state dimension of 2, while the number of qubits is 3. These must equal each other."},
    {"name": "new_zero_wires", "source": "src/circuit.rs:48-53", "n": 0, "calls": [],
     "message": "The initialised circuit must have at least one wire."},
    {"name": "position_out_of_bounds", "source": "src/circuit.rs:157-166", "n": 3,
     "calls": [["add_gate", X, 3]],
     "message": "The position, 3, is out of bounds for the circuit with 3 qubits."},
    {"name": "add_gates_wrong_length", "source": "src/circuit.rs:210-216", "n": 3,
     "calls": [["add_gates", [X, X]]],
     "message": "The number of gates, 2, does not match the number of wires, 3. All wires must have gates added."},
]

# Index <-> basis-state convention (product_states.rs:274-321): wire 0 is the MSB.
INDEX = [{"bits": "110", "index": 6, "source": "src/circuit/states/product_states.rs:274-280"},
         {"bits": "101", "index": 5, "source": "src/circuit/states/product_states.rs:302-321"},
         {"bits": "111", "index": 7, "source": "src/circuit/states/product_states.rs:302-321"},
         {"bits": "10", "index": 2, "source": "src/circuit/states/product_states.rs:302-321"},
         {"bits": "1011", "index": 11, "source": "src/circuit/states/product_states.rs:302-321"}]

if __name__ == "__main__":
    out = {"generator": "tests/golden/make_golden.py", "reference": "quantr 0.6.0",
           "cases": CASES, "tables": TABLES, "layouts": LAYOUTS, "errors": ERRORS, "index": INDEX}
    with open(os.path.join(HERE, "reference_vectors.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(CASES), "cases,", len(LAYOUTS), "layouts,", len(ERRORS), "error cases")
