"""Writes tests/golden/printer_diagrams.json: the two circuit diagrams asserted by the
reference's Printer tests (src/circuit/printer.rs:395-416 and :418-440), as data.

The circuits are described below as builder calls (transcribed from those tests); the expected
strings are read from the asserts of that file when this script is run in a checkout that has
the reference at /root/reference.

Run: python tests/golden/make_printer_golden.py
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/circuit/printer.rs"

CASES = [
    {"name": "producing_string_circuit", "cite": "src/circuit/printer.rs:395-416", "n": 4,
     "calls": [["add_gate", "H", 3], ["add_repeating_gate", "Y", [0, 1]], ["add_gate", "Toffoli(0,3)", 1],
               ["add_gate", "CNot(1)", 3], ["add_gate", "CNot(2)", 0], ["add_gate", "CNot(2)", 1]]},
    {"name": "producing_string_circuit_custom", "cite": "src/circuit/printer.rs:418-440", "n": 4,
     "calls": [["add_gate", "H", 3], ["add_gates", ["H", "Custom(example_cnot,[3],\"Custom CNot\")", "Id", "X"]],
               ["add_repeating_gate", "Y", [0, 1]], ["add_gate", "Toffoli(0,3)", 1], ["add_gate", "CNot(1)", 3],
               ["add_gate", "CNot(2)", 0], ["add_gate", "CNot(2)", 1]]},
]


def main():
    text = open(SRC, encoding="utf-8").read()
    found = re.findall(r'assert_eq!\(circuit_printer\.get_diagram\(\), "((?:[^"\\]|\\.)*)"\.to_string\(\)\);', text)
    assert len(found) == len(CASES), len(found)
    for case, lit in zip(CASES, found):
        case["diagram"] = lit.encode("utf-8").decode("unicode_escape").encode("latin-1").decode("utf-8")
    with open(os.path.join(HERE, "printer_diagrams.json"), "w", encoding="utf-8") as f:
        json.dump({"cases": CASES}, f, ensure_ascii=False, indent=1)


if __name__ == "__main__":
    main()
