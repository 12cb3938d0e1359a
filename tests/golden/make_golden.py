"""Writes tests/golden/statevector_cases.json.

The cases are transcribed from the closed-form statevector targets that the
reference's own tests hold (test/terra/reference/ref_*.py); each entry cites
its source.  Circuits are written as Aer operation lists: [name, qubits,
params] for gates, ["unitary", qubits, row-major matrix] and
["diagonal", qubits, diagonal].  Nothing of the reference is imported or run:
the targets below are the closed forms printed in those files.

Run: python tests/golden/make_golden.py
"""
import json
import math
import pathlib

import numpy as np

s2 = 1 / math.sqrt(2)
Y = [[0, -1j], [1j, 0]]
CX = [[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]]

cases = []


def add(name, src, n, ops, target):
    t = np.asarray(target, dtype=complex).reshape(-1)
    cases.append({"name": name, "source": src, "num_qubits": n, "ops": ops,
                  "target_re": t.real.tolist(), "target_im": t.imag.tolist()})


def mat(m):
    m = np.asarray(m, dtype=complex)
    return {"re": m.real.tolist(), "im": m.imag.tolist()}


# ref_non_clifford.py:71-80 (T, T.X) and :159-170 (T.H, X.T.H, H.T.T.H)
add("t", "ref_non_clifford.py:71-80", 1, [["t", [0], []]], [1, 0])
add("t.x", "ref_non_clifford.py:71-80", 1, [["x", [0], []], ["t", [0], []]], np.array([0, 1 + 1j]) / math.sqrt(2))
add("t.h", "ref_non_clifford.py:159-170", 1, [["h", [0], []], ["t", [0], []]], [s2, 0.5 + 0.5j])
add("x.t.h", "ref_non_clifford.py:159-170", 1, [["h", [0], []], ["t", [0], []], ["x", [0], []]], [0.5 + 0.5j, s2])
add("h.t.t.h", "ref_non_clifford.py:159-170", 1,
    [["h", [0], []], ["t", [0], []], ["t", [0], []], ["h", [0], []]], np.array([1 + 1j, 1 - 1j]) / 2)
# ref_non_clifford.py:581-588 (CCX nondeterministic)
add("ccx(0,1,2)", "ref_non_clifford.py:581-588", 3,
    [["h", [0], []], ["x", [1], []], ["ccx", [0, 1, 2], []], ["x", [1], []]],
    np.array([1, 0, 0, 0, 0, 1, 0, 0]) / math.sqrt(2))
add("ccx(2,1,0)", "ref_non_clifford.py:581-588", 3,
    [["h", [1], []], ["x", [2], []], ["ccx", [2, 1, 0], []], ["x", [2], []]],
    np.array([1, 0, 0, 1, 0, 0, 0, 0]) / math.sqrt(2))
# ref_non_clifford.py:1638-1690 (CU deterministic: first column of each unitary)
pi = math.pi
cu_cases = [
    ("I^X.CI.I^X", [["x", [0], []], ["cu", [0, 1], [0, 0, 0, 0]], ["x", [0], []]], [1, 0, 0, 0]),
    ("CX", [["cu", [0, 1], [pi, 0, pi, 0]]], [1, 0, 0, 0]),
    ("I^X.CX.I^X", [["x", [0], []], ["cu", [0, 1], [pi, 0, pi, 0]], ["x", [0], []]], [0, 0, 1, 0]),
    ("H^X.CH.I^X", [["x", [0], []], ["cu", [0, 1], [pi / 2, 0, pi, 0]], ["x", [0], []], ["h", [1], []]],
     [1, 0, 0, 0]),
    ("I^X.CRX(pi).I^X", [["x", [0], []], ["cu", [0, 1], [pi, -pi / 2, pi / 2, 0]], ["x", [0], []]],
     [0, 0, -1j, 0]),
    ("I^X.CRY(pi).I^X", [["x", [0], []], ["cu", [0, 1], [pi, 0, 0, 0]], ["x", [0], []]], [0, 0, 1, 0]),
]
for nm, ops, t in cu_cases:
    add("cu " + nm, "ref_non_clifford.py:1638-1690", 2, ops, t)
# ref_unitary_gate.py:124-141
uops = [
    [["unitary", [0, 1], mat(CX)]],
    [["unitary", [1, 0], mat(CX)]],
    [["unitary", [1], mat(Y)], ["unitary", [0, 1], mat(CX)]],
    [["unitary", [0], mat(Y)], ["unitary", [1, 0], mat(CX)]],
    [["unitary", [0], mat(Y)], ["unitary", [0, 1], mat(CX)]],
    [["unitary", [1], mat(Y)], ["unitary", [1, 0], mat(CX)]],
]
utargets = [[1, 0, 0, 0], [1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1j], [0, 0, 0, 1j]]
for i, (o, t) in enumerate(zip(uops, utargets)):
    add(f"unitary {i}", "ref_unitary_gate.py:124-141", 2, o, t)
# ref_diagonal_gate.py:97-110
for q in (0, 1):
    add(f"diagonal [1,-1] on {q}", "ref_diagonal_gate.py:97-110", 2,
        [["h", [q], []], ["diagonal", [q], mat([1, -1])], ["h", [q], []]],
        [0, 1, 0, 0] if q == 0 else [0, 0, 1, 0])
add("diagonal [1,-1,-1,1]", "ref_diagonal_gate.py:97-110", 2,
    [["h", [0], []], ["h", [1], []], ["diagonal", [0, 1], mat([1, -1, -1, 1])], ["h", [0], []], ["h", [1], []]],
    [0, 0, 0, 1])
for k in (10, 100, 1000, 10000):
    ph = np.exp(-1j * np.pi / k)
    add(f"diagonal phase pi/{k}", "ref_diagonal_gate.py:97-110", 2,
        [["x", [0], []], ["x", [1], []], ["diagonal", [0, 1], mat([1, 1, 1, ph])]], [0, 0, 0, ph])

out = pathlib.Path(__file__).with_name("statevector_cases.json")
out.write_text(json.dumps(cases, indent=1))
print(f"wrote {len(cases)} cases to {out}")
