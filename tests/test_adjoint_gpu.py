"""Saved expectation values (probes in tile passes) and adjoint gradients.

Saved Pauli expectation values must equal the oracle's expval_pauli at the
same point of the gate sequence.  Adjoint energies and gradients are checked
against the oracle-evaluated energy, central finite differences (h = 1e-4)
and the parameter-shift rule, all on the CPU oracle.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from qiskit_aer_amd import circuits
from qiskit_aer_amd.adjoint import AdjointEstimator, adjoint_expectation, bind
from qiskit_aer_amd.circuit import Circuit
from tests import dense as D

pytestmark = pytest.mark.gpu


def tfim(n, h=0.7):
    terms = []
    for i in range(n - 1):
        lab = ["I"] * n
        lab[n - 1 - i] = "Z"
        lab[n - 2 - i] = "Z"
        terms.append(("".join(lab), -1.0))
    for i in range(n):
        lab = ["I"] * n
        lab[n - 1 - i] = "X"
        terms.append(("".join(lab), -h))
    return terms


def ansatz(n, layers, seed):
    """RY and RZ on every qubit then a CZ brickwork, per layer; returns the
    circuit, its param_index and a random parameter row."""
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    idx = []
    p = 0
    for layer in range(layers):
        for q in range(n):
            c.gate("ry", [q], [0.0]); idx.append(p); p += 1
            c.gate("rz", [q], [0.0]); idx.append(p); p += 1
        for a in range(layer % 2, n - 1, 2):
            c.gate("cz", [a, a + 1]); idx.append(-1)
    return c, idx, p, rng.uniform(-math.pi, math.pi, size=p)


def oracle_energy(c, idx, row, ham):
    bc = bind(c, idx, row)
    _, sv, _ = O.run_circuit(bc, shots=0, seed=0, fusion=False)
    v = O.OracleQubitVector(c.num_qubits)
    v.initialize_from_vector(sv)
    q = list(range(c.num_qubits))
    return sum(coef * v.expval_pauli(q, lab) for lab, coef in ham)


def test_saved_expvals_match_oracle():
    from qiskit_aer_amd import QubitVector
    n = 16
    print("case: saved expvals n=16", flush=True)
    rng = np.random.default_rng(3)
    c = circuits.brickwork_u3_cx(n, 6, seed=3, measure=False)
    e = QubitVector(n)
    e.initialize()
    o = O.OracleQubitVector(n)
    o.initialize()
    expected = []
    half = len(c.ops) // 2
    labels = ["".join(rng.choice(list("IXYZ"), 5)) for _ in range(12)]
    qs = [int(x) for x in rng.choice(n, 5, replace=False)]
    for part in (c.ops[:half], c.ops[half:]):
        sub = Circuit(n)
        sub.ops = part
        e.apply_ops(sub)
        for op in part:
            if op.name == "u3":
                o.apply_mcu(list(op.qubits), D.u3(*op.params))
            else:
                o.apply_mcx(list(op.qubits))
        for lab in labels:
            e.save_expval_pauli(qs, lab, 0.5 - 0.25j)
            expected.append(o.expval_pauli(qs, lab, 0.5 - 0.25j))
    got = e.saved_expvals()
    assert np.max(np.abs(got - np.array(expected))) < 1e-12
    assert e.stats()["passes"] < len(c.ops)


@pytest.mark.parametrize("n,layers", [(6, 3), (13, 2)])
def test_adjoint_vs_finite_differences_and_parameter_shift(n, layers):
    print(f"case: adjoint n={n} layers={layers}", flush=True)
    c, idx, npar, row = ansatz(n, layers, seed=n)
    ham = tfim(n)
    est = AdjointEstimator(c, idx, npar, ham)
    energy, grad = est.energy_and_gradient(row[None, :])
    assert abs(energy[0] - oracle_energy(c, idx, row, ham)) < 1e-10
    h = 1e-4
    check = list(range(npar)) if npar <= 40 else list(range(0, npar, max(1, npar // 24)))
    for j in check:
        rp, rm = row.copy(), row.copy()
        rp[j] += h
        rm[j] -= h
        fd = (oracle_energy(c, idx, rp, ham) - oracle_energy(c, idx, rm, ham)) / (2 * h)
        ps = 0.5 * (oracle_energy(c, idx, row + (np.arange(npar) == j) * math.pi / 2, ham)
                    - oracle_energy(c, idx, row - (np.arange(npar) == j) * math.pi / 2, ham))
        assert abs(grad[0, j] - ps) < 1e-9, (j, grad[0, j], ps)
        assert abs(grad[0, j] - fd) < 1e-6, (j, grad[0, j], fd)


def test_adjoint_batch_and_torch_autograd():
    import torch
    n = 8
    print("case: adjoint batch + torch", flush=True)
    c, idx, npar, _ = ansatz(n, 2, seed=1)
    ham = tfim(n)
    est = AdjointEstimator(c, idx, npar, ham)
    rows = np.random.default_rng(5).uniform(-3, 3, size=(4, npar))
    e, g = est.energy_and_gradient(rows)
    for b in range(4):
        e1, g1 = est.energy_and_gradient(rows[b][None, :])
        assert abs(e[b] - e1[0]) < 1e-12 and np.max(np.abs(g[b] - g1[0])) < 1e-12
    theta = torch.tensor(rows, dtype=torch.float64, requires_grad=True)
    energies = adjoint_expectation(theta, est)
    energies.sum().backward()
    assert np.max(np.abs(theta.grad.numpy() - g)) < 1e-12
    ps = est.parameter_shift(rows[:1], which=[0, 5, npar - 1])
    assert np.max(np.abs(ps[0] - g[0, [0, 5, npar - 1]])) < 1e-9


@pytest.mark.parametrize("n,precision,tol", [(12, "double", 1e-10), (17, "double", 1e-10),
                                             (17, "single", 2e-4)])
def test_bound_rows_in_one_launch_vs_oracle(n, precision, tol):
    """aerhip_bind_pauli_expval: every row's energy against the oracle and,
    bit for bit in c128, against the row run on its own.  Row 0 binds every
    angle to 0 (its rotations are identities, so its program has another
    shape and runs apart from the rows that share one)."""
    print(f"case: bound rows n={n} {precision}", flush=True)
    c, idx, npar, _ = ansatz(n, 2, seed=n + 1)
    ham = tfim(n)
    est = AdjointEstimator(c, idx, npar, ham, precision=precision)
    rows = np.random.default_rng(n).uniform(-math.pi, math.pi, size=(6, npar))
    rows[0] = 0.0
    ms = []
    e = est.energy(rows, kernel_ms=ms)
    assert ms and ms[0] > 0
    want = np.array([oracle_energy(c, idx, r, ham) for r in rows])
    assert np.max(np.abs(e - want)) < tol, (e, want)
    if precision == "double":
        one = est.energy_one_by_one(rows)
        assert np.max(np.abs(e - one)) < 1e-12


def test_adjoint_bindings_together_match_one_by_one():
    """Five bindings of a 15-qubit ansatz (16 qubits with the flag, 32 tiles
    per state) run together; each row equals its own single-row run."""
    import torch
    n = 15
    print("case: adjoint 5 bindings together, n=15", flush=True)
    c, idx, npar, _ = ansatz(n, 2, seed=4)
    ham = tfim(n)
    est = AdjointEstimator(c, idx, npar, ham)
    rows = np.random.default_rng(9).uniform(-3, 3, size=(5, npar))
    e, g = est.energy_and_gradient(rows)
    for b in (0, 3):
        e1, g1 = est.energy_and_gradient(rows[b][None, :])
        assert abs(e[b] - e1[0]) < 1e-12 and np.max(np.abs(g[b] - g1[0])) < 1e-12
    assert abs(e[2] - oracle_energy(c, idx, rows[2], ham)) < 1e-10
    theta = torch.tensor(rows, dtype=torch.float64)
    with torch.no_grad():
        e_ng = adjoint_expectation(theta, est)
    assert np.max(np.abs(e_ng.numpy() - e)) < 1e-12


def test_bound_rows_and_gradients_with_swap_layers():
    """Bound rows (several states per launch) and the adjoint sweep through
    circuits holding qubit reversals, which run as bit-permutation passes:
    energies against the oracle, gradients against parameter shift."""
    n = 14
    print(f"case: bound rows with swap layers n={n}", flush=True)
    c, idx, npar, _ = ansatz(n, 2, seed=5)
    c2 = Circuit(n)
    idx2 = []
    half = len(c.ops) // 2
    for k, op in enumerate(c.ops):
        if k == half:
            for q in range(n // 2):
                c2.gate("swap", [q, n - 1 - q]); idx2.append(-1)
        c2.gate(op.name, list(op.qubits), list(op.params)); idx2.append(idx[k])
    for q in range(1, n // 2):
        c2.gate("swap", [q, n - q]); idx2.append(-1)
    ham = tfim(n)
    est = AdjointEstimator(c2, idx2, npar, ham)
    rows = np.random.default_rng(3).uniform(-math.pi, math.pi, size=(4, npar))
    e = est.energy(rows)
    want = np.array([oracle_energy(c2, idx2, r, ham) for r in rows])
    assert np.max(np.abs(e - want)) < 1e-10, (e, want)
    eg, g = est.energy_and_gradient(rows[:2])
    assert np.max(np.abs(eg - want[:2])) < 1e-10
    ps = est.parameter_shift(rows[:1], which=[0, 7, npar - 1])
    assert np.max(np.abs(ps[0] - g[0, [0, 7, npar - 1]])) < 1e-9
