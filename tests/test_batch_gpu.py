"""Batched circuits (one workgroup per circuit, one launch) against the oracle.

Each circuit is a hardware-efficient ansatz (RY, RZ on every qubit, then a CX
ladder) on its own qubit count; its energy is a random Pauli sum's
expectation value.  complex128 energies match the oracle's to 1e-10, complex64
to 1e-4; the batched launch and the circuit-by-circuit path agree.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from qiskit_aer_amd import circuits
from qiskit_aer_amd.batch import batch_pauli_expval

pytestmark = pytest.mark.gpu


def oracle_energy(c, ham):
    _, sv, _ = O.run_circuit(c, shots=0, seed=0, fusion=False)
    v = O.OracleQubitVector(c.num_qubits)
    v.initialize_from_vector(sv)
    q = list(range(c.num_qubits))
    return sum(coef * v.expval_pauli(q, lab) for lab, coef in ham)


def workload(sizes, seed):
    rng = np.random.default_rng(seed)
    cs, hs = [], []
    for i, n in enumerate(sizes):
        c, _ = circuits.hea(n, int(rng.integers(2, 5)), seed + i)
        cs.append(c)
        hs.append(circuits.random_pauli_sum(n, 8, seed + 100 + i))
    return cs, hs


@pytest.mark.parametrize("precision", ["double", "single"])
def test_batch_matches_oracle_and_sequential(precision):
    sizes = [3, 8, 10, 11, 12, 13, 14, 16, 12, 5]
    print(f"case: batch {precision} sizes={sizes}", flush=True)
    cs, hs = workload(sizes, 7)
    e_batch, ms = batch_pauli_expval(cs, hs, precision, "batch")
    e_seq, _ = batch_pauli_expval(cs, hs, precision, "sequential")
    want = np.array([oracle_energy(c, h) for c, h in zip(cs, hs)])
    tol = 1e-10 if precision == "double" else 1e-4
    assert np.max(np.abs(e_batch - want)) < tol
    assert np.max(np.abs(e_seq - want)) < tol
    assert ms > 0


def test_batch_edge_cases():
    print("case: batch edge cases", flush=True)
    # no terms for one circuit, a 1-qubit circuit, an empty circuit
    from qiskit_aer_amd.circuit import Circuit
    c0 = Circuit(1)
    c0.gate("h", [0])
    c1 = Circuit(4)
    c2, _ = circuits.hea(12, 2, 3)
    e, _ = batch_pauli_expval([c0, c1, c2], [[("X", 2.0)], [("ZZZZ", 1.0), ("XIII", 0.5)], []])
    assert abs(e[0] - 2.0) < 1e-12
    assert abs(e[1] - 1.0) < 1e-12
    assert e[2] == 0.0


@pytest.mark.parametrize("precision", ["double", "single"])
def test_batch_with_swap_layers(precision):
    """Circuits ending in a qubit reversal (a SWAP layer run as one
    bit-permutation pass) inside the batched launch."""
    sizes = [6, 12, 13, 16]
    print(f"case: batch swap layers {precision} sizes={sizes}", flush=True)
    cs, hs = workload(sizes, 11)
    for c in cs:
        n = c.num_qubits
        for q in range(n // 2):
            c.gate("swap", [q, n - 1 - q])
        c.gate("h", [0])
    e_batch, _ = batch_pauli_expval(cs, hs, precision, "batch")
    want = np.array([oracle_energy(c, h) for c, h in zip(cs, hs)])
    tol = 1e-10 if precision == "double" else 1e-4
    assert np.max(np.abs(e_batch - want)) < tol


@pytest.mark.parametrize("precision", ["double", "single"])
def test_batch_with_staged_tables(precision):
    """QFT ladders (fused diagonals with lane tables staged in LDS) and SU(4)
    layers in circuits of many tiles, where one workgroup walks all tiles of
    its circuit and restages the tables per tile."""
    sizes = [12, 14, 16, 13]
    print(f"case: batch staged tables {precision} sizes={sizes}", flush=True)
    cs, hs = [], []
    for i, n in enumerate(sizes):
        c = circuits.qft(n, swaps=False)
        c.ops += circuits.su4_brickwork(n, 3, 50 + i).ops
        cs.append(c)
        hs.append(circuits.random_pauli_sum(n, 8, 70 + i))
    e_batch, _ = batch_pauli_expval(cs, hs, precision, "batch")
    e_seq, _ = batch_pauli_expval(cs, hs, precision, "sequential")
    want = np.array([oracle_energy(c, h) for c, h in zip(cs, hs)])
    tol = 1e-10 if precision == "double" else 1e-4
    assert np.max(np.abs(e_batch - want)) < tol, (e_batch, want)
    assert np.max(np.abs(e_seq - want)) < tol, (e_seq, want)
