"""End-to-end parity of the controller + engine with the oracle's runner.

The 20-qubit U3+CX brickwork (20 layers, 1024 shots) of the reference's
CPU-runnable configuration: amplitudes within 1e-12 and identical counts for
the same seed.  QFT checked against the DFT (QFT psi = sqrt(N) ifft(psi)).
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from qiskit_aer_amd import AerSimulator, circuits
from qiskit_aer_amd.circuit import Circuit
from tests import dense as D

pytestmark = pytest.mark.gpu


def with_save(c: Circuit) -> Circuit:
    """Insert save_statevector before the first measurement."""
    out = Circuit(c.num_qubits, c.num_clbits, c.global_phase)
    saved = False
    for op in c.ops:
        if op.name == "measure" and not saved:
            out.save_statevector()
            saved = True
        out.ops.append(op)
    return out


@pytest.mark.parametrize("precision", ["double", "single"])
def test_brickwork_20q_amplitudes_and_counts(precision):
    print(f"case: brickwork 20q 20 layers {precision}", flush=True)
    c = circuits.brickwork_u3_cx(20, 20, seed=2024)
    sim = AerSimulator(precision=precision)
    res = sim.run(with_save(c), shots=1024, seed_simulator=9000)
    mem_o, sv_o, _ = O.run_circuit(c, shots=1024, seed=9000, precision="double")
    sv = res.get_statevector()
    if precision == "double":
        assert np.max(np.abs(sv - sv_o)) <= 1e-12
        assert np.array_equal(res.results[0].memory, mem_o)
    else:
        assert 1 - abs(np.vdot(sv, sv_o)) ** 2 <= 1e-6
    assert sum(res.get_counts().values()) == 1024


def test_qft_random_state_is_dft():
    n = 16
    print("case: qft 16q", flush=True)
    rng = np.random.default_rng(31)
    psi = D.random_state(n, rng)
    c = Circuit(n)
    c.initialize(psi, list(range(n)))
    c.ops.extend(circuits.qft(n).ops)
    c.save_statevector()
    res = AerSimulator().run(c, shots=0)
    assert np.max(np.abs(res.get_statevector() - np.sqrt(1 << n) * np.fft.ifft(psi))) <= 1e-12


def test_mid_circuit_measure_shots():
    print("case: mid-circuit measure", flush=True)
    c = Circuit(3, num_clbits=3)
    c.gate("h", [0]); c.gate("cx", [0, 1])
    c.measure([0], [0])
    c.gate("h", [2]); c.gate("cx", [2, 1])
    c.measure([1, 2], [1, 2])
    res = AerSimulator().run(c, shots=64, seed_simulator=5)
    mem_o, _, _ = O.run_circuit(c, shots=64, seed=5, statevector=False)
    assert np.array_equal(res.results[0].memory, mem_o)


def test_global_phase_and_reset():
    print("case: global phase / reset", flush=True)
    c = Circuit(4, global_phase=0.7)
    c.gate("h", [0]); c.gate("cx", [0, 3]); c.reset([3]); c.gate("ry", [2], [0.4])
    c.save_statevector()
    res = AerSimulator().run(c, shots=1, seed_simulator=11)
    _, sv_o, _ = O.run_circuit(c, shots=1, seed=11)
    assert np.max(np.abs(res.get_statevector() - sv_o)) <= 1e-12
