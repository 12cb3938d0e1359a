"""Repeated flushes replayed as HIP graphs.

A flush whose pass launches repeat the previous flush's (same kernels, grids
and program offsets) is captured into a HIP graph the second time and
replayed after that, with the new program (other angles) uploaded to the same
device buffer first.  Every replay is checked against the oracle, with eager
work (readouts, re-initialisation) between the replays."""
import numpy as np
import pytest

from oracle import pyoracle as O
from qiskit_aer_amd import circuits
from tests import dense as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("single", 2e-5)])
def test_same_shape_other_angles_replays_graph(precision, tol):
    from qiskit_aer_amd import QubitVector
    n = 17
    print(f"case: graph replays, n={n} {precision}", flush=True)
    qv = QubitVector(n, precision)
    for it in range(6):
        c = circuits.brickwork_u3_cx(n, 6, seed=100 + it, measure=False)
        qv.initialize()
        qv.apply_ops(c)
        got = qv.copy_to_vector()
        _, want, _ = O.run_circuit(c, shots=0, seed=0)
        err = float(np.max(np.abs(got - want)))
        assert err < tol, (it, err)
        assert abs(qv.norm() - 1.0) < 10 * tol  # eager work between replays
    captures, replays = qv.graph_stats()
    qv.close()
    assert captures >= 1 and replays >= 3, (captures, replays)


def test_repeated_circuit_with_saved_values_replays_graph():
    """The same circuit applied again and again (no re-initialisation), with a
    saved expectation value after each application: the probe partials are
    written by the replayed kernels and summed outside the graph."""
    from qiskit_aer_amd import QubitVector
    n = 16
    print("case: repeated circuit with saved expectation values", flush=True)
    c = circuits.brickwork_u3_cx(n, 4, seed=7, measure=False)
    qv = QubitVector(n)
    o = O.OracleQubitVector(n)
    qv.initialize()
    o.initialize()
    qs = [0, 3, 8, 15]
    for it in range(5):
        qv.apply_ops(c)
        qv.save_expval_pauli(qs, "XZYZ")
        got = qv.saved_expvals()
        for op in c.ops:
            if op.name == "u3":
                o.apply_mcu(list(op.qubits), D.u3(*op.params))
            else:
                o.apply_mcx(list(op.qubits))
        want = o.expval_pauli(qs, "XZYZ")
        assert got.size == 1 and abs(got[0] - want) < 1e-12, (it, got, want)
    assert np.max(np.abs(qv.copy_to_vector() - o.copy_to_vector())) < 1e-12
    captures, replays = qv.graph_stats()
    qv.close()
    assert captures >= 1 and replays >= 2, (captures, replays)
