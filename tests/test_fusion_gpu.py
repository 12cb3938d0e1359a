"""Block fusion and the MFMA GEMM that applies dense 3..5-qubit matrices.

Fused circuits (every block fused, no cost check, widths 2..5) against the
oracle run without fusion; dense 3..5-qubit matrices on random qubits (the
GEMM on the LDS tile) against the oracle; the autotuner's measured choice."""
import numpy as np
import pytest

from oracle import pyoracle as O
from qiskit_aer_amd import circuits

pytestmark = pytest.mark.gpu


def haar(d, rng):
    z = (rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("single", 2e-5)])
@pytest.mark.parametrize("width", [2, 3, 4, 5])
@pytest.mark.parametrize("workload", ["brickwork", "su4"])
def test_fused_circuits_vs_oracle(workload, width, precision, tol):
    from qiskit_aer_amd import QubitVector
    n = 18
    print(f"case: block fusion {workload} width {width} {precision}", flush=True)
    if workload == "brickwork":
        c = circuits.brickwork_u3_cx(n, 8, seed=width, measure=False)
    else:
        c = circuits.su4_brickwork(n, 6, seed=width)
    qv = QubitVector(n, precision)
    qv.set_block_fusion(width, cost_check=False)
    qv.initialize()
    qv.apply_ops(c)
    got = qv.copy_to_vector()
    fused = qv.fused_count()
    qv.close()
    _, want, _ = O.run_circuit(c, shots=0, seed=0, fusion=False)
    err = float(np.max(np.abs(got - want)))
    print(f"  fused {fused} gates, max|engine - oracle| = {err:.2e}", flush=True)
    # SU(4) brickwork gates never share a pair with a neighbour: nothing fits width 2
    assert fused > 0 or (workload == "su4" and width == 2)
    assert err < tol


@pytest.mark.parametrize("precision,tol", [("double", 1e-12), ("single", 2e-5)])
def test_dense_matrices_on_3_to_5_qubits_gemm(precision, tol):
    from qiskit_aer_amd import QubitVector
    n = 16
    rng = np.random.default_rng(31)
    print(f"case: GEMM dense 3..5 qubits {precision}", flush=True)
    qv = QubitVector(n, precision)
    o = O.OracleQubitVector(n)
    qv.initialize_random(9)
    o.initialize_from_vector(qv.copy_to_vector())
    for it in range(18):
        k = 3 + it % 3
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        m = haar(1 << k, rng)
        qv.apply_matrix(qs, m)
        o.apply_matrix(qs, m)
    err = float(np.max(np.abs(qv.copy_to_vector() - o.copy_to_vector())))
    qv.close()
    print(f"  max|engine - oracle| = {err:.2e}", flush=True)
    assert err < tol


def test_autotuner_measures_every_width():
    from qiskit_aer_amd.autotune import tune
    print("case: autotune 18q brickwork", flush=True)
    c = circuits.brickwork_u3_cx(18, 8, seed=3, measure=False)
    best, times = tune(c, "double", widths=(0, 3, 4, 5))
    assert set(times) == {0, 3, 4, 5}
    assert all(t > 0 for t in times.values())
    assert best == min(times, key=times.get)
