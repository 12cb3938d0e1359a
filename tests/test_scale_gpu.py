"""The workloads at their defining sizes on one MI355X.

- 30 qubits (16 GiB, complex128): the seeded U3+CX brickwork followed by two
  Haar-SU(4) layers, against the CPU oracle (max-abs <= 1e-12); QFT-30 of a
  basis state against its closed form; QFT-30 then its inverse on an i.i.d.
  complex Gaussian state recovers the input.
- 33 qubits (128 GiB, every index path past 2^32): U then U^dagger returns
  |0...0>; marginals and sampled indices of a state with amplitude on qubit 32
  against a host CDF; Pauli expectation values on qubit 32.
- 32 qubits: complex64 against complex128 on the 20-layer Haar-SU(4) brickwork,
  1 - F <= 1e-6, mean |d<Z_i Z_i+1>| stated and bounded.
- Batched HEA circuits with 10..20 qubits and 64 Pauli terms each against the
  oracle; 24-qubit adjoint gradients against the oracle's central
  differences and parameter shift on a parameter subset.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from qiskit_aer_amd import circuits
from qiskit_aer_amd.circuit import Circuit

pytestmark = pytest.mark.gpu

SEED = 2024  # the seed of the benchmark workloads


def engine(n, precision="double"):
    from qiskit_aer_amd import QubitVector
    return QubitVector(n, precision)


def device_bytes_free():
    from qiskit_aer_amd.qubitvector import device_memory
    return device_memory(0)[0]


def need_device_gib(gib):
    if device_bytes_free() < gib * (1 << 30):
        pytest.skip(f"needs {gib} GiB of free device memory")


def max_abs_diff(a, b, chunk=1 << 26):
    m = 0.0
    for o in range(0, a.size, chunk):
        m = max(m, float(np.max(np.abs(a[o:o + chunk] - b[o:o + chunk]))))
    return m


def inverse_brickwork(c):
    """U^dagger of a U3+CX circuit: reversed order, U3(t, p, l)^-1 = U3(-t, -l, -p)."""
    inv = Circuit(c.num_qubits)
    for op in reversed(c.ops):
        if op.name == "u3":
            t, p, l = op.params
            inv.gate("u3", list(op.qubits), [-t, -l, -p])
        else:
            inv.gate(op.name, list(op.qubits), list(op.params))
    return inv


def inverse_qft(c):
    inv = Circuit(c.num_qubits)
    for op in reversed(c.ops):
        if op.name == "cp":
            inv.gate("cp", list(op.qubits), [-op.params[0]])
        else:
            inv.gate(op.name, list(op.qubits), list(op.params))
    return inv


def test_c2_brickwork_and_su4_30q_vs_oracle():
    n = 30
    print("case: 30q c128 brickwork (2 layers) + 2 SU(4) layers vs oracle", flush=True)
    need_device_gib(17)
    c = circuits.brickwork_u3_cx(n, 2, SEED, measure=False)
    for op in circuits.su4_brickwork(n, 2, SEED).ops:
        c.ops.append(op)
    e = engine(n)
    e.initialize()
    e.apply_ops(c)
    a = e.copy_to_vector()
    e.close()
    print("  engine done; oracle (Aer fusion on, as the reference at 30 qubits)", flush=True)
    _, b, _ = O.run_circuit(c, shots=0, seed=0)
    err = max_abs_diff(a, b)
    print(f"  max|engine - oracle| = {err:.3e}", flush=True)
    assert err <= 1e-12


def test_c2_qft30_basis_state_closed_form():
    """QFT|x> = N^-1/2 sum_k exp(2 pi i x k / N) |k> (qiskit's QFT with the final
    swaps is sqrt(N) * ifft)."""
    n = 30
    N = 1 << n
    x = 0x2B5C1F37 & (N - 1)
    print(f"case: QFT-30 of |{x}> vs closed form", flush=True)
    need_device_gib(17)
    c = Circuit(n)
    for q in range(n):
        if (x >> q) & 1:
            c.gate("x", [q])
    for op in circuits.qft(n).ops:
        c.ops.append(op)
    e = engine(n)
    e.initialize()
    e.apply_ops(c)
    a = e.copy_to_vector()
    e.close()
    err = 0.0
    chunk = 1 << 25
    s = 1.0 / math.sqrt(N)
    for o in range(0, N, chunk):
        k = np.arange(o, o + chunk, dtype=np.uint64)
        ph = ((k * np.uint64(x)) & np.uint64(N - 1)).astype(np.float64) * (2.0 * math.pi / N)
        ref = s * np.exp(1j * ph)
        err = max(err, float(np.max(np.abs(a[o:o + chunk] - ref))))
    print(f"  max|engine - closed form| = {err:.3e}", flush=True)
    assert err <= 1e-12


def test_c2_qft30_roundtrip_gaussian():
    n = 30
    print("case: QFT-30 then inverse on a Gaussian state", flush=True)
    need_device_gib(17)
    e = engine(n)
    e.initialize_random(SEED)
    a = e.copy_to_vector()
    q = circuits.qft(n)
    e.apply_ops(q)
    mid = e.probability(0)
    e.apply_ops(inverse_qft(q))
    b = e.copy_to_vector()
    e.close()
    err = max_abs_diff(a, b)
    print(f"  max|psi - QFT^-1 QFT psi| = {err:.3e}", flush=True)
    assert err <= 1e-12
    assert abs(mid - abs(a[0]) ** 2) > 1e-12  # the QFT did move amplitude


def test_c2_33q_roundtrip_and_high_indices():
    n = 33
    print("case: 33q c128 (128 GiB): U U^dagger, marginals and samples past 2^32", flush=True)
    need_device_gib(129)
    e = engine(n)
    e.initialize()
    c = circuits.brickwork_u3_cx(n, 3, SEED, measure=False)
    e.apply_ops(c)
    p0 = e.probability(0)
    assert abs(e.norm() - 1) < 1e-12
    e.apply_ops(inverse_brickwork(c))
    back = e.probability(0)
    print(f"  after U: p(0) = {p0:.3e}; after U U^dagger: 1 - p(0) = {1 - back:.3e}", flush=True)
    assert p0 < 0.5 and abs(1 - back) < 1e-12

    # amplitude on qubit 32 (index bit 32): H on 0, 31, 32, X on 20
    e.initialize()
    for q in (0, 31, 32):
        e.apply_mcu([q], np.array([[1, 1], [1, -1]]) / math.sqrt(2))
    e.apply_mcx([20])
    m32 = e.probabilities([32])
    assert np.max(np.abs(m32 - 0.5)) < 1e-15
    marg = e.probabilities([0, 20, 31, 32])
    want = np.zeros(16)
    for b0 in (0, 1):
        for b31 in (0, 1):
            for b32 in (0, 1):
                want[b0 | 2 | (b31 << 2) | (b32 << 3)] = 0.125
    assert np.max(np.abs(marg - want)) < 1e-15
    support = sorted(b0 | (1 << 20) | (b31 << 31) | (b32 << 32)
                     for b0 in (0, 1) for b31 in (0, 1) for b32 in (0, 1))
    probs = np.array([marg[(i & 1) | 2 | (((i >> 31) & 1) << 2) | (((i >> 32) & 1) << 3)]
                      for i in support])
    cdf = np.cumsum(probs)
    rnds = np.random.default_rng(5).uniform(size=4096)
    rnds = rnds[np.min(np.abs(rnds[:, None] - cdf[None, :]), axis=1) > 1e-9]
    got = e.sample_measure(rnds)
    want_idx = np.array(support, dtype=np.uint64)[np.searchsorted(cdf, rnds, side="right")]
    assert np.array_equal(got, want_idx)
    assert np.count_nonzero(got >= (1 << 32)) > 0
    assert abs(e.expval_pauli([32], "X") - 1) < 1e-14
    assert abs(e.expval_pauli([32], "Z")) < 1e-14
    assert abs(e.expval_pauli([20, 32], "XZ") + 1) < 1e-14  # Z on |1> of qubit 20
    e.close()


def test_c3_su4_32q_c64_vs_c128():
    n = 32
    print("case: 32q Haar-SU(4) brickwork, complex64 vs complex128", flush=True)
    need_device_gib(97)
    c = circuits.su4_brickwork(n, 20, SEED)
    d = engine(n, "double")
    s = engine(n, "single")
    for qv in (d, s):
        qv.initialize()
        qv.apply_ops(c)
    ov = d.inner_product(s)
    infid = 1 - abs(ov) ** 2
    zz = []
    for i in range(n - 1):
        zz.append(abs(d.expval_pauli([i, i + 1], "ZZ") - s.expval_pauli([i, i + 1], "ZZ")))
    print(f"  1 - F = {infid:.3e}, mean |d<ZZ>| = {np.mean(zz):.3e}, max {np.max(zz):.3e}, "
          f"norm c64 = {s.norm():.9f}", flush=True)
    assert infid <= 1e-6
    assert np.mean(zz) <= 1e-6
    assert abs(d.norm() - 1) < 1e-12
    d.close()
    s.close()


def test_c4_batch_up_to_20q_64_terms_vs_oracle():
    from qiskit_aer_amd.batch import batch_pauli_expval
    print("case: batched HEA circuits 10..20 qubits, 64 terms", flush=True)
    rng = np.random.default_rng(SEED)
    ns = [10, 20, 13, 17, 11, 19, 15, 12, 18, 14, 16, 20]
    cs, hs = [], []
    for i, n in enumerate(ns):
        c, _ = circuits.hea(n, int(rng.integers(2, 9)), SEED + i)
        cs.append(c)
        hs.append(circuits.random_pauli_sum(n, 64, SEED + 7919 + i))
    e128, _ = batch_pauli_expval(cs, hs, "double", "batch")
    e64, _ = batch_pauli_expval(cs, hs, "single", "batch")
    want = []
    for c, h in zip(cs, hs):
        _, sv, _ = O.run_circuit(c, shots=0, seed=0)
        v = O.OracleQubitVector(c.num_qubits)
        v.initialize_from_vector(sv)
        q = list(range(c.num_qubits))
        want.append(sum(coef * v.expval_pauli(q, lab) for lab, coef in h))
    want = np.array(want)
    err128 = np.max(np.abs(e128 - want))
    err64 = np.max(np.abs(e64 - want))
    print(f"  max |E - oracle|: c128 {err128:.3e}, c64 {err64:.3e}", flush=True)
    assert err128 <= 1e-10
    assert err64 <= 1e-4


def test_c5_adjoint_24q_vs_oracle_subset():
    """24-qubit RY/RZ + CZ ansatz, 12 layers (576 parameters), TFIM: the
    adjoint energy against the oracle, and the gradient at parameters 0, 287
    and 575 against the oracle's central differences (h = 1e-4) and
    parameter shift."""
    from qiskit_aer_amd.adjoint import AdjointEstimator, bind
    from tests.test_adjoint_gpu import ansatz, tfim
    n, layers = 24, 12
    print("case: adjoint 24q, 576 params vs oracle on 3 params", flush=True)
    c, idx, npar, row = ansatz(n, layers, seed=SEED)
    assert npar == 576
    ham = tfim(n)
    est = AdjointEstimator(c, idx, npar, ham)
    energy, grad = est.energy_and_gradient(row[None, :])

    def oracle_energy(r):
        _, sv, _ = O.run_circuit(bind(c, idx, r), shots=0, seed=0)
        v = O.OracleQubitVector(n)
        v.initialize_from_vector(sv)
        q = list(range(n))
        return sum(coef * v.expval_pauli(q, lab) for lab, coef in ham)

    e0 = oracle_energy(row)
    print(f"  energy {energy[0]:.12f}, oracle {e0:.12f}", flush=True)
    assert abs(energy[0] - e0) < 1e-10
    h = 1e-4
    for j in (0, 287, 575):
        unit = np.arange(npar) == j
        fd = (oracle_energy(row + h * unit) - oracle_energy(row - h * unit)) / (2 * h)
        ps = 0.5 * (oracle_energy(row + unit * math.pi / 2) - oracle_energy(row - unit * math.pi / 2))
        print(f"  param {j}: adjoint {grad[0, j]:.12f}, shift {ps:.12f}, fd {fd:.12f}", flush=True)
        assert abs(grad[0, j] - ps) < 1e-9
        assert abs(grad[0, j] - fd) < 1e-6
