"""Pins the CPU oracle (oracle/) before it is trusted as the engine's checker.

Every restated function is checked against something independent of it:
the closed-form targets of the reference's own tests (tests/golden/), the
reference's own compilable code (oracle/_ref: Transformer, TransformerAVX2,
VMatrix, RngEngine, indexes), brute-force dense numpy operators, and
closed-form physics (expectation values of product / Bell states, QFT = DFT).
"""
import json
import math
import pathlib

import numpy as np
import pytest

from oracle import pyoracle as O
from qiskit_aer_amd import circuits
from qiskit_aer_amd.circuit import Circuit
from tests import dense as D

GOLDEN = pathlib.Path(__file__).with_name("golden") / "statevector_cases.json"


def circuit_from_case(case) -> Circuit:
    c = Circuit(case["num_qubits"])
    for name, qubits, params in case["ops"]:
        if name == "unitary":
            c.unitary(np.array(params["re"]) + 1j * np.array(params["im"]), qubits)
        elif name == "diagonal":
            c.diagonal(np.array(params["re"]) + 1j * np.array(params["im"]), qubits)
        else:
            c.gate(name, qubits, params)
    return c


@pytest.mark.parametrize("case", json.loads(GOLDEN.read_text()), ids=lambda c: c["name"])
def test_golden_statevectors(case):
    c = circuit_from_case(case)
    _, sv, _ = O.run_circuit(c, shots=0, seed=1)
    target = np.array(case["target_re"]) + 1j * np.array(case["target_im"])
    assert np.max(np.abs(sv - target)) < 1e-12, case["source"]


# ---- reference's own code (oracle/_ref) -----------------------------------

def aligned64(a: np.ndarray) -> np.ndarray:
    """Copy of `a` on a 64-byte boundary: the reference's AVX2 kernels use
    aligned loads on the state, which it allocates with posix_memalign(64)
    (qubitvector.hpp:951-967)."""
    buf = np.empty(a.nbytes + 64, dtype=np.uint8)
    off = (-buf.ctypes.data) % 64
    out = buf[off:off + a.nbytes].view(a.dtype).reshape(a.shape)
    out[...] = a
    return out


def ref_or_skip():
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    return R


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("avx", [0, 1])
def test_apply_matrix_vs_reference_transformer(k, avx):
    import ctypes as C
    R = ref_or_skip()
    rng = np.random.default_rng(100 + k)
    n = 9
    for trial in range(3):
        qubits = [int(q) for q in rng.choice(n, size=k, replace=False)]
        mat = D.random_unitary(1 << k, rng)
        psi = D.random_state(n, rng)
        v = O.OracleQubitVector(n)
        v.initialize_from_vector(psi)
        v.apply_matrix(qubits, mat)
        mine = v.copy_to_vector()
        theirs = aligned64(psi)
        q = np.array(qubits, dtype=np.uint64)
        m = np.ascontiguousarray(mat.reshape(-1, order="F"))
        assert R.ref_apply_matrix(theirs.ctypes.data, n, q.ctypes.data_as(C.POINTER(C.c_uint64)), k,
                                  m.ctypes.data, avx) == 0
        # Same arithmetic up to FMA contraction and (AVX2) summation order.
        assert np.max(np.abs(mine - theirs)) < 1e-15


@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_apply_diagonal_vs_reference_transformer(k):
    import ctypes as C
    R = ref_or_skip()
    rng = np.random.default_rng(7 + k)
    n = 8
    qubits = [int(q) for q in rng.choice(n, size=k, replace=False)]
    diag = np.exp(1j * rng.uniform(0, 2 * np.pi, size=1 << k))
    psi = D.random_state(n, rng)
    v = O.OracleQubitVector(n)
    v.initialize_from_vector(psi)
    v.apply_diagonal_matrix(qubits, diag)
    for avx in (0, 1):
        theirs = aligned64(psi)
        q = np.array(qubits, dtype=np.uint64)
        assert R.ref_apply_diagonal_matrix(theirs.ctypes.data, n, q.ctypes.data_as(C.POINTER(C.c_uint64)),
                                           k, np.ascontiguousarray(diag).ctypes.data, avx) == 0
        assert np.max(np.abs(v.copy_to_vector() - theirs)) < 1e-15


def gate_matrix_via_oracle(name, nq, params):
    """Columns of the gate's matrix obtained by applying it to basis states."""
    d = 1 << nq
    cols = []
    for j in range(d):
        c = Circuit(nq)
        e = np.zeros(d, dtype=complex)
        e[j] = 1
        c.initialize(e, list(range(nq)))
        c.gate(name, list(range(nq)), params)
        _, sv, _ = O.run_circuit(c, shots=0, seed=0)
        cols.append(sv)
    return np.array(cols).T


@pytest.mark.parametrize("name,ref_name,nq,params", [
    ("u3", "u3", 1, [0.3, 1.1, -2.2]),
    ("u", "u3", 1, [2.9, -0.4, 0.7]),
    ("r", "r", 1, [0.7, 0.2]),
    ("rx", "rx", 1, [1.3]),
    ("ry", "ry", 1, [-2.1]),
    ("rz", "rz", 1, [0.9]),
    ("rxx", "rxx", 2, [0.4]),
    ("ryy", "ryy", 2, [1.4]),
    ("rzx", "rzx", 2, [-0.8]),
    ("cu", "cu", 2, [0.7, -1.3, 2.1, 0.45]),
    ("cp", "cphase", 2, [1.25]),
    ("mcu", "u4", 1, [1.9, 0.3, -0.8, 1.1]),
    ("sx", "sx", 1, []),
    ("sxdg", "sxdg", 1, []),
    ("ecr", "ecr", 2, []),
    ("h", "h", 1, []),
])
def test_gate_matrices_vs_reference_vmatrix(name, ref_name, nq, params):
    import ctypes as C
    R = ref_or_skip()
    out = np.zeros(2 * 16)
    p = np.array(params + [0.0] * 4)
    cnt = R.ref_vmatrix(ref_name.encode(), p.ctypes.data, out.ctypes.data)
    assert cnt == 4 ** nq
    ref = (out[0:2 * cnt:2] + 1j * out[1:2 * cnt:2]).reshape((1 << nq, 1 << nq), order="F")
    mine = gate_matrix_via_oracle(name, nq, params)
    # The statevector dispatch builds H as u4(pi/2, 0, pi, 0), not Matrix::H.
    tol = 1e-15 if name != "h" else 2e-16
    assert np.max(np.abs(mine - ref)) <= tol


def test_rzz_diagonal_vs_reference_vmatrix():
    R = ref_or_skip()
    theta = 0.83
    out = np.zeros(2 * 16)
    p = np.array([theta, 0.0, 0.0, 0.0])
    cnt = R.ref_vmatrix(b"rzz_diag", p.ctypes.data, out.ctypes.data)
    assert cnt == 4
    ref = out[0:8:2] + 1j * out[1:8:2]
    mine = gate_matrix_via_oracle("rzz", 2, [theta])
    assert np.max(np.abs(mine - np.diag(ref))) <= 1e-15


def test_rng_draws_and_sampler_vs_reference_rngengine():
    """Counts of H|0> with seed s: draws r_i from RngEngine(s); outcome 1 iff
    r_i >= |<0|psi>|^2; memory is stored in reverse shot order."""
    R = ref_or_skip()
    shots, seed = 200, 12345
    draws = np.zeros(shots)
    R.ref_rng_uniform(seed, shots, draws.ctypes.data)
    c = Circuit(1)
    c.gate("h", [0])
    c.measure_all()
    mem, sv, _ = O.run_circuit(c, shots=shots, seed=seed)
    p0 = abs(sv[0]) ** 2
    expected = (draws >= p0).astype(np.uint64)[::-1]
    assert np.array_equal(mem, expected)


def test_discrete_distribution_vs_reference():
    import ctypes as C
    R = ref_or_skip()
    probs = np.array([0.1, 0.2, 0.3, 0.4])
    out = np.zeros(50, dtype=np.uint64)
    R.ref_rng_discrete(99, probs.ctypes.data, 4, 50, out.ctypes.data)
    # mid-circuit measurement of a 2-qubit state with those probabilities
    amps = np.sqrt(probs)
    got = []
    for shot in range(50):
        c = Circuit(2, num_clbits=2)
        c.initialize(amps, [0, 1])
        c.measure([0, 1], [0, 1])
        c.gate("x", [0])  # forces shot-by-shot execution
        mem, _, _ = O.run_circuit(c, shots=1, seed=99, statevector=False)
        got.append(mem[0])
    # shot s uses seed + s; the first draw of RngEngine(99) is out[0]
    assert got[0] == out[0]


def test_indexes_vs_reference():
    import ctypes as C
    R = ref_or_skip()
    rng = np.random.default_rng(3)
    for _ in range(20):
        k = int(rng.integers(1, 5))
        q = np.array(rng.choice(12, size=k, replace=False), dtype=np.uint64)
        kk = int(rng.integers(0, 1 << (12 - k)))
        out = np.zeros(1 << k, dtype=np.uint64)
        R.ref_indexes(q.ctypes.data_as(C.POINTER(C.c_uint64)), k, kk, out.ctypes.data)
        # independent: insert zeros at sorted positions, then OR bits of the position
        base = kk
        for s in sorted(q.tolist()):
            base = ((base >> s) << (s + 1)) | (base & ((1 << s) - 1))
        exp = [base | sum(((i >> j) & 1) << int(q[j]) for j in range(k)) for i in range(1 << k)]
        assert out.tolist() == exp


# ---- brute force ----------------------------------------------------------

def test_random_gates_vs_dense():
    rng = np.random.default_rng(11)
    n = 6
    psi = D.random_state(n, rng)
    v = O.OracleQubitVector(n)
    v.initialize_from_vector(psi)
    ref = psi.copy()
    for step in range(60):
        kind = step % 8
        if kind == 0:
            q = int(rng.integers(n)); m = D.random_unitary(2, rng)
            v.apply_matrix([q], m); ref = D.apply(ref, m, [q])
        elif kind == 1:
            qs = [int(x) for x in rng.choice(n, 3, replace=False)]; m = D.random_unitary(8, rng)
            v.apply_matrix(qs, m); ref = D.apply(ref, m, qs)
        elif kind == 2:
            qs = [int(x) for x in rng.choice(n, 3, replace=False)]
            v.apply_mcx(qs); ref = D.apply(ref, D.controlled(D.X, 2), qs)
        elif kind == 3:
            qs = [int(x) for x in rng.choice(n, 2, replace=False)]
            v.apply_mcy(qs); ref = D.apply(ref, D.controlled(D.Y, 1), qs)
        elif kind == 4:
            qs = [int(x) for x in rng.choice(n, 3, replace=False)]
            ph = np.exp(1j * rng.uniform(0, 6.28))
            v.apply_mcphase(qs, ph); ref = D.apply(ref, D.controlled(np.diag([1, ph]), 2), qs)
        elif kind == 5:
            qs = [int(x) for x in rng.choice(n, 3, replace=False)]
            sw = np.eye(4)[[0, 2, 1, 3]]
            v.apply_mcswap(qs); ref = D.apply(ref, D.controlled(sw, 1), qs)
        elif kind == 6:
            qs = [int(x) for x in rng.choice(n, 2, replace=False)]; m = D.random_unitary(2, rng)
            v.apply_mcu(qs, m); ref = D.apply(ref, D.controlled(m, 1), qs)
        else:
            qs = [int(x) for x in rng.choice(n, 3, replace=False)]
            lab = "".join(rng.choice(list("IXYZ"), 3))
            v.apply_pauli(qs, lab); ref = D.apply(ref, D.pauli_matrix(lab), qs)
    assert np.max(np.abs(v.copy_to_vector() - ref)) < 1e-13


def test_probabilities_and_expval_vs_dense():
    rng = np.random.default_rng(5)
    n = 7
    psi = D.random_state(n, rng)
    v = O.OracleQubitVector(n)
    v.initialize_from_vector(psi)
    for qs in ([3], [0, 5], [6, 1, 2], list(range(n)), [4, 3, 2, 1, 0, 6, 5]):
        p = v.probabilities(qs)
        t = np.abs(psi.reshape([2] * n)) ** 2
        ref = np.zeros(1 << len(qs))
        for idx in range(1 << n):
            m = sum(((idx >> q) & 1) << j for j, q in enumerate(qs))
            ref[m] += abs(psi[idx]) ** 2
        assert np.max(np.abs(p - ref)) < 1e-15
    for trial in range(10):
        k = int(rng.integers(1, 5))
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        lab = "".join(rng.choice(list("IXYZ"), k))
        e = v.expval_pauli(qs, lab)
        ref = np.vdot(psi, D.apply(psi, D.pauli_matrix(lab), qs)).real
        assert abs(e - ref) < 1e-14


def test_expval_closed_forms():
    for theta in (0.3, 1.7, -2.5):
        v = O.OracleQubitVector(2)
        v.initialize()
        v.apply_rotation([1], "y", theta)
        assert abs(v.expval_pauli([1], "Z") - math.cos(theta)) < 1e-15
        assert abs(v.expval_pauli([1], "X") - math.sin(theta)) < 1e-15
        assert abs(v.expval_pauli([0, 1], "ZI") - math.cos(theta)) < 1e-15
    bell = O.OracleQubitVector(2)
    bell.initialize()
    bell.apply_mcu([0], D.H)
    bell.apply_mcx([0, 1])
    for lab, val in (("ZZ", 1), ("XX", 1), ("YY", -1), ("ZI", 0), ("XY", 0)):
        assert abs(bell.expval_pauli([0, 1], lab) - val) < 1e-15


def test_sample_measure_vs_sequential_cdf():
    rng = np.random.default_rng(21)
    for n in (4, 9, 10, 12):
        psi = D.random_state(n, rng)
        v = O.OracleQubitVector(n)
        v.initialize_from_vector(psi)
        rnds = rng.uniform(size=300)
        got = v.sample_measure(rnds)
        cdf = np.cumsum(np.abs(psi) ** 2)
        exp = np.minimum(np.searchsorted(cdf, rnds, side="right"), (1 << n) - 1)
        assert np.array_equal(got, exp.astype(np.uint64))


def test_qft_is_dft():
    n = 8
    rng = np.random.default_rng(4)
    psi = D.random_state(n, rng)
    c = Circuit(n)
    c.initialize(psi, list(range(n)))
    for op in circuits.qft(n).ops:
        c.ops.append(op)
    _, sv, _ = O.run_circuit(c, shots=0, seed=0, fusion=False)
    assert np.max(np.abs(sv - np.sqrt(1 << n) * np.fft.ifft(psi))) < 1e-13


def test_fused_equals_unfused():
    c = circuits.brickwork_u3_cx(16, 6, seed=3, measure=False)
    _, a, _ = O.run_circuit(c, shots=0, fusion=True)
    _, b, _ = O.run_circuit(c, shots=0, fusion=False)
    assert np.max(np.abs(a - b)) < 1e-13
    widths = O.fused_widths(c)
    assert max(widths) <= 5 and len(widths) < c.count_gates()


def test_single_precision_close_to_double():
    c = circuits.brickwork_u3_cx(12, 10, seed=9, measure=False)
    _, a, _ = O.run_circuit(c, shots=0, fusion=False)
    _, b, _ = O.run_circuit(c, shots=0, fusion=False, precision="single")
    assert 1 - abs(np.vdot(a, b)) ** 2 < 1e-6


# ---- the remaining QubitVector methods: dense brute force / closed forms ----

def test_permutation_matrix_vs_dense():
    """apply_permutation_matrix (qubitvector.hpp:1350-1434) swaps the listed
    pairs of local basis states in order: new[i] = old[src[i]]."""
    rng = np.random.default_rng(31)
    n = 7
    for k in (1, 2, 3, 5, 6):
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        d = 1 << k
        pairs = [tuple(int(x) for x in rng.choice(d, 2, replace=False)) for _ in range(4)]
        psi = D.random_state(n, rng)
        v = O.OracleQubitVector(n)
        v.initialize_from_vector(psi)
        v.apply_permutation_matrix(qs, pairs)
        src = list(range(d))
        for a, b in pairs:
            src[a], src[b] = src[b], src[a]
        P = np.zeros((d, d))
        for i in range(d):
            P[i, src[i]] = 1
        assert np.array_equal(v.copy_to_vector(), D.apply(psi, P, qs))


def test_multi_swaps_vs_dense():
    """apply_multi_swaps (qubitvector.hpp:1843-1871): in groups of ten
    qubits, SWAP of the pairs (q0, q1), (q2, q3), ... of each group."""
    rng = np.random.default_rng(32)
    n = 12
    SW = np.eye(4)[[0, 2, 1, 3]]
    for qs in ([3, 8, 0, 11], [int(x) for x in rng.permutation(n)]):
        psi = D.random_state(n, rng)
        v = O.OracleQubitVector(n)
        v.initialize_from_vector(psi)
        v.apply_multi_swaps(qs)
        ref = psi
        for g in range(0, len(qs), 10):
            grp = qs[g:g + 10]
            for j in range(0, len(grp) - 1, 2):
                ref = D.apply(ref, SW, [grp[j], grp[j + 1]])
        assert np.array_equal(v.copy_to_vector(), ref)


def test_multiplexer_vs_dense():
    """apply_multiplexer (qubitvector.hpp:1306-1340): the control value b
    (bit i = controls[i]) selects the block M_b(i, j) = mat[(b*cols + i) +
    DIM*j] acting on the targets; as one operator on targets + controls it is
    block diagonal."""
    rng = np.random.default_rng(33)
    n = 7
    for nt, nc in ((1, 1), (1, 2), (2, 1), (2, 2), (3, 1)):
        qs = [int(x) for x in rng.choice(n, nt + nc, replace=False)]
        tq, cq = qs[:nt], qs[nt:]
        cols, blocks = 1 << nt, 1 << nc
        mats = [D.random_unitary(cols, rng) for _ in range(blocks)]
        stacked = np.vstack(mats)  # DIM x cols, block b in rows b*cols ...
        psi = D.random_state(n, rng)
        v = O.OracleQubitVector(n)
        v.initialize_from_vector(psi)
        v.apply_multiplexer(cq, tq, stacked)
        full = np.zeros((cols * blocks, cols * blocks), dtype=complex)
        for b in range(blocks):
            full[b * cols:(b + 1) * cols, b * cols:(b + 1) * cols] = mats[b]
        assert np.max(np.abs(v.copy_to_vector() - D.apply(psi, full, tq + cq))) < 1e-14


def test_initialize_component_closed_form():
    """initialize_component (qubitvector.hpp:879-896): within every group,
    new[i] = old[index0] * state[i] (bit j of i = qubits[j])."""
    rng = np.random.default_rng(34)
    n = 8
    for qs in ([2], [5, 1], [0, 7, 3], [6, 2, 4, 1, 0, 7]):
        k = len(qs)
        psi = D.random_state(n, rng)
        state = D.random_state(k, rng)
        v = O.OracleQubitVector(n)
        v.initialize_from_vector(psi)
        v.initialize_component(qs, state)
        ref = psi.copy()
        qmask = sum(1 << q for q in qs)
        for idx in range(1 << n):
            i = sum(((idx >> q) & 1) << j for j, q in enumerate(qs))
            ref[idx] = psi[idx & ~qmask] * state[i]
        assert np.max(np.abs(v.copy_to_vector() - ref)) < 1e-15


def _collapse_case(kind, qubits, seed):
    n = 5
    rng = np.random.default_rng(40 + seed)
    psi = D.random_state(n, rng)
    c = Circuit(n, num_clbits=len(qubits))
    c.initialize(psi, list(range(n)))
    if kind == "measure":
        c.measure(qubits, list(range(len(qubits))))
    else:
        c.reset(qubits)
    c.save_statevector()
    mem, sv, _ = O.run_circuit(c, shots=1, seed=seed)
    return psi, mem, sv


def _draw_outcome(psi, qubits, seed):
    """First RngEngine(seed).rand_int(probs) of the reference (rng.hpp:94-96)."""
    import ctypes as C
    R = ref_or_skip()
    probs = np.zeros(1 << len(qubits))
    for idx in range(psi.size):
        probs[sum(((idx >> q) & 1) << j for j, q in enumerate(qubits))] += abs(psi[idx]) ** 2
    out = np.zeros(1, dtype=np.uint64)
    R.ref_rng_discrete(seed, probs.ctypes.data, probs.size, 1, out.ctypes.data)
    return int(out[0]), probs


@pytest.mark.parametrize("qubits,seed", [([2], 1), ([1, 3], 2), ([4, 0, 2], 3)])
def test_measure_collapse_closed_form(qubits, seed):
    """Post-measurement state = P_m psi / sqrt(p_m) with m the reference's
    own draw (statevector_state.hpp:936-1014)."""
    psi, mem, sv = _collapse_case("measure", qubits, seed)
    m, probs = _draw_outcome(psi, qubits, seed)
    assert int(mem[0]) == m
    ref = np.zeros_like(psi)
    for idx in range(psi.size):
        if sum(((idx >> q) & 1) << j for j, q in enumerate(qubits)) == m:
            ref[idx] = psi[idx] / math.sqrt(probs[m])
    assert np.max(np.abs(sv - ref)) < 1e-15


@pytest.mark.parametrize("qubits,seed", [([2], 5), ([1, 3], 6), ([4, 0, 2], 7)])
def test_reset_closed_form(qubits, seed):
    """Reset = projection on the drawn outcome m, renormalised, then the basis
    state m of the reset qubits moved to |0...0> (statevector_state.hpp:
    952-1011)."""
    psi, _, sv = _collapse_case("reset", qubits, seed)
    m, probs = _draw_outcome(psi, qubits, seed)
    qmask = sum(1 << q for q in qubits)
    ref = np.zeros_like(psi)
    for idx in range(psi.size):
        if sum(((idx >> q) & 1) << j for j, q in enumerate(qubits)) == m:
            ref[idx & ~qmask] = psi[idx] / math.sqrt(probs[m])
    assert np.max(np.abs(sv - ref)) < 1e-15


def test_reference_transformer_hook_matches_restatement():
    """The CPU baseline runs the oracle's driver with the reference's own
    compiled Transformer(AVX2) for matrices and diagonals: it must give the
    restatement's state (fused brickwork and QFT, 16 qubits, 8 threads)."""
    ref_or_skip()
    c = circuits.brickwork_u3_cx(16, 6, seed=3, measure=False)
    q = circuits.qft(16)
    for op in q.ops:
        c.ops.append(op)
    _, a, _ = O.run_circuit(c, shots=0, threads=8)
    with O.reference_transformer() as rt:
        _, b, _ = O.run_circuit(c, shots=0, threads=8)
    _, a2, _ = O.run_circuit(c, shots=0, threads=8)
    assert np.max(np.abs(a - b)) < 1e-13
    assert np.array_equal(a, a2)  # the hook is removed again
    assert isinstance(rt.avx2, bool)
