"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

complex128: max |engine - oracle| <= 1e-12 on amplitudes and FP64 readouts;
index / permutation / sign work (mcx, mcswap, Pauli flips, sampled indices)
bit-exact; complex64: max error <= 1e-5 against the complex128 oracle.
"""
import numpy as np
import pytest

from oracle import pyoracle as O
from tests import dense as D

pytestmark = pytest.mark.gpu

TOL = {"double": 1e-12, "single": 2e-5}


def engine(n, precision="double"):
    from qiskit_aer_amd import QubitVector
    return QubitVector(n, precision)


def pair(n, seed, precision="double"):
    rng = np.random.default_rng(seed)
    psi = D.random_state(n, rng)
    e = engine(n, precision)
    e.initialize_from_vector(psi)
    o = O.OracleQubitVector(n)
    o.initialize_from_vector(psi)
    return e, o, rng


def close(e, o, precision="double"):
    a, b = e.copy_to_vector(), o.copy_to_vector()
    err = np.max(np.abs(a - b))
    assert err <= TOL[precision], err


def test_initialize_and_copy():
    print("case: initialize n=12", flush=True)
    e = engine(12)
    e.initialize()
    v = e.copy_to_vector()
    assert v[0] == 1 and np.count_nonzero(v) == 1
    rng = np.random.default_rng(0)
    psi = D.random_state(12, rng)
    e.initialize_from_vector(psi)
    assert np.array_equal(e.copy_to_vector(), psi)


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("n", [3, 11, 14])
def test_single_qubit_matrices_every_target(n, precision):
    print(f"case: 1q matrices n={n} {precision}", flush=True)
    e, o, rng = pair(n, 1000 + n, precision)
    for q in range(n):
        m = D.random_unitary(2, rng)
        e.apply_matrix([q], m)
        o.apply_matrix([q], m)
    close(e, o, precision)


@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_dense_matrices(k):
    n = 14
    print(f"case: dense k={k} n={n}", flush=True)
    e, o, rng = pair(n, 2000 + k)
    for trial in range(6):
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        m = D.random_unitary(1 << k, rng)
        e.apply_matrix(qs, m)
        o.apply_matrix(qs, m)
    close(e, o)


@pytest.mark.parametrize("k", [6, 7])
def test_dense_group_matrices(k):
    n = 12
    print(f"case: dense group k={k}", flush=True)
    e, o, rng = pair(n, 2100 + k)
    qs = [int(x) for x in rng.choice(n, k, replace=False)]
    m = D.random_unitary(1 << k, rng)
    e.apply_matrix(qs, m)
    o.apply_matrix(qs, m)
    close(e, o)


def test_diagonals_and_phases():
    n = 15
    print("case: diagonals n=15", flush=True)
    e, o, rng = pair(n, 3)
    for k in (1, 2, 3, 6):
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        d = np.exp(1j * rng.uniform(0, 6.3, 1 << k))
        e.apply_diagonal_matrix(qs, d)
        o.apply_diagonal_matrix(qs, d)
    for k in (1, 2, 4):
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        ph = np.exp(1j * rng.uniform(0, 6.3))
        e.apply_mcphase(qs, ph)
        o.apply_mcphase(qs, ph)
    e.apply_mcphase([14, 2], -1)
    o.apply_mcphase([14, 2], -1)
    close(e, o)


def test_permutations_bit_exact():
    n = 15
    print("case: permutations n=15", flush=True)
    e, o, rng = pair(n, 4)
    for k in (1, 2, 3, 4):
        for _ in range(3):
            qs = [int(x) for x in rng.choice(n, k, replace=False)]
            e.apply_mcx(qs)
            o.apply_mcx(qs)
    for k in (2, 3):
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        e.apply_mcswap(qs)
        o.apply_mcswap(qs)
    e.apply_permutation_matrix([3, 12], [(0, 3), (1, 2)])
    o.apply_permutation_matrix([3, 12], [(0, 3), (1, 2)])
    e.apply_multi_swaps([0, 13, 5, 9])
    o.apply_multi_swaps([0, 13, 5, 9])
    assert np.array_equal(e.copy_to_vector(), o.copy_to_vector())


def test_mcy_mcu_rotation_pauli():
    n = 13
    print("case: mcy/mcu/rotation/pauli n=13", flush=True)
    e, o, rng = pair(n, 5)
    for k in (1, 2, 3):
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        e.apply_mcy(qs); o.apply_mcy(qs)
        m = D.random_unitary(2, rng)
        e.apply_mcu(qs, m); o.apply_mcu(qs, m)
        e.apply_mcu(qs, np.diag([1, 1j])); o.apply_mcu(qs, np.diag([1, 1j]))
        e.apply_mcu(qs, np.diag([-1j, 0.5])); o.apply_mcu(qs, np.diag([-1j, 0.5]))
    for axis in ("x", "y", "z"):
        qs = [int(x) for x in rng.choice(n, 2, replace=False)]
        e.apply_rotation(qs, axis, 0.37); o.apply_rotation(qs, axis, 0.37)
    for axis in ("xx", "yy", "zz", "zx"):
        qs = [int(x) for x in rng.choice(n, 2, replace=False)]
        e.apply_rotation(qs, axis, -1.1); o.apply_rotation(qs, axis, -1.1)
    for _ in range(6):
        k = int(rng.integers(1, 6))
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        lab = "".join(rng.choice(list("IXYZ"), k))
        e.apply_pauli(qs, lab, 0.3 - 0.7j); o.apply_pauli(qs, lab, 0.3 - 0.7j)
    close(e, o)


def test_multiplexer_and_initialize_component():
    n = 12
    print("case: multiplexer n=12", flush=True)
    e, o, rng = pair(n, 6)
    mats = [D.random_unitary(2, rng) for _ in range(4)]
    stacked = np.vstack(mats)  # (8 x 2): rows blocked by control value
    e.apply_multiplexer([9, 2], [5], stacked); o.apply_multiplexer([9, 2], [5], stacked)
    st = D.random_state(2, rng)
    e.initialize_component([1, 10], st); o.initialize_component([1, 10], st)
    close(e, o)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_long_random_sequence_many_passes(precision):
    """Hundreds of mixed gates on 18 qubits: exercises the planner across tiles."""
    n = 18
    print(f"case: long sequence n=18 {precision}", flush=True)
    e, o, rng = pair(n, 7, precision)
    for step in range(400):
        r = step % 6
        if r in (0, 1):
            q = int(rng.integers(n)); m = D.random_unitary(2, rng)
            e.apply_matrix([q], m); o.apply_matrix([q], m)
        elif r == 2:
            qs = [int(x) for x in rng.choice(n, 2, replace=False)]
            e.apply_mcx(qs); o.apply_mcx(qs)
        elif r == 3:
            qs = [int(x) for x in rng.choice(n, 2, replace=False)]
            ph = np.exp(1j * rng.uniform(0, 6.3))
            e.apply_mcphase(qs, ph); o.apply_mcphase(qs, ph)
        elif r == 4:
            qs = [int(x) for x in rng.choice(n, 2, replace=False)]; m = D.random_unitary(4, rng)
            e.apply_matrix(qs, m); o.apply_matrix(qs, m)
        else:
            qs = [int(x) for x in rng.choice(n, 3, replace=False)]
            e.apply_mcu(qs, D.random_unitary(2, rng) if False else D.H)
            o.apply_mcu(qs, D.H)
    close(e, o, precision)
    st = e.stats()
    assert st["passes"] < st["gates"]


@pytest.mark.parametrize("precision", ["double", "single"])
def test_brickwork_light_cone_passes(precision):
    """A nearest-neighbour U3+CX brickwork on 20 qubits: the planner seeds
    tiles with windows of consecutive qubits, so a pass runs a whole light
    cone (8 passes for complex128, 7 for complex64; taking primitives in
    program order from the low qubits alone gives 14 and 10)."""
    n, layers = 20, 16
    print(f"case: brickwork light cones n={n} {precision}", flush=True)
    e, o, rng = pair(n, 11, precision)
    for layer in range(layers):
        for q in range(n):
            m = D.u3(*rng.uniform(0.0, 2.0 * np.pi, 3))
            e.apply_matrix([q], m); o.apply_matrix([q], m)
        for a in range(layer % 2, n - 1, 2):
            e.apply_mcx([a, a + 1]); o.apply_mcx([a, a + 1])
    close(e, o, precision)
    assert e.stats()["passes"] <= 8


def test_readouts():
    n = 16
    print("case: readouts n=16", flush=True)
    e, o, rng = pair(n, 8)
    assert abs(e.norm() - o.norm()) <= 1e-13
    for qs in ([], [0], [15], [3, 9], [12, 1, 7], list(range(16)), [5, 4, 3, 2, 1, 0, 15, 14, 13, 12, 11, 10, 9, 8, 7, 6],
               list(range(1, 16))):
        a, b = e.probabilities(qs), o.probabilities(qs)
        assert np.max(np.abs(a - b)) <= 1e-13, qs
    for _ in range(12):
        k = int(rng.integers(1, 8))
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        lab = "".join(rng.choice(list("IXYZ"), k))
        ph = complex(*rng.standard_normal(2))
        assert abs(e.expval_pauli(qs, lab, ph) - o.expval_pauli(qs, lab, ph)) <= 1e-12
    for i in (0, 12345, (1 << n) - 1):
        assert abs(e.probability(i) - abs(o.copy_to_vector()[i]) ** 2) <= 1e-15


@pytest.mark.parametrize("n", [6, 10, 17])
def test_sample_measure_identical(n):
    print(f"case: sample_measure n={n}", flush=True)
    e, o, rng = pair(n, 9 + n)
    rnds = rng.uniform(size=4096)
    a, b = e.sample_measure(rnds), o.sample_measure(rnds)
    assert np.array_equal(a, b)


def test_errors_are_raised():
    from qiskit_aer_amd import _lib as L
    print("case: errors", flush=True)
    e = engine(5)
    with pytest.raises(L.AerhipInvalidArgument):
        e.apply_mcx([0, 7])
    with pytest.raises(L.AerhipInvalidArgument):
        e.apply_matrix([1, 1], np.eye(4))
    with pytest.raises(L.AerhipError):
        e.initialize_from_vector(np.ones(7))
    with pytest.raises(L.AerhipInvalidArgument):
        e.apply_pauli([0], "Q")


def test_random_state_is_normalised_gaussian():
    print("case: initialize_random n=16", flush=True)
    e = engine(16)
    e.initialize_random(123)
    v = e.copy_to_vector()
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    z = v * np.sqrt(v.size)  # unit-variance complex Gaussian samples
    assert abs(np.mean(np.abs(z) ** 2) - 1) < 0.02
    assert abs(np.mean(z.real)) < 0.02 and abs(np.mean(z.imag)) < 0.02
    e.initialize_random(123)
    assert np.array_equal(e.copy_to_vector(), v)


@pytest.mark.parametrize("precision,reg_bits", [("double", 0), ("double", 3), ("double", 4),
                                                ("single", 0), ("single", 4), ("single", 5)])
def test_register_layouts_match_oracle(precision, reg_bits):
    """Every register-blocking setting gives the oracle's state (brickwork
    U3+CX plus RY/RZ/CZ/SU(4)/swap layers on 16 qubits)."""
    from qiskit_aer_amd import circuits
    n = 16
    print(f"case: register layouts {precision} R={reg_bits}", flush=True)
    e, o, rng = pair(n, 40 + reg_bits, precision)
    e.set_register_bits(reg_bits)
    c = circuits.brickwork_u3_cx(n, 8, seed=5, measure=False)
    e.apply_ops(c)
    for op in c.ops:
        if op.name == "u3":
            o.apply_mcu(list(op.qubits), D.u3(*op.params))
        else:
            o.apply_mcx(list(op.qubits))
    for q in range(n):
        t = rng.uniform(-3, 3)
        e.apply_rotation([q], "y", t); o.apply_rotation([q], "y", t)
        e.apply_rotation([q], "z", -t); o.apply_rotation([q], "z", -t)
    for a in range(0, n - 1, 2):
        e.apply_mcphase([a, a + 1], -1); o.apply_mcphase([a, a + 1], -1)
        m = D.random_unitary(4, rng)
        e.apply_matrix([a + 1, a], m); o.apply_matrix([a + 1, a], m)
    e.apply_mcswap([3, 12]); o.apply_mcswap([3, 12])
    for q in range(n // 2):  # a reversal: one bit-permutation pass
        e.apply_mcswap([q, n - 1 - q]); o.apply_mcswap([q, n - 1 - q])
    # controlled gates with controls inside and outside the tile, Y, and a
    # 2-qubit diagonal
    for ctrl, tgt in ((9, 1), (15, 2), (0, 14), (13, 7)):
        u = D.random_unitary(2, rng)
        e.apply_mcu([ctrl, tgt], u); o.apply_mcu([ctrl, tgt], u)
    e.apply_mcx([2, 15, 5]); o.apply_mcx([2, 15, 5])
    e.apply_mcy([6, 1]); o.apply_mcy([6, 1])
    d = np.exp(1j * rng.uniform(-3, 3, 4))
    e.apply_diagonal_matrix([4, 11], d); o.apply_diagonal_matrix([4, 11], d)
    for q in (0, 5, 10, 15):
        m = D.random_unitary(2, rng)
        e.apply_matrix([q], m); o.apply_matrix([q], m)
    labels = ["XYZIX", "ZZIIY", "IXXYZ"]
    for lab in labels:
        e.save_expval_pauli([1, 4, 9, 12, 15], lab)
    got = e.saved_expvals()
    want = [o.expval_pauli([1, 4, 9, 12, 15], lab) for lab in labels]
    assert np.max(np.abs(np.array(got) - np.array(want))) < (1e-12 if precision == "double" else 1e-5)
    close(e, o, precision)
    if reg_bits:
        assert e.stats()["gathers"] > 0


@pytest.mark.parametrize("pa,pb", [("double", "double"), ("single", "double"), ("double", "single")])
def test_inner_product(pa, pb):
    """<a|b> on the device equals numpy's vdot of the copied-out states."""
    n = 14
    print(f"case: inner product {pa}/{pb}", flush=True)
    rng = np.random.default_rng(11)
    u, v = D.random_state(n, rng), D.random_state(n, rng)
    a, b = engine(n, pa), engine(n, pb)
    a.initialize_from_vector(u)
    b.initialize_from_vector(v)
    got = a.inner_product(b)
    want = np.vdot(a.copy_to_vector(), b.copy_to_vector())
    assert abs(got - want) < 1e-12
    assert abs(a.inner_product(a) - 1) < (1e-12 if pa == "double" else 1e-6)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_pauli_strings_wider_than_a_tile(precision):
    """apply_pauli and saved expectation values with X/Y on more qubits than a
    tile holds (split into a sign/phase part and single-qubit X; probe sweeps)."""
    n = 15
    print(f"case: wide Pauli strings {precision}", flush=True)
    e, o, rng = pair(n, 12, precision)
    qs = list(range(n))
    lab = "".join(rng.choice(list("XY"), n - 1)) + "Z"
    e.apply_pauli(qs, lab, 0.6 + 0.8j)
    o.apply_pauli(qs, lab, 0.6 + 0.8j)
    close(e, o, precision)
    labels = ["".join(rng.choice(list("XYZ"), n)) for _ in range(3)] + ["Y" * n, "I" * (n - 1) + "X"]
    for lab in labels:
        e.save_expval_pauli(qs, lab, 0.3 - 0.4j)
    m = D.random_unitary(2, rng)
    e.apply_matrix([2], m)
    o2 = [o.expval_pauli(qs, lab, 0.3 - 0.4j) for lab in labels]
    got = e.saved_expvals()
    assert np.max(np.abs(np.array(got) - np.array(o2))) < (1e-12 if precision == "double" else 1e-5)


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_register_relabelling_random_mix(precision, seed):
    """Random mixes that drive the register relabelling: X/CX between
    register qubits (absorbed), 2x2 gates on relabelled qubits, diagonals and
    phases on relabelled and thread bits, CX controlled by a qubit outside the
    tile (a whole-tile X), controlled 2x2 and 4x4 gates (which need a fresh
    layout), and saved Pauli values read through a relabelled layout."""
    n = 16
    print(f"case: relabelling mix {precision} seed={seed}", flush=True)
    e, o, rng = pair(n, 700 + seed, precision)
    near = [0, 1, 2, 3, 4, 5]
    far = [12, 15]
    for _ in range(240):
        kind = rng.integers(0, 10)
        a, b = (int(x) for x in rng.choice(near, 2, replace=False))
        if kind <= 2:
            e.apply_mcx([a, b]); o.apply_mcx([a, b])
        elif kind == 3:
            e.apply_mcx([b]); o.apply_mcx([b])
        elif kind <= 5:
            m = D.random_unitary(2, rng)
            e.apply_matrix([a], m); o.apply_matrix([a], m)
        elif kind == 6:
            t = rng.uniform(-3, 3)
            e.apply_rotation([a], "z", t); o.apply_rotation([a], "z", t)
            ph = np.exp(1j * rng.uniform(-3, 3))
            e.apply_mcphase([a, b], ph); o.apply_mcphase([a, b], ph)
        elif kind == 7:
            c = int(rng.choice(far))
            e.apply_mcx([c, a]); o.apply_mcx([c, a])
        elif kind == 8:
            u = D.random_unitary(2, rng)
            e.apply_mcu([b, a], u); o.apply_mcu([b, a], u)
        else:
            m = D.random_unitary(4, rng)
            e.apply_matrix([a, b], m); o.apply_matrix([a, b], m)
    if seed == 2:
        labels = ["XYZIXZ", "ZZIIYX", "IXXYZY"]
        for lab in labels:
            e.save_expval_pauli(near, lab)
        got = e.saved_expvals()
        want = [o.expval_pauli(near, lab) for lab in labels]
        assert np.max(np.abs(np.array(got) - np.array(want))) <= 10 * TOL[precision]
    close(e, o, precision)


@pytest.mark.parametrize("precision", ["double", "single"])
def test_phase_and_diagonal_register_masks(precision):
    """Phases and single-qubit diagonals under register relabelling, with
    predicates that mix register bits, thread bits and qubits fixed by the
    tile: each such op carries a register mask that the host computes from
    the relabelling in force."""
    n = 16
    print(f"case: phase/diagonal register masks {precision}", flush=True)
    e, o, rng = pair(n, 900, precision)
    near = [0, 1, 2, 3, 4, 5, 6, 7]
    far = [11, 13, 15]
    for _ in range(300):
        kind = rng.integers(0, 8)
        a, b = (int(x) for x in rng.choice(near, 2, replace=False))
        if kind <= 2:
            e.apply_mcx([a, b]); o.apply_mcx([a, b])
        elif kind == 3:
            e.apply_mcx([b]); o.apply_mcx([b])
        elif kind == 4:
            k = int(rng.integers(1, 4))
            qs = [int(x) for x in rng.choice(near + far, k, replace=False)]
            ph = np.exp(1j * rng.uniform(-3, 3))
            e.apply_mcphase(qs, ph); o.apply_mcphase(qs, ph)
        elif kind == 5:
            q = int(rng.choice(near + far))
            t = rng.uniform(-3, 3)
            e.apply_rotation([q], "z", t); o.apply_rotation([q], "z", t)
        elif kind == 6:
            d = np.exp(1j * rng.uniform(-3, 3, 2))
            e.apply_diagonal_matrix([a], d); o.apply_diagonal_matrix([a], d)
        else:
            m = D.random_unitary(2, rng)
            e.apply_matrix([a], m); o.apply_matrix([a], m)
    close(e, o, precision)



def test_multiplexer_many_targets_with_controls():
    """6 targets with 2 controls runs as a group pass per control value; each
    applies only where the controls hold their value."""
    n = 12
    print("case: multiplexer 6 targets + 2 controls", flush=True)
    e, o, rng = pair(n, 13)
    targets, controls = [0, 3, 5, 7, 8, 11], [2, 9]
    mats = [D.random_unitary(64, rng) for _ in range(4)]
    stacked = np.vstack(mats)
    e.apply_multiplexer(controls, targets, stacked)
    o.apply_multiplexer(controls, targets, stacked)
    close(e, o)
    st = D.random_state(7, rng)
    qs = [4, 10, 1, 6, 9, 0, 11]
    e.initialize_component(qs, st)
    o.initialize_component(qs, st)
    close(e, o)


def test_dense_group_pass_at_30_qubits():
    """A 6-qubit dense matrix on a 30-qubit state launches more groups than one
    workgroup each could cover: the kernel loops over groups.  The matrix is
    X on qubit 29 written densely, so the result must equal apply_mcx."""
    from qiskit_aer_amd.qubitvector import device_memory
    n = 30
    print("case: dense 6-qubit matrix at 30 qubits", flush=True)
    if device_memory(0)[0] < 34 * (1 << 30):
        pytest.skip("needs 34 GiB of free device memory")
    a, b = engine(n), engine(n)
    a.initialize_random(77)
    b.initialize_from_device(a.device_pointer(), 1 << n, "double")
    qs = [0, 5, 29, 13, 2, 21]
    M = np.zeros((64, 64))
    for j in range(64):
        M[j ^ 4, j] = 1
    a.apply_mcx([29])
    b.apply_matrix(qs, M)
    ov = a.inner_product(b)
    assert abs(ov - 1) < 1e-12, ov
    # a multiplexer on the same 6 targets with a control outside them
    b.apply_multiplexer([7], qs, np.vstack([np.eye(64), M]))
    a.apply_mcx([7, 29])
    assert abs(a.inner_product(b) - 1) < 1e-12
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["double", "single"])
def test_wide_pauli_with_odd_y_count(precision):
    """The split form of a Pauli string wider than a tile takes its sign from
    the index each amplitude moves to, as the reference does: an odd number of
    Y letters exposes a sign error."""
    n = 15
    print(f"case: wide Pauli, odd Y count {precision}", flush=True)
    e, o, rng = pair(n, 14, precision)
    qs = list(range(n))
    for lab in ("Y" * n, "Y" * 3 + "X" * 12):
        e.apply_pauli(qs, lab, 1j)
        o.apply_pauli(qs, lab, 1j)
        close(e, o, precision)


def test_probe_not_moved_across_non_unitary_ops():
    """A saved value queued before an mcphase with |phase| != 1 (or a
    projector) reads the state before it."""
    n = 14
    print("case: probes around non-unitary ops", flush=True)
    e, o, rng = pair(n, 15)
    labels = ["ZX", "XY", "IZ"]
    want = []
    for lab in labels:
        e.save_expval_pauli([1, 3], lab)
        want.append(o.expval_pauli([1, 3], lab))
    e.apply_mcphase([9, 12], 0.5)
    o.apply_mcphase([9, 12], 0.5)
    e.apply_diagonal_matrix([6], [1.0, 0.0])
    o.apply_diagonal_matrix([6], [1.0, 0.0])
    for lab in labels:
        e.save_expval_pauli([1, 3], lab)
        want.append(o.expval_pauli([1, 3], lab))
    m = D.random_unitary(2, rng)
    e.apply_matrix([1], m)
    o.apply_matrix([1], m)
    got = e.saved_expvals()
    assert np.max(np.abs(np.array(got) - np.array(want))) < 1e-12
    close(e, o)


def test_large_measure_reset_initialize_vs_oracle():
    """Mid-circuit measurement of 17 qubits, reset of 12 and initialize of 7
    qubits on a 20-qubit state (projectors beyond one diagonal, X moves for the
    reset, the two-sweep initialize_component)."""
    from qiskit_aer_amd import AerSimulator
    n = 20
    print("case: 20q measure 17 / reset 12 / initialize 7 vs oracle", flush=True)
    rng = np.random.default_rng(16)
    c = circuits_layer(n, rng)
    mq = [int(x) for x in rng.choice(n, 17, replace=False)]
    c.measure(mq, list(range(17)))
    c.gate("h", [int(mq[0])])
    c.reset([int(x) for x in rng.choice(n, 12, replace=False)])
    for op in circuits_layer(n, rng).ops:
        c.ops.append(op)
    c.initialize(D.random_state(7, rng), [int(x) for x in rng.choice(n, 7, replace=False)])
    c.save_statevector()
    res = AerSimulator(precision="double").run(c, shots=1, seed_simulator=11)
    mem_o, sv_o, _ = O.run_circuit(c, shots=1, seed=11)
    assert int(res.results[0].memory[0]) == int(mem_o[0])
    assert np.max(np.abs(res.get_statevector() - sv_o)) <= 1e-12


def circuits_layer(n, rng):
    from qiskit_aer_amd.circuit import Circuit
    c = Circuit(n, num_clbits=17)
    for q in range(n):
        c.gate("u3", [q], list(rng.uniform(0, 6.28, 3)))
    for a in range(0, n - 1, 2):
        c.gate("cx", [a, a + 1])
    return c


@pytest.mark.parametrize("precision", ["double", "single"])
def test_norm_of_matrix_and_diagonal_and_amplitude_access(precision):
    """norm(qubits, mat) = |M psi|^2 and norm_diagonal against a dense
    brute force (Kraus probabilities, statevector_state.hpp:1136);
    get_state / set_state on single amplitudes."""
    n = 14
    print(f"case: norm(qubits, mat) {precision}", flush=True)
    e, o, rng = pair(n, 17, precision)
    psi = e.copy_to_vector()
    tol = 1e-12 if precision == "double" else 1e-5
    for k in (1, 2, 3, 6):
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        m = D.random_unitary(1 << k, rng) * rng.uniform(0.2, 0.9)
        want = np.linalg.norm(D.apply(psi, m, qs)) ** 2
        assert abs(e.norm_matrix(qs, m) - want) < tol, k
        d = rng.standard_normal(1 << k) + 1j * rng.standard_normal(1 << k)
        want = np.linalg.norm(D.apply(psi, np.diag(d), qs)) ** 2
        assert abs(e.norm_diagonal(qs, d) - want) < 10 * tol, k
    for pos in (0, 77, (1 << n) - 1):
        assert abs(e.get_state(pos) - psi[pos]) < tol
    e.set_state(123, 0.25 - 0.5j)
    psi[123] = 0.25 - 0.5j
    assert np.max(np.abs(e.copy_to_vector() - psi)) < tol


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("max_qubit", [64, 3])
@pytest.mark.parametrize("more_registers", [False, True])
def test_fused_diagonal_runs_match_oracle(precision, max_qubit, more_registers):
    """Runs of phase gates (cp / cz / p on pairs and single qubits), RZ, S, T
    and single-qubit projectors between 2x2 gates and saved values, with
    qubits on register bits, thread bits and qubits fixed by the tile: the
    engine fuses them into table-driven diagonal ops (or, with max_qubit 3,
    into short ones) and must give the oracle's state and saved values; also
    with one more register bit (128-thread workgroups staging the tables)."""
    n = 17
    print(f"case: fused diagonals {precision} max_qubit={max_qubit} more_registers={more_registers}",
          flush=True)
    e, o, rng = pair(n, 18, precision)
    if more_registers:
        e.set_register_bits(4 if precision == "double" else 5)
    e.set_fusion(True, max_qubit, 0)
    want = []
    for rnd in range(6):
        for q in rng.choice(n, 6, replace=False):
            m = D.random_unitary(2, rng)
            e.apply_matrix([int(q)], m); o.apply_matrix([int(q)], m)
        for _ in range(40):
            kind = rng.integers(0, 6)
            if kind <= 1:
                a, b = (int(x) for x in rng.choice(n, 2, replace=False))
                ph = np.exp(1j * rng.uniform(-3, 3)) if kind == 0 else -1
                e.apply_mcphase([a, b], ph); o.apply_mcphase([a, b], ph)
            elif kind == 2:
                q = int(rng.integers(n))
                ph = np.exp(1j * rng.uniform(-3, 3))
                e.apply_mcphase([q], ph); o.apply_mcphase([q], ph)
            elif kind == 3:
                q = int(rng.integers(n))
                t = rng.uniform(-3, 3)
                e.apply_rotation([q], "z", t); o.apply_rotation([q], "z", t)
            elif kind == 4:
                q = int(rng.integers(n))
                d = [1, np.exp(1j * rng.choice([np.pi / 2, np.pi / 4, -np.pi / 4]))]
                e.apply_diagonal_matrix([q], d); o.apply_diagonal_matrix([q], d)
            else:
                j = int(rng.integers(4, n))
                for k in rng.choice(j, min(j, 9), replace=False):
                    th = rng.uniform(-3, 3)
                    e.apply_mcphase([int(k), j], np.exp(1j * th))
                    o.apply_mcphase([int(k), j], np.exp(1j * th))
        if rnd == 3:
            q = int(rng.integers(n))
            e.apply_diagonal_matrix([q], [1.0, 0.0]); o.apply_diagonal_matrix([q], [1.0, 0.0])
        lab = "".join(rng.choice(list("IXYZ"), 4))
        qs = [int(x) for x in rng.choice(n, 4, replace=False)]
        e.save_expval_pauli(qs, lab)
        want.append(o.expval_pauli(qs, lab))
    got = e.saved_expvals()
    close(e, o, precision)
    assert np.max(np.abs(np.array(got) - np.array(want))) < 10 * TOL[precision]
    assert e.fused_count() > 0


@pytest.mark.parametrize("precision", ["double", "single"])
def test_qft_fused_vs_unfused_vs_dft(precision):
    """QFT-18 with fusion on and off: both equal sqrt(N) ifft(psi)."""
    from qiskit_aer_amd import circuits
    n = 18
    print(f"case: QFT fused/unfused {precision}", flush=True)
    rng = np.random.default_rng(19)
    psi = D.random_state(n, rng)
    want = np.sqrt(1 << n) * np.fft.ifft(psi)
    tol = 1e-12 if precision == "double" else 1e-5
    for fuse in (True, False):
        e = engine(n, precision)
        e.set_fusion(fuse)
        e.initialize_from_vector(psi)
        e.apply_ops(circuits.qft(n))
        assert np.max(np.abs(e.copy_to_vector() - want)) < tol, fuse
        assert (e.fused_count() > 0) == fuse


def swap_layer(n, k, rng):
    qs = [int(x) for x in rng.permutation(n)[:2 * k]]
    return [(qs[2 * i], qs[2 * i + 1]) for i in range(k)]


@pytest.mark.parametrize("precision", ["double", "single"])
@pytest.mark.parametrize("n", [9, 13, 20])
def test_swap_layers_as_permutation_passes(precision, n):
    """Layers of SWAPs on disjoint pairs (merged into one bit-permutation pass
    when they span more qubits than a tile pass holds): bit-exact against the
    oracle's SWAPs one by one, alone and between gates; the reversal of all
    qubits (the QFT's final layer) included."""
    print(f"case: swap layers n={n} {precision}", flush=True)
    e, o, rng = pair(n, 40 + n, precision)
    if precision == "single":
        o.initialize_from_vector(e.copy_to_vector().astype(np.complex128))
    layers = [swap_layer(n, k, rng) for k in (2, min(16, n // 2), min(7, n // 2))]
    layers.append([(q, n - 1 - q) for q in range(n // 2)])
    for layer in layers:
        for a, b in layer:
            e.apply_mcswap([a, b])
            o.apply_mcswap([a, b])
    assert np.array_equal(e.copy_to_vector(), o.copy_to_vector().astype(e.copy_to_vector().dtype))
    # between gates, with controlled swaps (never merged) in the run
    for layer in layers:
        q = int(rng.integers(n))
        m = D.u3(*rng.uniform(0, 2 * np.pi, 3))
        e.apply_mcu([q], m)
        o.apply_mcu([q], m)
        for i, (a, b) in enumerate(layer):
            e.apply_mcswap([a, b])
            o.apply_mcswap([a, b])
            if i == 1:
                c = [x for x in range(n) if x not in (a, b)][0]
                e.apply_mcswap([c, a, b])
                o.apply_mcswap([c, a, b])
    close(e, o, precision)


@pytest.mark.parametrize("tile_bits,low_bits", [(6, 2), (8, 3), (9, 1), (11, 3)])
def test_swap_layers_under_tile_settings(tile_bits, low_bits):
    """Bit-permutation passes under other tile sizes and low-qubit counts (the
    permutation tile holds at most tile_bits - 1 qubits), between gates."""
    n = 14
    print(f"case: swap layers tile={tile_bits} low={low_bits}", flush=True)
    e, o, rng = pair(n, 60 + tile_bits)
    e.set_tile_bits(tile_bits, low_bits)
    for layer in range(3):
        for q in range(n):
            m = D.u3(*rng.uniform(0, 2 * np.pi, 3))
            e.apply_mcu([q], m)
            o.apply_mcu([q], m)
        for a, b in swap_layer(n, 7 - layer, rng):
            e.apply_mcswap([a, b])
            o.apply_mcswap([a, b])
    close(e, o)
