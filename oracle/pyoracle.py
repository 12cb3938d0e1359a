"""ORACLE -- test infrastructure only.

ctypes binding of the CPU restatement (oracle/_build/libaer_oracle.so) and of
the reference's own compilable parts (oracle/_ref/libaer_ref.so).  Imported
by tests/, __graft_entry__.smoke() and bench.py's CPU baseline only.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ORACLE_SO = HERE / "_build" / "libaer_oracle.so"
REF_SO = HERE / "_ref" / "libaer_ref.so"

_u64p = C.POINTER(C.c_uint64)
_oracle = None
_ref = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def build_ref() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE / "ref")], check=True)


def lib():
    global _oracle
    if _oracle is None:
        if not ORACLE_SO.exists():
            build()
        L = C.CDLL(str(ORACLE_SO))
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_qv_create.restype = C.c_void_p
        L.oracle_qv_create.argtypes = [C.c_uint64, C.c_int, C.c_int]
        L.oracle_qv_destroy.argtypes = [C.c_void_p]
        for name in ["oracle_qv_initialize"]:
            getattr(L, name).argtypes = [C.c_void_p]
        L.oracle_qv_initialize_from_vector.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        L.oracle_qv_copy_to_vector.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]
        L.oracle_qv_set_sample_measure_index_size.argtypes = [C.c_void_p, C.c_int]
        L.oracle_qv_initialize_component.argtypes = [C.c_void_p, _u64p, C.c_uint64, C.c_void_p]
        L.oracle_apply_matrix.argtypes = [C.c_void_p, _u64p, C.c_uint64, C.c_void_p]
        L.oracle_apply_diagonal_matrix.argtypes = [C.c_void_p, _u64p, C.c_uint64, C.c_void_p]
        for name in ["oracle_apply_mcx", "oracle_apply_mcy", "oracle_apply_mcswap",
                     "oracle_apply_multi_swaps"]:
            getattr(L, name).argtypes = [C.c_void_p, _u64p, C.c_uint64]
        L.oracle_apply_mcphase.argtypes = [C.c_void_p, _u64p, C.c_uint64, C.c_double, C.c_double]
        L.oracle_apply_mcu.argtypes = [C.c_void_p, _u64p, C.c_uint64, C.c_void_p]
        L.oracle_apply_rotation.argtypes = [C.c_void_p, _u64p, C.c_uint64, C.c_int, C.c_double]
        L.oracle_apply_pauli.argtypes = [C.c_void_p, _u64p, C.c_uint64, C.c_char_p, C.c_double, C.c_double]
        L.oracle_apply_permutation_matrix.argtypes = [C.c_void_p, _u64p, C.c_uint64, _u64p, C.c_uint64]
        L.oracle_apply_multiplexer.argtypes = [C.c_void_p, _u64p, C.c_uint64, _u64p, C.c_uint64, C.c_void_p]
        L.oracle_norm.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        L.oracle_probabilities.argtypes = [C.c_void_p, _u64p, C.c_uint64, C.c_void_p]
        L.oracle_sample_measure.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
        L.oracle_expval_pauli.argtypes = [C.c_void_p, _u64p, C.c_uint64, C.c_char_p, C.c_double,
                                          C.c_double, C.POINTER(C.c_double)]
        L.oracle_run_circuit.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_void_p,
                                         C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_fuse_count.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                        _u64p, C.c_void_p, C.c_uint64]
        _oracle = L
    return _oracle


def ref():
    """The reference's own transformer / VMatrix / RNG (None if not built)."""
    global _ref
    if _ref is None:
        if not REF_SO.exists():
            return None
        L = C.CDLL(str(REF_SO))
        L.ref_apply_matrix.argtypes = [C.c_void_p, C.c_uint64, _u64p, C.c_uint64, C.c_void_p, C.c_int]
        L.ref_apply_diagonal_matrix.argtypes = [C.c_void_p, C.c_uint64, _u64p, C.c_uint64, C.c_void_p, C.c_int]
        L.ref_indexes.argtypes = [_u64p, C.c_uint64, C.c_uint64, C.c_void_p]
        L.ref_vmatrix.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p]
        L.ref_rng_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_void_p]
        L.ref_rng_discrete.argtypes = [C.c_uint64, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p]
        L.ref_avx2_supported.restype = C.c_int
        _ref = L
    return _ref


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().oracle_last_error().decode())


def _q(qubits):
    a = np.ascontiguousarray(np.asarray(qubits, dtype=np.uint64).reshape(-1))
    return a, a.ctypes.data_as(_u64p), a.size


def _c(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.complex128).reshape(-1))


def _colmajor(m):
    m = np.asarray(m, dtype=np.complex128)
    return np.ascontiguousarray(m.reshape(-1, order="F")) if m.ndim == 2 else _c(m)


_ROT = {"x": 0, "y": 1, "z": 2, "xx": 3, "yy": 4, "zz": 5, "zx": 6}


class OracleQubitVector:
    """Same method set as qiskit_aer_amd.QubitVector, on the CPU oracle."""

    def __init__(self, num_qubits, precision="double", threads=0):
        import os
        th = threads or os.cpu_count() or 1
        self._L = lib()
        self._h = C.c_void_p(self._L.oracle_qv_create(num_qubits, 0 if precision == "double" else 1, th))
        self._n = num_qubits

    def __del__(self):
        try:
            self._L.oracle_qv_destroy(self._h)
        except Exception:
            pass

    def size(self):
        return 1 << self._n

    def initialize(self):
        _check(self._L.oracle_qv_initialize(self._h))

    def initialize_from_vector(self, v):
        v = _c(v)
        _check(self._L.oracle_qv_initialize_from_vector(self._h, v.ctypes.data, v.size))

    def initialize_component(self, qubits, state):
        a, qp, n = _q(qubits)
        s = _c(state)
        _check(self._L.oracle_qv_initialize_component(self._h, qp, n, s.ctypes.data))

    def copy_to_vector(self):
        out = np.empty(self.size(), dtype=np.complex128)
        _check(self._L.oracle_qv_copy_to_vector(self._h, out.ctypes.data, out.size))
        return out

    def set_sample_measure_index_size(self, n):
        _check(self._L.oracle_qv_set_sample_measure_index_size(self._h, n))

    def apply_matrix(self, qubits, mat):
        a, qp, n = _q(qubits)
        m = _colmajor(mat)
        _check(self._L.oracle_apply_matrix(self._h, qp, n, m.ctypes.data))

    def apply_diagonal_matrix(self, qubits, diag):
        a, qp, n = _q(qubits)
        d = _c(diag)
        _check(self._L.oracle_apply_diagonal_matrix(self._h, qp, n, d.ctypes.data))

    def apply_mcx(self, qubits):
        a, qp, n = _q(qubits)
        _check(self._L.oracle_apply_mcx(self._h, qp, n))

    def apply_mcy(self, qubits):
        a, qp, n = _q(qubits)
        _check(self._L.oracle_apply_mcy(self._h, qp, n))

    def apply_mcswap(self, qubits):
        a, qp, n = _q(qubits)
        _check(self._L.oracle_apply_mcswap(self._h, qp, n))

    def apply_mcphase(self, qubits, phase):
        a, qp, n = _q(qubits)
        phase = complex(phase)
        _check(self._L.oracle_apply_mcphase(self._h, qp, n, phase.real, phase.imag))

    def apply_mcu(self, qubits, mat):
        a, qp, n = _q(qubits)
        m = _colmajor(mat)
        _check(self._L.oracle_apply_mcu(self._h, qp, n, m.ctypes.data))

    def apply_rotation(self, qubits, axis, theta):
        a, qp, n = _q(qubits)
        _check(self._L.oracle_apply_rotation(self._h, qp, n, _ROT[axis], float(theta)))

    def apply_pauli(self, qubits, pauli, coeff=1.0):
        a, qp, n = _q(qubits)
        coeff = complex(coeff)
        _check(self._L.oracle_apply_pauli(self._h, qp, n, pauli.encode(), coeff.real, coeff.imag))

    def apply_permutation_matrix(self, qubits, pairs):
        a, qp, n = _q(qubits)
        p = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint64).reshape(-1))
        _check(self._L.oracle_apply_permutation_matrix(self._h, qp, n, p.ctypes.data_as(_u64p), p.size // 2))

    def apply_multiplexer(self, controls, targets, mat):
        c, cp, nc = _q(controls)
        t, tp, nt = _q(targets)
        m = _colmajor(mat)
        _check(self._L.oracle_apply_multiplexer(self._h, cp, nc, tp, nt, m.ctypes.data))

    def apply_multi_swaps(self, qubits):
        a, qp, n = _q(qubits)
        _check(self._L.oracle_apply_multi_swaps(self._h, qp, n))

    def norm(self):
        out = C.c_double()
        _check(self._L.oracle_norm(self._h, C.byref(out)))
        return out.value

    def probabilities(self, qubits=()):
        a, qp, n = _q(qubits)
        out = np.empty(1 << (n if n else self._n))
        _check(self._L.oracle_probabilities(self._h, qp, n, out.ctypes.data))
        return out

    def sample_measure(self, rnds):
        r = np.ascontiguousarray(np.asarray(rnds, dtype=np.float64))
        out = np.empty(r.size, dtype=np.uint64)
        _check(self._L.oracle_sample_measure(self._h, r.ctypes.data, r.size, out.ctypes.data))
        return out

    def expval_pauli(self, qubits, pauli, phase=1.0):
        a, qp, n = _q(qubits)
        phase = complex(phase)
        out = C.c_double()
        _check(self._L.oracle_expval_pauli(self._h, qp, n, pauli.encode(), phase.real, phase.imag,
                                           C.byref(out)))
        return out.value


class reference_transformer:
    """Context in which the oracle's complex128 apply_matrix /
    apply_diagonal_matrix run the reference's own compiled Transformer /
    TransformerAVX2 (oracle/_ref, chosen by the reference's CPUID check) instead
    of the restatement.  Used by the timed CPU baseline."""

    def __enter__(self):
        R = ref()
        if R is None:
            raise RuntimeError("oracle/_ref is not built")
        L = lib()
        L.oracle_set_transformer.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_set_transformer(C.cast(R.ref_qv_apply_matrix, C.c_void_p),
                                 C.cast(R.ref_qv_apply_diagonal_matrix, C.c_void_p))
        self.avx2 = bool(R.ref_avx2_supported())
        return self

    def __exit__(self, *exc):
        lib().oracle_set_transformer(None, None)
        return False


def run_circuit(circ, shots=1024, seed=0, precision="double", fusion=True, threads=0,
                fusion_max_qubit=5, fusion_threshold=14, index_size=10, statevector=True):
    """Runs a qiskit_aer_amd.Circuit through the oracle's controller.

    Returns (memory array, statevector or None, timing[4])."""
    from qiskit_aer_amd import _lib as L
    cfg = L.AerhipRunConfig()
    cfg.precision = 0 if precision == "double" else 1
    cfg.shots = shots
    cfg.seed = seed
    cfg.fusion_enable = int(fusion)
    cfg.fusion_max_qubit = fusion_max_qubit
    cfg.fusion_threshold = fusion_threshold
    cfg.sample_index_size = index_size
    cfg.global_phase = circ.global_phase
    ops, keep = circ.pack()
    mem = np.zeros(max(shots, 1), dtype=np.uint64)
    sv = np.empty(1 << circ.num_qubits, dtype=np.complex128) if statevector else None
    timing = np.zeros(4)
    _check(lib().oracle_run_circuit(C.cast(ops, C.c_void_p), len(circ.ops), circ.num_qubits,
                                    circ.num_clbits, C.cast(C.pointer(cfg), C.c_void_p), threads,
                                    mem.ctypes.data, sv.ctypes.data if sv is not None else None,
                                    timing.ctypes.data))
    del keep
    return mem[:shots], sv, timing


def fused_widths(circ, max_qubit=5, threshold=14):
    """Qubit counts of Aer's fused op list for `circ` (fusion.hpp restated)."""
    ops, keep = circ.pack()
    num = C.c_uint64()
    cap = len(circ.ops) + 1
    widths = np.zeros(cap, dtype=np.uint64)
    _check(lib().oracle_fuse_count(C.cast(ops, C.c_void_p), len(circ.ops), circ.num_qubits, max_qubit,
                                   threshold, C.byref(num), widths.ctypes.data, cap))
    return widths[: num.value].tolist()
