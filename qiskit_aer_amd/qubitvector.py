"""Python mirror of the reference's ``QubitVector`` method family.

``QubitVector(num_qubits, precision)`` holds a statevector in HIP device
memory and exposes the methods the reference's ``Statevector::State`` calls
on ``QV::QubitVector`` (src/simulators/statevector/qubitvector.hpp:61-651):
apply_matrix, apply_diagonal_matrix, apply_mcx/mcy/mcswap/mcphase/mcu,
apply_rotation, apply_pauli, norm, probabilities, sample_measure,
expval_pauli.  Conventions (little-endian qubits, column-major matrices,
Pauli character N-1-i on qubits[i]) are the reference's.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib as L

_ROT = {"x": L.ROT_X, "y": L.ROT_Y, "z": L.ROT_Z, "xx": L.ROT_XX, "yy": L.ROT_YY,
        "zz": L.ROT_ZZ, "zx": L.ROT_ZX}


def _q(qubits: Sequence[int]):
    a = np.ascontiguousarray(np.asarray(qubits, dtype=np.uint64).reshape(-1))
    return a, a.ctypes.data_as(C.POINTER(C.c_uint64)), a.size


def _c(values) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(values, dtype=np.complex128).reshape(-1))


def _colmajor(mat) -> np.ndarray:
    """Column-major vectorisation of a square matrix (Utils::vectorize_matrix)."""
    m = np.asarray(mat, dtype=np.complex128)
    if m.ndim == 2:
        return np.ascontiguousarray(m.reshape(-1, order="F"))
    return _c(m)


class QubitVector:
    """A statevector on one MI355X (HIP device), c128 or c64."""

    def __init__(self, num_qubits: int, precision: str = "double", device: int = 0):
        self._L = L.lib()
        prec = {"double": L.PRECISION_DOUBLE, "single": L.PRECISION_SINGLE}[precision]
        h = self._L.aerhip_qv_create(int(num_qubits), prec, int(device))
        if not h:
            raise L.AerhipError(self._L.aerhip_last_error().decode())
        self._h = C.c_void_p(h)
        self._n = int(num_qubits)
        self.precision = precision

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.aerhip_qv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # lifecycle --------------------------------------------------------------
    def num_qubits(self) -> int:
        return self._n

    def size(self) -> int:
        return 1 << self._n

    def initialize(self) -> None:
        L.check(self._L.aerhip_qv_initialize(self._h))

    def zero(self) -> None:
        L.check(self._L.aerhip_qv_zero(self._h))

    def initialize_from_vector(self, vec) -> None:
        v = _c(vec)
        L.check(self._L.aerhip_qv_initialize_from_vector(self._h, v.ctypes.data, v.size))

    def initialize_from_device(self, ptr: int, length: int, precision: str) -> None:
        p = {"double": L.PRECISION_DOUBLE, "single": L.PRECISION_SINGLE}[precision]
        L.check(self._L.aerhip_qv_initialize_from_device(self._h, C.c_void_p(ptr), length, p))

    def initialize_random(self, seed: int) -> None:
        """psi = g/|g| with i.i.d. complex Gaussian g generated on the device."""
        L.check(self._L.aerhip_qv_initialize_random(self._h, int(seed)))

    def initialize_component(self, qubits, state) -> None:
        a, qp, n = _q(qubits)
        s = _c(state)
        L.check(self._L.aerhip_qv_initialize_component(self._h, qp, n, s.ctypes.data))

    def copy_to_vector(self) -> np.ndarray:
        out = np.empty(self.size(), dtype=np.complex128)
        L.check(self._L.aerhip_qv_copy_to_vector(self._h, out.ctypes.data, out.size))
        return out

    def synchronize(self) -> None:
        L.check(self._L.aerhip_qv_synchronize(self._h))

    def set_sample_measure_index_size(self, n: int) -> None:
        L.check(self._L.aerhip_qv_set_sample_measure_index_size(self._h, int(n)))

    def set_tile_bits(self, tile_bits: int = 0, low_bits: int = 0) -> None:
        L.check(self._L.aerhip_qv_set_tile_bits(self._h, int(tile_bits), int(low_bits)))

    def device_pointer(self) -> int:
        p = C.c_void_p()
        L.check(self._L.aerhip_qv_device_pointer(self._h, C.byref(p)))
        return p.value or 0

    def set_event_timing(self, enable: bool = True) -> None:
        L.check(self._L.aerhip_qv_set_event_timing(self._h, int(bool(enable))))

    def pass_timing(self) -> tuple:
        """(device milliseconds of pass kernels, launches) while event timing was on."""
        ms = C.c_double()
        n = C.c_uint64()
        L.check(self._L.aerhip_qv_pass_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def stats(self) -> dict:
        out = (C.c_uint64 * 4)()
        L.check(self._L.aerhip_qv_stats(self._h, out))
        return {"passes": out[0], "launches": out[1], "gates": out[2], "gathers": out[3]}

    def set_fusion(self, enable: bool = True, max_qubit: int = 64, threshold: int = 0) -> None:
        """Engine-side fusion of diagonal runs (Aer's fusion_enable /
        fusion_max_qubit / fusion_threshold)."""
        L.check(self._L.aerhip_qv_set_fusion(self._h, int(bool(enable)), int(max_qubit), int(threshold)))

    def set_block_fusion(self, width: int, cost_check: bool = True) -> None:
        """Fuses runs of unitary gates on at most `width` qubits (2..5, 0: off)
        into dense matrices (GEMMs on MFMA for 3..5 qubits)."""
        L.check(self._L.aerhip_qv_set_block_fusion(self._h, int(width), int(bool(cost_check))))

    def graph_stats(self) -> tuple:
        """(flushes captured into a HIP graph, flushes that replayed it)."""
        c, r = C.c_uint64(), C.c_uint64()
        L.check(self._L.aerhip_qv_graph_stats(self._h, C.byref(c), C.byref(r)))
        return int(c.value), int(r.value)

    def fused_count(self) -> int:
        out = C.c_uint64()
        L.check(self._L.aerhip_qv_fused_count(self._h, C.byref(out)))
        return out.value

    def set_register_bits(self, reg_bits: int) -> None:
        L.check(self._L.aerhip_qv_set_register_bits(self._h, int(reg_bits)))

    # gates ------------------------------------------------------------------
    def apply_matrix(self, qubits, mat) -> None:
        a, qp, n = _q(qubits)
        m = _colmajor(mat)
        L.check(self._L.aerhip_apply_matrix(self._h, qp, n, m.ctypes.data))

    def apply_diagonal_matrix(self, qubits, diag) -> None:
        a, qp, n = _q(qubits)
        d = _c(diag)
        L.check(self._L.aerhip_apply_diagonal_matrix(self._h, qp, n, d.ctypes.data))

    def apply_mcx(self, qubits) -> None:
        a, qp, n = _q(qubits)
        L.check(self._L.aerhip_apply_mcx(self._h, qp, n))

    def apply_mcy(self, qubits) -> None:
        a, qp, n = _q(qubits)
        L.check(self._L.aerhip_apply_mcy(self._h, qp, n))

    def apply_mcswap(self, qubits) -> None:
        a, qp, n = _q(qubits)
        L.check(self._L.aerhip_apply_mcswap(self._h, qp, n))

    def apply_mcphase(self, qubits, phase: complex) -> None:
        a, qp, n = _q(qubits)
        phase = complex(phase)
        L.check(self._L.aerhip_apply_mcphase(self._h, qp, n, phase.real, phase.imag))

    def apply_mcu(self, qubits, mat) -> None:
        a, qp, n = _q(qubits)
        m = _colmajor(mat)
        L.check(self._L.aerhip_apply_mcu(self._h, qp, n, m.ctypes.data))

    def apply_rotation(self, qubits, axis: str, theta: float) -> None:
        a, qp, n = _q(qubits)
        L.check(self._L.aerhip_apply_rotation(self._h, qp, n, _ROT[axis], float(theta)))

    def apply_pauli(self, qubits, pauli: str, coeff: complex = 1.0) -> None:
        a, qp, n = _q(qubits)
        coeff = complex(coeff)
        L.check(self._L.aerhip_apply_pauli(self._h, qp, n, pauli.encode(), coeff.real, coeff.imag))

    def apply_permutation_matrix(self, qubits, pairs) -> None:
        a, qp, n = _q(qubits)
        p = np.ascontiguousarray(np.asarray(pairs, dtype=np.uint64).reshape(-1))
        L.check(self._L.aerhip_apply_permutation_matrix(
            self._h, qp, n, p.ctypes.data_as(C.POINTER(C.c_uint64)), p.size // 2))

    def apply_multiplexer(self, controls, targets, mat) -> None:
        c, cp, nc = _q(controls)
        t, tp, nt = _q(targets)
        m = _colmajor(mat)
        L.check(self._L.aerhip_apply_multiplexer(self._h, cp, nc, tp, nt, m.ctypes.data))

    def apply_multi_swaps(self, qubits) -> None:
        a, qp, n = _q(qubits)
        L.check(self._L.aerhip_apply_multi_swaps(self._h, qp, n))

    def apply_ops(self, circuit) -> None:
        """Applies the unitary operations of a Circuit (State::apply_ops)."""
        ops, keep = circuit.pack() if not isinstance(circuit, tuple) else circuit
        L.check(self._L.aerhip_apply_ops(self._h, ops, len(ops)))
        del keep

    # readout ----------------------------------------------------------------
    def norm(self) -> float:
        out = C.c_double()
        L.check(self._L.aerhip_norm(self._h, C.byref(out)))
        return out.value

    def norm_matrix(self, qubits, mat) -> float:
        """|M psi|^2 for M on `qubits`, not applied (QubitVector::norm(qubits, mat))."""
        a, qp, n = _q(qubits)
        m = _colmajor(mat)
        out = C.c_double()
        L.check(self._L.aerhip_norm_matrix(self._h, qp, n, m.ctypes.data_as(C.POINTER(C.c_double)),
                                           C.byref(out)))
        return out.value

    def norm_diagonal(self, qubits, diag) -> float:
        """|D psi|^2 for the diagonal D on `qubits` (QubitVector::norm_diagonal)."""
        a, qp, n = _q(qubits)
        d = _c(diag)
        out = C.c_double()
        L.check(self._L.aerhip_norm_diagonal(self._h, qp, n, d.ctypes.data_as(C.POINTER(C.c_double)),
                                             C.byref(out)))
        return out.value

    def get_state(self, pos: int) -> complex:
        out = np.zeros(2)
        L.check(self._L.aerhip_qv_get_state(self._h, int(pos), out.ctypes.data_as(C.POINTER(C.c_double))))
        return complex(out[0], out[1])

    def set_state(self, pos: int, value: complex) -> None:
        value = complex(value)
        L.check(self._L.aerhip_qv_set_state(self._h, int(pos), value.real, value.imag))

    def inner_product(self, other: "QubitVector") -> complex:
        """<self|other> accumulated in FP64 (the states may differ in precision)."""
        re, im = C.c_double(), C.c_double()
        L.check(self._L.aerhip_qv_inner_product(self._h, other._h, C.byref(re), C.byref(im)))
        return complex(re.value, im.value)

    def probability(self, outcome: int) -> float:
        out = C.c_double()
        L.check(self._L.aerhip_probability(self._h, int(outcome), C.byref(out)))
        return out.value

    def probabilities(self, qubits=()) -> np.ndarray:
        a, qp, n = _q(qubits)
        out = np.empty(1 << (n if n else self._n), dtype=np.float64)
        L.check(self._L.aerhip_probabilities(self._h, qp, n, out.ctypes.data))
        return out

    def sample_measure(self, rnds) -> np.ndarray:
        r = np.ascontiguousarray(np.asarray(rnds, dtype=np.float64).reshape(-1))
        out = np.empty(r.size, dtype=np.uint64)
        L.check(self._L.aerhip_sample_measure(self._h, r.ctypes.data, r.size, out.ctypes.data))
        return out

    def save_expval_pauli(self, qubits, pauli: str, phase: complex = 1.0) -> int:
        """Queues Re<psi|phase P|psi> at this point of the gate sequence; returns its slot."""
        a, qp, n = _q(qubits)
        phase = complex(phase)
        slot = C.c_uint64()
        L.check(self._L.aerhip_save_expval_pauli(self._h, qp, n, pauli.encode(), phase.real,
                                                 phase.imag, C.byref(slot)))
        return slot.value

    def saved_expvals(self, capacity: int = 1 << 16) -> np.ndarray:
        """Flushes and returns every saved value in slot order (then clears them)."""
        out = np.empty(capacity, dtype=np.float64)
        cnt = C.c_uint64()
        L.check(self._L.aerhip_saved_expvals(self._h, out.ctypes.data, capacity, C.byref(cnt)))
        return out[: cnt.value].copy()

    def expval_pauli_multi(self, qubits, labels) -> np.ndarray:
        """<P_t> for many Pauli strings on the same qubits, evaluated in one sweep."""
        for lab in labels:
            self.save_expval_pauli(qubits, lab)
        return self.saved_expvals(max(1, len(labels)))

    def expval_pauli(self, qubits, pauli: str, phase: complex = 1.0) -> float:
        a, qp, n = _q(qubits)
        phase = complex(phase)
        out = C.c_double()
        L.check(self._L.aerhip_expval_pauli(self._h, qp, n, pauli.encode(), phase.real,
                                            phase.imag, C.byref(out)))
        return out.value


def device_memory(device: int = 0) -> tuple:
    """(free, total) bytes of device memory."""
    free, total = C.c_uint64(), C.c_uint64()
    L.check(L.lib().aerhip_device_memory(int(device), C.byref(free), C.byref(total)))
    return free.value, total.value


def device_count() -> int:
    """Number of visible HIP devices."""
    return int(L.lib().aerhip_device_count())
