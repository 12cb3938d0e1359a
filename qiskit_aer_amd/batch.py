"""Many circuits with Pauli-sum expectation values in one launch.

``batch_pauli_expval(circuits, hamiltonians)`` returns, for each circuit
(run from |0> on its own qubit count), E_i = sum_t c_t <P_t>.  With
``mode="batch"`` every circuit runs in a single kernel launch (one workgroup
per circuit); ``mode="sequential"`` runs them one after the other with one
launch per pass, the way the reference's controller executes a list of
circuits (src/controllers/aer_controller.hpp:309-437).
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _lib as L
from .circuit import Circuit


def batch_pauli_expval(circuits: Sequence[Circuit], hamiltonians: Sequence[Sequence[tuple]],
                       precision: str = "double", mode: str = "batch",
                       device: int = 0) -> tuple[np.ndarray, float]:
    """Returns (energies, kernel_ms); kernel_ms is 0 in sequential mode."""
    if len(circuits) != len(hamiltonians):
        raise ValueError("one Pauli sum per circuit")
    nc = len(circuits)
    all_ops = [op for c in circuits for op in c.ops]
    merged = Circuit(max([c.num_qubits for c in circuits] + [1]))
    merged.ops = all_ops
    arr, keep = merged.pack()
    op_off = np.zeros(nc + 1, dtype=np.uint64)
    term_off = np.zeros(nc + 1, dtype=np.uint64)
    nq = np.ascontiguousarray(np.asarray([c.num_qubits for c in circuits], dtype=np.uint32))
    labels, coeffs = [], []
    for i, (c, h) in enumerate(zip(circuits, hamiltonians)):
        op_off[i + 1] = op_off[i] + len(c.ops)
        term_off[i + 1] = term_off[i] + len(h)
        for lab, coef in h:
            if len(lab) != c.num_qubits:
                raise ValueError("every label needs one character per qubit of its circuit")
            labels.append(lab)
            coeffs.append(float(coef))
    lab_bytes = "".join(labels).encode()
    coeffs = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
    energies = np.zeros(nc, dtype=np.float64)
    ms = C.c_double(0)
    prec = L.PRECISION_DOUBLE if precision == "double" else L.PRECISION_SINGLE
    L.check(L.lib().aerhip_batch_pauli_expval(
        arr, op_off.ctypes.data_as(C.POINTER(C.c_uint64)), nq.ctypes.data_as(C.POINTER(C.c_uint32)),
        nc, lab_bytes, coeffs.ctypes.data_as(C.POINTER(C.c_double)),
        term_off.ctypes.data_as(C.POINTER(C.c_uint64)), prec, int(device),
        0 if mode == "batch" else 1, energies.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ms)))
    del keep
    return energies, ms.value
