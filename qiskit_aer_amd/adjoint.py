"""Adjoint-method gradients of Pauli-sum expectation values, and a
``torch.autograd.Function`` around them.

``AdjointEstimator(circuit, param_index, num_params, hamiltonian)`` evaluates
E(theta) = <psi(theta)|H|psi(theta)> and dE/dtheta for a batch of parameter
rows on the MI355X engine (``aerhip_adjoint_expval_gradient``): one forward
sweep and one reverse sweep that uncomputes the gates on psi and on
lambda = H psi, with the gradient inner products evaluated inside the
engine's tile passes.  ``parameter_shift`` computes the same gradient the way
the reference does, by re-running the circuit at theta +- pi/2
(qiskit_aer/primitives/estimator_v2.py:112-159 driven per shifted binding).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Sequence

import numpy as np

from . import _lib as L
from .circuit import Circuit


class AdjointEstimator:
    def __init__(self, circuit: Circuit, param_index: Sequence[int], num_params: int,
                 hamiltonian: Sequence[tuple], precision: str = "double", device: int = 0):
        if len(param_index) != len(circuit.ops):
            raise ValueError("param_index needs one entry per operation")
        self.circuit = circuit
        self.num_qubits = circuit.num_qubits
        self.param_index = np.ascontiguousarray(np.asarray(param_index, dtype=np.int64))
        self.num_params = int(num_params)
        labels = [lab for lab, _ in hamiltonian]
        if any(len(lab) != self.num_qubits for lab in labels):
            raise ValueError("every Hamiltonian label must have one character per qubit")
        self.labels = "".join(labels).encode()
        self.coeffs = np.ascontiguousarray(np.asarray([c for _, c in hamiltonian], dtype=np.float64))
        self.precision = L.PRECISION_DOUBLE if precision == "double" else L.PRECISION_SINGLE
        self.device = int(device)

    def energy_and_gradient(self, params) -> tuple[np.ndarray, np.ndarray]:
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1, self.num_params))
        nb = p.shape[0]
        energy = np.empty(nb, dtype=np.float64)
        grad = np.empty((nb, self.num_params), dtype=np.float64)
        ops, keep = self.circuit.pack()
        L.check(L.lib().aerhip_adjoint_expval_gradient(
            ops, len(self.circuit.ops), self.num_qubits, self.param_index.ctypes.data,
            self.num_params, nb, p.ctypes.data, self.labels, self.coeffs.ctypes.data,
            self.coeffs.size, self.precision, self.device, energy.ctypes.data, grad.ctypes.data))
        del keep
        return energy, grad

    def energy(self, params, kernel_ms: list | None = None) -> np.ndarray:
        """E(theta) for every parameter row, the rows advancing through the
        circuit together (``aerhip_bind_pauli_expval``: one kernel launch per
        tile pass for all rows of a launch group)."""
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1, self.num_params))
        out = np.empty(p.shape[0], dtype=np.float64)
        ms = C.c_double(0.0)
        ops, keep = self.circuit.pack()
        L.check(L.lib().aerhip_bind_pauli_expval(
            ops, len(self.circuit.ops), self.num_qubits, self.param_index.ctypes.data,
            self.num_params, p.shape[0], p.ctypes.data, self.labels, self.coeffs.ctypes.data,
            self.coeffs.size, self.precision, self.device, out.ctypes.data, C.byref(ms)))
        del keep
        if kernel_ms is not None:
            kernel_ms.append(ms.value)
        return out

    def energy_one_by_one(self, params) -> np.ndarray:
        """E(theta) rows one binding at a time on one statevector (forward run
        and saved expectation values per row)."""
        from .qubitvector import QubitVector
        p = np.asarray(params, dtype=np.float64).reshape(-1, self.num_params)
        qv = QubitVector(self.num_qubits, "double" if self.precision == 0 else "single", self.device)
        labels = [self.labels[i * self.num_qubits:(i + 1) * self.num_qubits].decode()
                  for i in range(self.coeffs.size)]
        out = np.empty(p.shape[0])
        try:
            for b in range(p.shape[0]):
                qv.initialize()
                qv.apply_ops(bind(self.circuit, self.param_index, p[b]))
                vals = qv.expval_pauli_multi(list(range(self.num_qubits)), labels)
                out[b] = float(np.dot(vals, self.coeffs))
        finally:
            qv.close()
        return out

    def parameter_shift(self, params, which: Sequence[int] | None = None) -> np.ndarray:
        """dE/dtheta_j = (E(theta + pi/2 e_j) - E(theta - pi/2 e_j)) / 2 for the
        parameters in `which` (all by default); valid when every parameter
        drives exactly one rotation exp(-i theta G/2) with G a Pauli string."""
        p = np.asarray(params, dtype=np.float64).reshape(-1, self.num_params)
        which = list(range(self.num_params)) if which is None else list(which)
        grad = np.zeros((p.shape[0], len(which)))
        for b in range(p.shape[0]):
            rows = []
            for j in which:
                for s in (+1, -1):
                    q = p[b].copy()
                    q[j] += s * math.pi / 2
                    rows.append(q)
            e = self.energy(np.array(rows))
            grad[b] = 0.5 * (e[0::2] - e[1::2])
        return grad


def bind(circuit: Circuit, param_index, row) -> Circuit:
    """Copy of `circuit` whose parametrized ops take their first angle from `row`."""
    from .circuit import Op
    out = Circuit(circuit.num_qubits, circuit.num_clbits, circuit.global_phase)
    for op, pi in zip(circuit.ops, param_index):
        if pi >= 0:
            params = (float(row[pi]),) + tuple(op.params[1:])
            out.ops.append(Op(op.type, op.name, op.qubits, params, op.matrix, op.label, op.memory))
        else:
            out.ops.append(op)
    return out


def _torch_function():
    import torch

    class AdjointExpectation(torch.autograd.Function):
        """E(theta) with dE/dtheta from the adjoint sweep.

        params: (..., num_params) tensor on any device; returns (...) energies
        on the same device and dtype.  When no gradient is needed (no_grad,
        or params that do not require one) only the energies are computed
        (the bindings' forward runs); otherwise the adjoint sweep computes
        energy and gradient together in forward."""

        @staticmethod
        def forward(ctx, params, estimator):
            p = params.detach().to("cpu", torch.float64).numpy()
            rows = p.reshape(-1, estimator.num_params)
            if ctx.needs_input_grad[0]:
                e, g = estimator.energy_and_gradient(rows)
                ctx.save_for_backward(torch.from_numpy(g.reshape(p.shape)).to(params.device, params.dtype))
            else:
                e = estimator.energy(rows)
            return torch.from_numpy(e.reshape(p.shape[:-1])).to(params.device, params.dtype)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out.unsqueeze(-1) * g, None

    return AdjointExpectation


def adjoint_expectation(params, estimator: AdjointEstimator):
    """Differentiable E(theta) for torch tensors (autograd through the adjoint sweep)."""
    return _torch_function().apply(params, estimator)
