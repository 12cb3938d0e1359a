"""Independent dense-matrix statevector helpers for brute-force checks.

They build every operator with numpy tensor products from textbook gate
definitions, sharing no code with the engine or the oracle.
"""
from __future__ import annotations

import math

import numpy as np


def apply(psi: np.ndarray, mat: np.ndarray, qubits) -> np.ndarray:
    """psi' = (mat on qubits) psi; bit j of the matrix index is qubits[j]."""
    n = int(round(math.log2(psi.size)))
    k = len(qubits)
    t = psi.reshape([2] * n)                      # axis a <-> qubit n-1-a
    m = np.asarray(mat, dtype=complex).reshape([2] * (2 * k))  # rows: q[k-1..0], cols: q[k-1..0]
    col_axes = list(range(k, 2 * k))
    st_axes = [n - 1 - qubits[k - 1 - i] for i in range(k)]
    out = np.tensordot(m, t, axes=(col_axes, st_axes))  # result axes: rows..., remaining state axes
    remaining = [a for a in range(n) if a not in st_axes]
    order = st_axes + remaining
    inv = np.argsort(order)
    return np.transpose(out, inv).reshape(-1)


def controlled(mat: np.ndarray, nctrl: int) -> np.ndarray:
    """Matrix of `mat` controlled on nctrl lower-index qubits (all ones)."""
    d = mat.shape[0]
    D = d << nctrl
    full = np.eye(D, dtype=complex)
    ones = (1 << nctrl) - 1
    idx = [ones | (i << nctrl) for i in range(d)]
    for a in range(d):
        for b in range(d):
            full[idx[a], idx[b]] = mat[a, b]
    return full


I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
H = np.array([[1, 1], [1, -1]], dtype=complex) / math.sqrt(2)


def u3(t, p, l):
    return np.array([[math.cos(t / 2), -np.exp(1j * l) * math.sin(t / 2)],
                     [np.exp(1j * p) * math.sin(t / 2), np.exp(1j * (p + l)) * math.cos(t / 2)]])


def ry(t):
    return np.array([[math.cos(t / 2), -math.sin(t / 2)], [math.sin(t / 2), math.cos(t / 2)]], dtype=complex)


def rz(t):
    return np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])


def rx(t):
    return np.array([[math.cos(t / 2), -1j * math.sin(t / 2)], [-1j * math.sin(t / 2), math.cos(t / 2)]])


def pauli_matrix(label: str) -> np.ndarray:
    """Matrix of a Pauli label; character N-1-i acts on local qubit i."""
    mats = {"I": I2, "X": X, "Y": Y, "Z": Z}
    out = np.array([[1.0 + 0j]])
    for ch in label:  # leftmost = highest local qubit
        out = np.kron(out, mats[ch])
    return out


def random_unitary(d: int, rng: np.random.Generator) -> np.ndarray:
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_state(n: int, rng: np.random.Generator) -> np.ndarray:
    g = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return g / np.linalg.norm(g)
