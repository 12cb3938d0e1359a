"""Seeded circuit families of the statevector workloads.

Each builder draws its parameters from ``numpy.random.default_rng(seed)`` in
a fixed order so that the same seed gives the same circuit everywhere.
"""
from __future__ import annotations

import math

import numpy as np

from .circuit import Circuit


def brickwork_u3_cx(n: int, layers: int, seed: int, measure: bool = True) -> Circuit:
    """Layer l: U3(theta, phi, lambda) on every qubit (angles uniform in
    [0, 2pi)), then CX on (0,1),(2,3),... for even l and (1,2),(3,4),... for
    odd l."""
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for layer in range(layers):
        ang = rng.uniform(0.0, 2.0 * math.pi, size=(n, 3))
        for q in range(n):
            c.gate("u3", [q], ang[q])
        for a in range(layer % 2, n - 1, 2):
            c.gate("cx", [a, a + 1])
    if measure:
        c.measure_all()
    return c


def qft(n: int, swaps: bool = True) -> Circuit:
    """QFT: H on q then CP(pi/2^k) from the lower qubits, for q = n-1..0,
    then the final swaps (the textbook circuit of qiskit's QFT)."""
    c = Circuit(n)
    for j in reversed(range(n)):
        c.gate("h", [j])
        for k in reversed(range(j)):
            c.gate("cp", [k, j], [math.pi / float(2 ** (j - k))])
    if swaps:
        for i in range(n // 2):
            c.gate("swap", [i, n - i - 1])
    return c


def random_gaussian_state(n: int, seed: int) -> np.ndarray:
    """psi = g / |g| with g i.i.d. complex Gaussian."""
    rng = np.random.default_rng(seed)
    g = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return g / np.linalg.norm(g)


def haar_su4(rng: np.random.Generator) -> np.ndarray:
    """Haar-random SU(4): QR of a complex Gaussian 4x4, column phases fixed by
    diag(R), determinant normalised to 1."""
    z = (rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    q = q * (d / np.abs(d))
    q = q / np.linalg.det(q) ** 0.25
    return q


def su4_brickwork(n: int, layers: int, seed: int) -> Circuit:
    """Layer l: Haar SU(4) on (2i, 2i+1) for even l, (2i+1, 2i+2) for odd l."""
    rng = np.random.default_rng(seed)
    c = Circuit(n)
    for layer in range(layers):
        for a in range(layer % 2, n - 1, 2):
            c.unitary(haar_su4(rng), [a, a + 1])
    return c


def hea(n: int, reps: int, seed: int) -> tuple[Circuit, np.ndarray]:
    """Hardware-efficient ansatz: per rep RY and RZ on every qubit (angles
    uniform in [-pi, pi]) then CX(i, i+1) for i = 0..n-2.  Returns the circuit
    and its angle array."""
    rng = np.random.default_rng(seed)
    ang = rng.uniform(-math.pi, math.pi, size=(reps, n, 2))
    c = Circuit(n)
    for r in range(reps):
        for q in range(n):
            c.gate("ry", [q], [ang[r, q, 0]])
            c.gate("rz", [q], [ang[r, q, 1]])
        for q in range(n - 1):
            c.gate("cx", [q, q + 1])
    return c, ang


def random_pauli_sum(n: int, terms: int, seed: int) -> list[tuple[str, float]]:
    """`terms` Pauli labels uniform over {I,X,Y,Z}^n with N(0,1) coefficients."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(terms):
        letters = rng.integers(0, 4, size=n)
        out.append(("".join("IXYZ"[i] for i in letters), float(rng.standard_normal())))
    return out
