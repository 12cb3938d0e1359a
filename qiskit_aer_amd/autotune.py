"""Autotuner of the engine's block fusion.

``tune(circuit, precision)`` runs the circuit on a statevector of its size
under each block-fusion width (0 = off, 2..5: gates on at most that many
qubits are fused into dense matrices, applied as MFMA GEMMs on the tile for
3..5 qubits) and returns the width with the least device time of the tile
passes, with every width's time.  The reference fixes its fusion width by
option (fusion_max_qubit, src/transpile/fusion.hpp:761) and weighs fusions by
a CPU cost model (:1105-1136); here the choice is measured on the device.
"""
from __future__ import annotations

from typing import Iterable, Optional

from .circuit import Circuit


def tune(circuit: Circuit, precision: str = "double", widths: Iterable[int] = (0, 2, 3, 4, 5),
         reps: int = 1, qv=None, cost_check: bool = True) -> tuple[int, dict]:
    """Returns (best width, {width: device ms per run}).  `qv` (optional) is a
    QubitVector of the circuit's size to time on (it is re-initialised); the
    circuit must hold only unitary operations."""
    from .qubitvector import QubitVector
    own = qv is None
    if own:
        qv = QubitVector(circuit.num_qubits, precision)
    packed = circuit.pack()
    times = {}
    try:
        for w in widths:
            qv.set_block_fusion(int(w), cost_check)
            qv.initialize()
            qv.apply_ops(packed)  # warm-up (plans, uploads, caches the launch sequence)
            qv.synchronize()
            qv.set_event_timing(True)
            ms0, _ = qv.pass_timing()
            for _ in range(max(1, reps)):
                qv.initialize()
                qv.apply_ops(packed)
                qv.synchronize()
            ms1, _ = qv.pass_timing()
            qv.set_event_timing(False)
            times[int(w)] = (ms1 - ms0) / max(1, reps)
    finally:
        qv.set_block_fusion(0)
        if own:
            qv.close()
    best = min(times, key=times.get)
    return best, times


def tuned_width(circuit: Circuit, precision: str = "double", proxy_qubits: Optional[int] = None,
                **kw) -> int:
    """The best width for `circuit`, timed on the circuit itself or, when
    proxy_qubits is given and smaller, on the same gates restricted to the
    qubits below proxy_qubits (a cheaper stand-in for uniform circuits such
    as brickworks)."""
    c = circuit
    if proxy_qubits and proxy_qubits < circuit.num_qubits:
        c = Circuit(proxy_qubits)
        c.ops = [op for op in circuit.ops if all(q < proxy_qubits for q in op.qubits)]
    return tune(c, precision, **kw)[0]
