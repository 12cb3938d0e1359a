"""``AerSimulator``-style front end for the HIP statevector engine.

``AerSimulator(method="statevector", device="GPU").run(circuits, shots=...,
seed_simulator=...)`` follows the reference's semantics for this method
(qiskit_aer/backends/aer_simulator.py; seeds per circuit are
seed_simulator + 2113 * i, src/controllers/controller_execute.hpp:189-211)
and returns counts keyed by hex memory strings as Aer does.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence, Union

import numpy as np

from . import _lib as L
from .circuit import Circuit


@dataclass
class ExperimentResult:
    counts: dict
    memory: np.ndarray
    statevector: Optional[np.ndarray]
    seed: int
    time_taken_gates: float
    time_taken_sampling: float
    passes: int
    launches: int
    metadata: dict = field(default_factory=dict)


class Result:
    def __init__(self, results: list):
        self.results = results

    def get_counts(self, i: int = 0) -> dict:
        return self.results[i].counts

    def get_memory(self, i: int = 0) -> list:
        return [hex(int(m)) for m in self.results[i].memory]

    def get_statevector(self, i: int = 0) -> np.ndarray:
        return self.results[i].statevector


class AerSimulator:
    """Statevector simulation on the MI355X engine."""

    def __init__(self, method: str = "statevector", device: str = "GPU",
                 precision: str = "double", seed_simulator: Optional[int] = None,
                 fusion_enable: bool = True, fusion_max_qubit: int = 5,
                 fusion_threshold: int = 14, statevector_sample_measure_opt: int = 10,
                 target_gpu: int = 0):
        if method != "statevector":
            raise ValueError("only method='statevector' is provided by this engine")
        if device != "GPU":
            raise ValueError("this engine runs on device='GPU' (MI355X)")
        if precision not in ("double", "single"):
            raise ValueError("precision must be 'double' or 'single'")
        self.options = dict(precision=precision, seed_simulator=seed_simulator,
                            fusion_enable=fusion_enable, fusion_max_qubit=fusion_max_qubit,
                            fusion_threshold=fusion_threshold,
                            statevector_sample_measure_opt=statevector_sample_measure_opt,
                            target_gpu=target_gpu)

    def set_options(self, **kw) -> None:
        for k, v in kw.items():
            if k not in self.options:
                raise ValueError(f"unknown option {k}")
            self.options[k] = v

    def run(self, circuits: Union[Circuit, Sequence[Circuit]], shots: int = 1024,
            seed_simulator: Optional[int] = None, save_statevector: bool = False) -> Result:
        if isinstance(circuits, Circuit):
            circuits = [circuits]
        seed = seed_simulator if seed_simulator is not None else self.options["seed_simulator"]
        if seed is None:
            seed = int(np.random.SeedSequence().entropy % (1 << 31))
        lib = L.lib()
        out = []
        for i, circ in enumerate(circuits):
            cfg = L.AerhipRunConfig()
            lib.aerhip_run_config_default(C.byref(cfg))
            cfg.precision = L.PRECISION_DOUBLE if self.options["precision"] == "double" else L.PRECISION_SINGLE
            cfg.device = int(self.options["target_gpu"])
            cfg.shots = int(shots)
            cfg.seed = int(seed) + 2113 * i
            cfg.fusion_enable = int(bool(self.options["fusion_enable"]))
            cfg.fusion_max_qubit = int(self.options["fusion_max_qubit"])
            cfg.fusion_threshold = int(self.options["fusion_threshold"])
            cfg.sample_index_size = int(self.options["statevector_sample_measure_opt"])
            cfg.global_phase = float(circ.global_phase)
            ops, keep = circ.pack()
            mem = np.zeros(max(int(shots), 1), dtype=np.uint64)
            sv = np.empty(1 << circ.num_qubits, dtype=np.complex128) if save_statevector or any(
                o.type == L.OP_SAVE_STATEVECTOR for o in circ.ops) else None
            timing = np.zeros(4, dtype=np.float64)
            L.check(lib.aerhip_run_circuit(ops, len(circ.ops), circ.num_qubits, circ.num_clbits,
                                           C.byref(cfg), mem.ctypes.data,
                                           sv.ctypes.data if sv is not None else None,
                                           timing.ctypes.data))
            del keep
            mem = mem[:shots]
            vals, cnt = np.unique(mem, return_counts=True)
            counts = {hex(int(v)): int(c) for v, c in zip(vals, cnt)} if circ.num_clbits else {}
            out.append(ExperimentResult(counts, mem, sv, int(cfg.seed), float(timing[0]),
                                        float(timing[1]), int(timing[2]), int(timing[3]),
                                        {"method": "statevector", "device": "GPU",
                                         "precision": self.options["precision"]}))
        return Result(out)
