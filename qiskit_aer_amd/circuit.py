"""A minimal circuit container with Aer's operation names.

``Circuit`` records operations the way Aer's ``AerCircuit`` holds them after
``aer_compiler.assemble_circuits`` (gate name, qubits, real parameters;
``unitary`` / ``diagonal`` matrices; ``measure`` with memory bits), and packs
them into the ``aerhip_op`` wire format of include/aerhip.h.
"""
from __future__ import annotations

import ctypes as C
import itertools
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib as L


@dataclass
class Op:
    type: int
    name: str = ""
    qubits: tuple = ()
    params: tuple = ()
    matrix: Optional[np.ndarray] = None  # complex, column-major vectorised or diagonal
    label: Optional[str] = None          # Pauli label
    memory: tuple = ()


@dataclass
class Circuit:
    num_qubits: int
    num_clbits: int = 0
    global_phase: float = 0.0
    ops: list = field(default_factory=list)

    # generic ------------------------------------------------------------------
    def gate(self, name: str, qubits: Sequence[int], params: Sequence[float] = ()) -> "Circuit":
        for q in qubits:
            if not 0 <= int(q) < self.num_qubits:
                raise ValueError(f"qubit {q} out of range")
        self.ops.append(Op(L.OP_GATE, name, tuple(int(q) for q in qubits),
                           tuple(float(p) for p in params)))
        return self

    def unitary(self, matrix, qubits: Sequence[int]) -> "Circuit":
        m = np.asarray(matrix, dtype=np.complex128)
        d = 1 << len(qubits)
        if m.shape != (d, d):
            raise ValueError("unitary has the wrong shape")
        self.ops.append(Op(L.OP_MATRIX, "unitary", tuple(int(q) for q in qubits),
                           matrix=np.ascontiguousarray(m.reshape(-1, order="F"))))
        return self

    def diagonal(self, diag, qubits: Sequence[int]) -> "Circuit":
        d = np.ascontiguousarray(np.asarray(diag, dtype=np.complex128).reshape(-1))
        if d.size != 1 << len(qubits):
            raise ValueError("diagonal has the wrong size")
        self.ops.append(Op(L.OP_DIAGONAL, "diagonal", tuple(int(q) for q in qubits), matrix=d))
        return self

    def initialize(self, state, qubits: Sequence[int]) -> "Circuit":
        s = np.ascontiguousarray(np.asarray(state, dtype=np.complex128).reshape(-1))
        self.ops.append(Op(L.OP_INITIALIZE, "initialize", tuple(int(q) for q in qubits), matrix=s))
        return self

    def measure(self, qubits: Sequence[int], clbits: Sequence[int]) -> "Circuit":
        qubits, clbits = list(qubits), list(clbits)
        self.num_clbits = max(self.num_clbits, max(clbits) + 1)
        self.ops.append(Op(L.OP_MEASURE, "measure", tuple(qubits), memory=tuple(clbits)))
        return self

    def measure_all(self) -> "Circuit":
        self.num_clbits = max(self.num_clbits, self.num_qubits)
        for q in range(self.num_qubits):
            self.ops.append(Op(L.OP_MEASURE, "measure", (q,), memory=(q,)))
        return self

    def reset(self, qubits) -> "Circuit":
        self.ops.append(Op(L.OP_RESET, "reset", tuple(np.atleast_1d(qubits).tolist())))
        return self

    def barrier(self, *qubits) -> "Circuit":
        q = qubits or tuple(range(self.num_qubits))
        self.ops.append(Op(L.OP_BARRIER, "barrier", tuple(q)))
        return self

    def save_statevector(self) -> "Circuit":
        self.ops.append(Op(L.OP_SAVE_STATEVECTOR, "save_statevector", tuple(range(self.num_qubits))))
        return self

    def pauli(self, label: str, qubits: Sequence[int]) -> "Circuit":
        self.ops.append(Op(L.OP_GATE, "pauli", tuple(qubits), label=label))
        return self

    # named gates (Aer spellings) ----------------------------------------------
    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        nparams = {"u3": 3, "u": 3, "u2": 2, "u1": 1, "p": 1, "rx": 1, "ry": 1, "rz": 1,
                   "r": 2, "cp": 1, "crx": 1, "cry": 1, "crz": 1, "cu1": 1, "cu2": 2,
                   "cu3": 3, "cu": 4, "rxx": 1, "ryy": 1, "rzz": 1, "rzx": 1, "mcp": 1,
                   "mcphase": 1, "mcu1": 1, "mcu2": 2, "mcu3": 3, "mcu": 4, "mcr": 2,
                   "mcrx": 1, "mcry": 1, "mcrz": 1}.get(name, 0)

        def method(*args):
            params, qubits = args[:nparams], args[nparams:]
            if len(qubits) == 1 and isinstance(qubits[0], (list, tuple)):
                qubits = qubits[0]
            return self.gate(name, qubits, params)
        return method

    def count_gates(self) -> int:
        return sum(1 for o in self.ops if o.type in (L.OP_GATE, L.OP_MATRIX, L.OP_DIAGONAL))

    # wire format ---------------------------------------------------------------
    def pack(self):
        """Returns (ctypes array of AerhipOp, keepalive list).

        The qubits, real parameters and names of all ops go into three flat
        buffers and the records are filled column by column, so packing costs
        a few Python-level passes over the op list rather than a handful of
        small arrays per op."""
        ops = self.ops
        nops = len(ops)
        arr = (L.AerhipOp * nops)()
        keep = []
        if nops == 0:
            return arr, keep
        rec = np.frombuffer(arr, dtype=_OP_DTYPE)
        types = np.fromiter((o.type for o in ops), dtype=np.int32, count=nops)
        rec["type"] = types
        # qubits
        nq = np.fromiter((len(o.qubits) for o in ops), dtype=np.int64, count=nops)
        qflat = np.fromiter(itertools.chain.from_iterable(o.qubits for o in ops), dtype=np.uint64,
                            count=int(nq.sum()))
        qflat = np.concatenate([qflat, np.zeros(1, dtype=np.uint64)])
        keep.append(qflat)
        qoff = np.concatenate([[0], np.cumsum(nq)[:-1]])
        rec["num_qubits"] = nq
        rec["qubits"] = np.uint64(qflat.ctypes.data) + (qoff * 8).astype(np.uint64)
        # real parameters: gate angles, or complex matrices as (re, im) pairs
        is_gate = types == L.OP_GATE
        mats = [i for i in range(nops) if not is_gate[i] and ops[i].matrix is not None]
        npar = np.fromiter((len(o.params) if o.type == L.OP_GATE else 0 for o in ops), dtype=np.int64,
                           count=nops)
        for i in mats:
            npar[i] = 2 * ops[i].matrix.size
        poff = np.concatenate([[0], np.cumsum(npar)[:-1]])
        pflat = np.zeros(int(npar.sum()) + 1, dtype=np.float64)
        gate_idx = np.nonzero(is_gate & (npar > 0))[0]
        if gate_idx.size:
            vals = np.fromiter(itertools.chain.from_iterable(ops[i].params for i in gate_idx),
                               dtype=np.float64, count=int(npar[gate_idx].sum()))
            starts = poff[gate_idx]
            lens = npar[gate_idx]
            pos = np.repeat(starts - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens) + np.arange(vals.size)
            pflat[pos] = vals
        for i in mats:
            pflat[poff[i]:poff[i] + npar[i]] = np.ascontiguousarray(
                ops[i].matrix.astype(np.complex128)).view(np.float64)
        keep.append(pflat)
        rec["num_params"] = np.where(is_gate, npar, npar // 2)
        rec["params"] = np.uint64(pflat.ctypes.data) + (poff * 8).astype(np.uint64)
        # names: one NUL-terminated copy of each distinct name
        names = sorted({o.name for o in ops})
        blob = b"".join(nm.encode() + b"\0" for nm in names)
        nbuf = np.frombuffer(blob + b"\0", dtype=np.uint8)
        keep.append(nbuf)
        addr, at = {}, 0
        for nm in names:
            addr[nm] = nbuf.ctypes.data + at
            at += len(nm.encode()) + 1
        rec["name"] = np.fromiter((addr[o.name] for o in ops), dtype=np.uint64, count=nops)
        # Pauli labels and measure clbits (rare)
        for i, op in enumerate(ops):
            if op.label is not None:
                lab = op.label.encode()
                keep.append(lab)
                arr[i].string_param = lab
            if op.memory:
                m = np.ascontiguousarray(np.asarray(op.memory, dtype=np.uint64))
                keep.append(m)
                arr[i].memory = m.ctypes.data_as(C.POINTER(C.c_uint64))
        return arr, keep


# Field layout of AerhipOp (include/aerhip.h) for filling packed records.
_OP_DTYPE = np.dtype({
    "names": ["type", "num_qubits", "name", "qubits", "num_params", "reserved", "params",
              "string_param", "memory"],
    "formats": ["<i4", "<u4", "<u8", "<u8", "<u4", "<u4", "<u8", "<u8", "<u8"],
    "offsets": [L.AerhipOp.type.offset, L.AerhipOp.num_qubits.offset, L.AerhipOp.name.offset,
                L.AerhipOp.qubits.offset, L.AerhipOp.num_params.offset, L.AerhipOp.reserved.offset,
                L.AerhipOp.params.offset, L.AerhipOp.string_param.offset, L.AerhipOp.memory.offset],
    "itemsize": C.sizeof(L.AerhipOp),
})
