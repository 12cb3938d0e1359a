"""aerhip_expval_pauli_noalloc (the entry point that allocates nothing)
against aerhip_expval_pauli on the same states: two calls in a row, the second
on a smaller state, every value bit-identical (same kernels, same summation
order).  The device output buffer is the test's own (hipMalloc through the HIP
runtime the engine links); the entry point itself allocates nothing."""
import ctypes as C

import numpy as np
import pytest

from qiskit_aer_amd import _lib as L

pytestmark = pytest.mark.gpu

LABEL_CASES = [
    ([0, 3, 7], "ZZZ"),
    ([1, 5], "XY"),
    ([0, 2, 4, 6], "YXZY"),
    ([2], "I"),
    ([0, 1, 2, 3, 4, 5, 6, 7, 8, 9], "XXYYZZIXYZ"),
]


def _hip():
    """The HIP runtime the engine library links (torch ships another copy
    under the same name, which may be loaded first)."""
    L.lib()
    path = "/opt/rocm/lib/libamdhip64.so.7"
    with open("/proc/self/maps") as f:
        for line in f:
            parts = line.split()
            if parts and "libamdhip64.so" in parts[-1] and "/torch/" not in parts[-1]:
                path = parts[-1]
                break
    h = C.CDLL(path)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _noalloc(sv_ptr, n, precision, qubits, label, phase, dout, hip):
    q = (C.c_uint64 * len(qubits))(*qubits)
    L.check(L.lib().aerhip_expval_pauli_noalloc(
        C.c_void_p(sv_ptr), n, precision, q, len(qubits), label.encode(), phase.real, phase.imag,
        dout, None))
    assert hip.hipDeviceSynchronize() == 0
    host = C.c_double(0.0)
    assert hip.hipMemcpy(C.byref(host), dout, 8, 2) == 0  # hipMemcpyDeviceToHost
    return host.value


@pytest.mark.parametrize("first,second", [((20, "double"), (12, "double")),
                                          ((18, "single"), (11, "single"))])
def test_noalloc_matches_expval_pauli(first, second):
    from qiskit_aer_amd import QubitVector
    hip = _hip()
    dout = C.c_void_p()
    assert hip.hipMalloc(C.byref(dout), 8) == 0
    try:
        for n, prec in (first, second):
            print(f"case: noalloc expval, {n} qubits, {prec}", flush=True)
            qv = QubitVector(n, prec)
            qv.initialize_random(1234 + n)
            qv.synchronize()
            ptr = qv.device_pointer()
            p = 0 if prec == "double" else 1  # AERHIP_PRECISION_DOUBLE / _SINGLE
            for qubits, label in LABEL_CASES:
                qubits = [x for x in qubits if x < n]
                label = label[len(label) - len(qubits):]
                for phase in (1.0 + 0j, 0.3 - 0.7j):
                    want = qv.expval_pauli(qubits, label, phase)
                    got = _noalloc(ptr, n, p, qubits, label, phase, dout, hip)
                    assert got == want, (n, prec, qubits, label, phase, got, want)
            qv.close()
    finally:
        hip.hipFree(dout)


def test_noalloc_rejects_bad_sizes():
    from qiskit_aer_amd import QubitVector
    qv = QubitVector(10, "double")
    qv.initialize()
    qv.synchronize()
    ptr = qv.device_pointer()
    q = (C.c_uint64 * 1)(0)
    lib = L.lib()
    with pytest.raises(L.AerhipInvalidArgument):
        L.check(lib.aerhip_expval_pauli_noalloc(C.c_void_p(ptr), 49, 0, q, 1, b"Z", 1.0, 0.0,
                                                C.c_void_p(ptr), None))
    q[0] = 10
    with pytest.raises(L.AerhipInvalidArgument):
        L.check(lib.aerhip_expval_pauli_noalloc(C.c_void_p(ptr), 10, 0, q, 1, b"Z", 1.0, 0.0,
                                                C.c_void_p(ptr), None))
    qv.close()
