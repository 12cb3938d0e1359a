"""ctypes binding of the engine's C ABI (include/aerhip.h).

The library is loaded from the package's ``lib/`` directory.  There is no
fallback: if the HIP library is missing every call raises, so no result can
silently come from another implementation.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import threading

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libaerhip.so"

PRECISION_DOUBLE = 0
PRECISION_SINGLE = 1

ROT_X, ROT_Y, ROT_Z, ROT_XX, ROT_YY, ROT_ZZ, ROT_ZX = range(7)

OP_GATE = 0
OP_MATRIX = 1
OP_DIAGONAL = 2
OP_MEASURE = 3
OP_BARRIER = 4
OP_RESET = 5
OP_INITIALIZE = 6
OP_SAVE_STATEVECTOR = 7

_u64p = C.POINTER(C.c_uint64)
_dp = C.POINTER(C.c_double)


class AerhipOp(C.Structure):
    """``aerhip_op`` of include/aerhip.h."""

    _fields_ = [
        ("type", C.c_int32),
        ("num_qubits", C.c_uint32),
        ("name", C.c_char_p),
        ("qubits", _u64p),
        ("num_params", C.c_uint32),
        ("reserved", C.c_uint32),
        ("params", _dp),
        ("string_param", C.c_char_p),
        ("memory", _u64p),
    ]


class AerhipRunConfig(C.Structure):
    """``aerhip_run_config`` of include/aerhip.h."""

    _fields_ = [
        ("precision", C.c_int32),
        ("device", C.c_int32),
        ("shots", C.c_uint64),
        ("seed", C.c_uint64),
        ("fusion_enable", C.c_int32),
        ("fusion_max_qubit", C.c_int32),
        ("fusion_threshold", C.c_int32),
        ("sample_index_size", C.c_int32),
        ("global_phase", C.c_double),
    ]


class AerhipError(RuntimeError):
    """Error reported by the engine (the reference raises std::runtime_error)."""


class AerhipInvalidArgument(AerhipError, ValueError):
    """Invalid argument reported by the engine (std::invalid_argument)."""


_SIGNATURES = {
    "aerhip_last_error": (C.c_char_p, []),
    "aerhip_device_count": (C.c_int, []),
    "aerhip_version": (C.c_char_p, []),
    "aerhip_device_memory": (C.c_int, [C.c_int, _u64p, _u64p]),
    "aerhip_qv_create": (C.c_void_p, [C.c_uint64, C.c_int, C.c_int]),
    "aerhip_qv_destroy": (None, [C.c_void_p]),
    "aerhip_qv_num_qubits": (C.c_int, [C.c_void_p, _u64p]),
    "aerhip_required_memory_mb": (C.c_uint64, [C.c_uint64, C.c_int]),
    "aerhip_qv_initialize": (C.c_int, [C.c_void_p]),
    "aerhip_qv_zero": (C.c_int, [C.c_void_p]),
    "aerhip_qv_initialize_from_vector": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "aerhip_qv_copy_to_vector": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "aerhip_qv_initialize_component": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, C.c_void_p]),
    "aerhip_qv_synchronize": (C.c_int, [C.c_void_p]),
    "aerhip_qv_set_sample_measure_index_size": (C.c_int, [C.c_void_p, C.c_int]),
    "aerhip_qv_initialize_from_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int]),
    "aerhip_qv_initialize_random": (C.c_int, [C.c_void_p, C.c_uint64]),
    "aerhip_qv_device_pointer":(C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "aerhip_qv_set_tile_bits": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "aerhip_qv_stats": (C.c_int, [C.c_void_p, _u64p]),
    "aerhip_qv_set_register_bits": (C.c_int, [C.c_void_p, C.c_int]),
    "aerhip_qv_set_event_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "aerhip_qv_pass_timing": (C.c_int, [C.c_void_p, _dp, _u64p]),
    "aerhip_apply_matrix": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, C.c_void_p]),
    "aerhip_apply_diagonal_matrix": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, C.c_void_p]),
    "aerhip_apply_mcx": (C.c_int, [C.c_void_p, _u64p, C.c_uint64]),
    "aerhip_apply_mcy": (C.c_int, [C.c_void_p, _u64p, C.c_uint64]),
    "aerhip_apply_mcswap": (C.c_int, [C.c_void_p, _u64p, C.c_uint64]),
    "aerhip_apply_mcphase": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, C.c_double, C.c_double]),
    "aerhip_apply_mcu": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, C.c_void_p]),
    "aerhip_apply_rotation": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, C.c_int, C.c_double]),
    "aerhip_apply_pauli": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, C.c_char_p, C.c_double, C.c_double]),
    "aerhip_apply_permutation_matrix": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, _u64p, C.c_uint64]),
    "aerhip_apply_multiplexer": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, _u64p, C.c_uint64, C.c_void_p]),
    "aerhip_apply_multi_swaps": (C.c_int, [C.c_void_p, _u64p, C.c_uint64]),
    "aerhip_norm": (C.c_int, [C.c_void_p, _dp]),
    "aerhip_qv_set_fusion": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "aerhip_qv_fused_count": (C.c_int, [C.c_void_p, _u64p]),
    "aerhip_qv_graph_stats": (C.c_int, [C.c_void_p, _u64p, _u64p]),
    "aerhip_qv_set_block_fusion": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "aerhip_norm_matrix": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, _dp, _dp]),
    "aerhip_norm_diagonal": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, _dp, _dp]),
    "aerhip_qv_get_state": (C.c_int, [C.c_void_p, C.c_uint64, _dp]),
    "aerhip_qv_set_state": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.c_double]),
    "aerhip_qv_inner_product": (C.c_int, [C.c_void_p, C.c_void_p, _dp, _dp]),
    "aerhip_probability": (C.c_int, [C.c_void_p, C.c_uint64, _dp]),
    "aerhip_probabilities": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, C.c_void_p]),
    "aerhip_sample_measure": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "aerhip_expval_pauli": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, C.c_char_p, C.c_double, C.c_double, _dp]),
    "aerhip_expval_pauli_noalloc": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int, _u64p, C.c_uint64,
                                              C.c_char_p, C.c_double, C.c_double, C.c_void_p,
                                              C.c_void_p]),
    "aerhip_save_expval_pauli": (C.c_int, [C.c_void_p, _u64p, C.c_uint64, C.c_char_p, C.c_double,
                                           C.c_double, _u64p]),
    "aerhip_saved_expvals": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, _u64p]),
    "aerhip_qv_pauli_sum_to_upper_half": (C.c_int, [C.c_void_p, C.c_char_p, C.c_void_p, C.c_uint64]),
    "aerhip_batch_pauli_expval": (C.c_int, [C.POINTER(AerhipOp), _u64p, C.POINTER(C.c_uint32),
                                            C.c_uint64, C.c_char_p, _dp, _u64p, C.c_int, C.c_int,
                                            C.c_int, _dp, _dp]),
    "aerhip_adjoint_expval_gradient": (C.c_int, [C.POINTER(AerhipOp), C.c_uint64, C.c_uint64,
                                                 C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                                 C.c_char_p, C.c_void_p, C.c_uint64, C.c_int,
                                                 C.c_int, C.c_void_p, C.c_void_p]),
    "aerhip_bind_pauli_expval": (C.c_int, [C.POINTER(AerhipOp), C.c_uint64, C.c_uint64, C.c_void_p,
                                           C.c_uint64, C.c_uint64, C.c_void_p, C.c_char_p, C.c_void_p,
                                           C.c_uint64, C.c_int, C.c_int, C.c_void_p, _dp]),
    "aerhip_run_config_default": (None, [C.POINTER(AerhipRunConfig)]),
    "aerhip_apply_ops": (C.c_int, [C.c_void_p, C.POINTER(AerhipOp), C.c_uint64]),
    "aerhip_run_circuit": (C.c_int, [C.POINTER(AerhipOp), C.c_uint64, C.c_uint64, C.c_uint64,
                                     C.POINTER(AerhipRunConfig), C.c_void_p, C.c_void_p, C.c_void_p]),
}

_lock = threading.Lock()
_lib = None


def lib() -> C.CDLL:
    """Loads (once) and returns the engine library; raises if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise AerhipError(
                    f"qiskit_aer_amd: the HIP engine library {LIB_PATH} is not built; "
                    "run `python -m qiskit_aer_amd.build`")
            L = C.CDLL(str(LIB_PATH))
            for name, (res, args) in _SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
        return _lib


def check(rc: int) -> None:
    """Raises the engine's error message when a call returned non-zero."""
    if rc != 0:
        msg = lib().aerhip_last_error().decode(errors="replace")
        if rc == 2:
            raise AerhipInvalidArgument(msg)
        raise AerhipError(msg)


def exported_symbols() -> list[str]:
    """Names of the C ABI functions this binding declares."""
    return sorted(_SIGNATURES)
