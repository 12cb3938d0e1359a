"""The controller_wrappers-compatible module: the surface the reference's
Python front end builds and executes circuits through
(qiskit_aer/backends/wrappers/aer_controller_binding.hpp:84-380,
aer_circuit_binding.hpp:150-190), checked against the oracle.

CPU tests cover the surface (AerConfig options as the compiler probes them
with hasattr, AerCircuit builders, unsupported instructions reported as
failed experiments, no silent CPU fallback); GPU tests compare counts, saved
statevectors and saved expectation values with the oracle."""
import math

import numpy as np
import pytest

from oracle import pyoracle as O
from qiskit_aer_amd import circuits
from qiskit_aer_amd import controller_wrappers as cw
from tests import dense as D


def to_aer_circuit(c, shots, memory_ops=True):
    """AerCircuit of a qiskit_aer_amd.Circuit, built the way the reference's
    aer_compiler.assemble_circuit calls the builder (aer_compiler.py:711-1044)."""
    a = cw.AerCircuit()
    a.num_qubits = c.num_qubits
    a.num_memory = c.num_clbits
    a.shots = shots
    a.global_phase_angle = c.global_phase
    a.set_header({"name": "circuit"})
    for op in c.ops:
        if op.name == "measure":
            if memory_ops:
                a.measure(list(op.qubits), list(op.memory), [])
        elif op.name == "barrier":
            a.barrier(list(op.qubits))
        else:
            a.gate(op.name, list(op.qubits), [complex(p) for p in op.params], [], -1, None, "")
    return a


def config(seed, **kw):
    cfg = cw.AerConfig()
    cfg.method = "statevector"
    cfg.device = "GPU"
    cfg.seed_simulator = seed
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def test_config_options_as_the_compiler_probes_them():
    cfg = cw.AerConfig()
    for key in ("shots", "method", "device", "precision", "seed_simulator", "fusion_enable",
                "fusion_max_qubit", "fusion_threshold", "statevector_sample_measure_opt",
                "statevector_parallel_threshold", "memory", "memory_slots", "n_qubits",
                "max_parallel_threads", "target_gpus", "cuStateVec_enable", "blocking_enable",
                "batched_shots_gpu", "parameterizations", "library_dir"):
        assert hasattr(cfg, key), key
    assert not hasattr(cfg, "not_an_option")
    assert (cfg.shots, cfg.precision, cfg.fusion_enable, cfg.statevector_sample_measure_opt) == \
        (1024, "double", True, 10)
    assert cfg.seed_simulator is None
    cfg.seed_simulator = 7
    cfg.fusion_max_qubit = 5
    assert cfg.seed_simulator == 7 and cfg.fusion_max_qubit == 5
    assert cw.AerConfig.GLOBAL_PHASE_POS == -1


def test_compiler_imports_resolve():
    # the names aer_compiler.py:47-60 imports from the module
    for name in ("AerUnaryExpr", "AerUnaryOp", "AerBinaryExpr", "AerBinaryOp", "AerUintValue",
                 "AerBoolValue", "AerUint", "AerBool", "AerCast", "AerVar", "AerCircuit", "AerConfig",
                 "aer_controller_execute", "aer_initialize_libraries"):
        assert hasattr(cw, name), name
    assert cw.AerBinaryOp.Equal != cw.AerBinaryOp.Less
    e = cw.AerBinaryExpr(cw.AerBinaryOp.Equal, cw.AerVar(cw.AerUint(2), [0, 1]), cw.AerUintValue(cw.AerUint(2), 1))
    a = cw.AerCircuit()
    a.num_qubits = 1
    a.gate("x", [0], [], [], -1, e, "")
    r = cw.aer_controller_execute().execute([a], {}, config(1))
    assert "conditional" in r["results"][0]["status"]


def test_circuit_builder_and_unsupported_instructions_fail_loudly():
    a = cw.AerCircuit()
    a.num_qubits = 2
    a.num_memory = 1
    a.gate("h", [0], [], [])
    a.unitary([0, 1], np.eye(4), -1, None, "u")
    a.diagonal([1], [1, 1j], -1, "d")
    a.kraus([0], [np.eye(2)])  # another method's instruction
    a.measure([0], [0], [])
    assert a.num_ops == 5
    r = cw.aer_controller_execute().execute([a], {}, config(1))
    assert r["success"] is False and "kraus" in r["results"][0]["status"]
    # conditional gates are outside this engine
    b = cw.AerCircuit()
    b.num_qubits = 1
    b.gate("x", [0], [], [], 0, None, "")
    r = cw.aer_controller_execute().execute([b], {}, config(1))
    assert "conditional" in r["results"][0]["status"]
    # a noise model is refused for every experiment
    c = cw.AerCircuit()
    c.num_qubits = 1
    r = cw.aer_controller_execute().execute([c], {"errors": [1]}, config(1))
    assert "noise" in r["results"][0]["status"]


def test_no_device_is_a_failed_experiment_not_a_fallback():
    from qiskit_aer_amd import _lib as L
    if L.lib().aerhip_device_count() > 0:
        pytest.skip("a HIP device is present")
    c = circuits.brickwork_u3_cx(4, 2, seed=1)
    r = cw.aer_controller_execute().execute([to_aer_circuit(c, 16)], {}, config(3))
    assert r["results"][0]["success"] is False


@pytest.mark.gpu
def test_counts_and_memory_match_the_oracle():
    print("case: controller_wrappers counts, two circuits", flush=True)
    cs = [circuits.brickwork_u3_cx(10, 5, seed=11), circuits.brickwork_u3_cx(7, 4, seed=12)]
    r = cw.aer_controller_execute().execute([to_aer_circuit(c, 300) for c in cs], {},
                                           config(41, memory=True))
    assert r["success"] and r["status"] == "COMPLETED"
    for i, c in enumerate(cs):
        res = r["results"][i]
        assert res["success"] and res["seed_simulator"] == 41 + 2113 * i
        mem, _, _ = O.run_circuit(c, shots=300, seed=41 + 2113 * i, statevector=False)
        assert res["data"]["memory"] == [hex(int(m)) for m in mem]
        vals, cnt = np.unique(mem, return_counts=True)
        assert res["data"]["counts"] == {hex(int(v)): int(k) for v, k in zip(vals, cnt)}
        assert res["header"] == {"name": "circuit"}


@pytest.mark.gpu
def test_saved_statevector_and_expvals_match_the_oracle():
    print("case: controller_wrappers save_statevector / save_expval", flush=True)
    n = 12
    c = circuits.brickwork_u3_cx(n, 6, seed=5, measure=False)
    a = to_aer_circuit(c, 1)
    a.save_state(list(range(n)), "save_statevector", "single", "statevector")
    r = cw.aer_controller_execute().execute([a], {}, config(0))
    assert r["success"], r["results"][0]["status"]
    _, sv, _ = O.run_circuit(c, shots=0, seed=0)
    assert np.max(np.abs(r["results"][0]["data"]["statevector"] - sv)) < 1e-12

    # save_expval (average) and save_expval_var in the middle and at the end
    half = len(c.ops) // 2
    first, second = c.ops[:half], c.ops[half:]
    terms = [("ZZIIII", 0.5), ("XIYIIZ", -1.25), ("IIIIXX", 0.75)]
    qs = [0, 2, 3, 5, 7, 11]
    b = cw.AerCircuit()
    b.num_qubits = n
    b.shots = 1
    for op in first:
        b.gate(op.name, list(op.qubits), [complex(p) for p in op.params], [])
    b.save_expval(qs, "save_expval", [t for t, _ in terms], [w for _, w in terms], [0.0] * 3,
                  "average", "mid")
    for op in second:
        b.gate(op.name, list(op.qubits), [complex(p) for p in op.params], [])
    sq = [w * w for _, w in terms]
    b.save_expval(qs, "save_expval_var", [t for t, _ in terms], [w for _, w in terms], sq,
                  "single", "end")
    r = cw.aer_controller_execute().execute([b], {}, config(0))
    res = r["results"][0]
    assert res["success"], res["status"]
    o = O.OracleQubitVector(n)
    o.initialize()
    want = {}
    for key, part in (("mid", first), ("end", second)):
        for op in part:
            if op.name == "u3":
                o.apply_mcu(list(op.qubits), D.u3(*op.params))
            else:
                o.apply_mcx(list(op.qubits))
        vals = [o.expval_pauli(qs, t) for t, _ in terms]
        want[key] = (sum(w * v for (_, w), v in zip(terms, vals)),
                     sum(s * v for s, v in zip(sq, vals)))
    assert abs(res["data"]["mid"] - want["mid"][0]) < 1e-12
    mean, var = res["data"]["end"]
    assert abs(mean - want["end"][0]) < 1e-12
    assert abs(var - (want["end"][1] - want["end"][0] ** 2)) < 1e-12
    assert not math.isnan(var)
