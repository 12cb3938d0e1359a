"""Throughput of the MI355X statevector engine on the 30-qubit workloads.

One step = one pass of the hot path over one batch: the whole 30-qubit
U3+CX brickwork circuit of depth 40 (1780 gates: 1200 U3 + 580 CX, angles
from a fixed seed) applied to a complex128 state already resident in HBM.
Prints one JSON line: gates/s (value), achieved HBM GB/s, the roofline of the
tile-pass kernel (HIP events on the engine's stream), the complex64 variant,
the 30-qubit QFT, and a CPU baseline (the oracle restatement of the
reference's OpenMP CPU path on a bounded sample of the same workload).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--qubits 30]
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
FP64_PEAK_TFLOPS = 78.6  # MI355X published FP64 peak (vector and matrix)

# HBM bytes per tile-pass launch of the default workload (30 qubits, c128),
# from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (profiles/reg_pass_30q_brickwork_c128.md):
# 2.0 x 8 388 975 KiB fetched (FETCH_SIZE counts half the bytes, calibrated on a
# kernel that reads the 16 GiB state once) + 16 977 477 KiB written, averaged over
# the passes of the block-fused step at width 2.
MEASURED_TRAFFIC_30Q_C128 = (2 * 8388975 + 16977477) * 1024


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def run_engine(n: int, precision: str, circ, steps: int, warmup: int, init=None,
               fusion_width: int = 0) -> dict:
    from qiskit_aer_amd import QubitVector
    qv = QubitVector(n, precision)
    if fusion_width:
        qv.set_block_fusion(fusion_width, cost_check=True)
    if init is None:
        qv.initialize()
    else:
        init(qv)
    packed = circ.pack()
    for _ in range(warmup):
        qv.apply_ops(packed)
        qv.synchronize()
    st0 = qv.stats()
    qv.set_event_timing(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        qv.apply_ops(packed)
        qv.synchronize()
    t1 = time.perf_counter()
    ms, launches = qv.pass_timing()
    st1 = qv.stats()
    norm = qv.norm()
    qv.close()
    amp = 16 if precision == "double" else 8
    sweep_bytes = 2 * (1 << n) * amp
    passes = st1["passes"] - st0["passes"]
    wall = t1 - t0
    gates = circ.count_gates() * steps
    return {
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / steps,
        "gates_per_s": gates / wall,
        "passes_per_step": passes / steps,
        "gathers_per_step": (st1["gathers"] - st0["gathers"]) / steps,
        "hbm_gbs": passes * sweep_bytes / wall / 1e9,
        "kernel_ms_avg": ms / max(launches, 1),
        "kernel_gbs": sweep_bytes / (ms / max(launches, 1) / 1e3) / 1e9 if launches else None,
        "sweep_bytes": sweep_bytes,
        "norm_after": norm,
    }


def tfim(n: int, h: float = 0.7) -> list:
    """-sum Z_i Z_{i+1} - h sum X_i as (label, coefficient) pairs; label
    character n-1-i acts on qubit i."""
    terms = []
    for i in range(n - 1):
        lab = ["I"] * n
        lab[n - 1 - i] = lab[n - 2 - i] = "Z"
        terms.append(("".join(lab), -1.0))
    for i in range(n):
        lab = ["I"] * n
        lab[n - 1 - i] = "X"
        terms.append(("".join(lab), -h))
    return terms


def ry_rz_cz_ansatz(n: int, layers: int):
    """RY and RZ on every qubit then a CZ brickwork, per layer: the circuit,
    the parameter index of each op (-1 for CZ) and the parameter count."""
    from qiskit_aer_amd.circuit import Circuit
    c = Circuit(n)
    idx, p = [], 0
    for layer in range(layers):
        for q in range(n):
            c.gate("ry", [q], [0.0]); idx.append(p); p += 1
            c.gate("rz", [q], [0.0]); idx.append(p); p += 1
        for a in range(layer % 2, n - 1, 2):
            c.gate("cz", [a, a + 1]); idx.append(-1)
    return c, idx, p


def run_adjoint(n: int, layers: int, batch: int, shift_params: int, seed: int) -> dict:
    """Energy + full gradient for `batch` parameter rows by the adjoint sweep,
    against the parameter-shift rule on the same engine for `shift_params`
    parameters of one row (two energy evaluations each)."""
    from qiskit_aer_amd.adjoint import AdjointEstimator
    c, idx, npar = ry_rz_cz_ansatz(n, layers)
    est = AdjointEstimator(c, idx, npar, tfim(n))
    rows = np.random.default_rng(seed).uniform(-np.pi, np.pi, size=(batch, npar))
    est.energy_and_gradient(rows[:1])  # warm-up
    t0 = time.perf_counter()
    energy, grad = est.energy_and_gradient(rows)
    t_adj = time.perf_counter() - t0
    which = list(range(0, npar, max(1, npar // shift_params)))[:shift_params]
    t0 = time.perf_counter()
    ps = est.parameter_shift(rows[:1], which=which)
    t_ps = time.perf_counter() - t0
    adj_rate = batch * npar / t_adj
    ps_rate = len(which) / t_ps
    return {"grads_per_s": adj_rate, "param_shift_grads_per_s": ps_rate,
            "speedup_vs_param_shift": adj_rate / ps_rate, "num_params": npar,
            "batch": batch, "gates": c.count_gates(), "terms": 2 * n - 1,
            "adjoint_s": t_adj, "param_shift_sample": f"{len(which)} params of row 0, {t_ps:.2f} s",
            "max_abs_adjoint_minus_shift": float(np.max(np.abs(grad[0, which] - ps[0])))}


def run_mixed_precision(n: int, layers: int, seed: int) -> dict:
    """Haar-SU(4) brickwork on n qubits in complex128 and complex64, both
    states resident on the device: gate throughput of each, 1 - |<psi64|psi128>|^2
    and the mean |error| of <Z_i Z_{i+1}>."""
    from qiskit_aer_amd import QubitVector, circuits
    from qiskit_aer_amd.autotune import tune
    c = circuits.su4_brickwork(n, layers, seed)
    packed = c.pack()
    out = {"gates": c.count_gates(), "num_qubits": n, "layers": layers}
    # Block-fusion width per precision, picked by the autotuner on the same
    # brickwork at 26 qubits (the 32-qubit runs below use the chosen width).
    proxy = circuits.su4_brickwork(min(n, 26), layers, seed)
    width = {}
    out["autotune"] = {}
    for prec in ("double", "single"):
        best, times = tune(proxy, prec, widths=(0, 2, 4, 5))
        width[prec] = best
        out["autotune"][prec] = {"proxy_qubits": proxy.num_qubits, "best_width": best,
                                 "ms_by_width": times}
    states = {}
    labels = []
    for i in range(n - 1):
        lab = ["I"] * n
        lab[n - 1 - i] = lab[n - 2 - i] = "Z"
        labels.append("".join(lab))
    zz = {}
    for prec in ("double", "single"):
        qv = QubitVector(n, prec)
        qv.set_block_fusion(width[prec])
        qv.initialize()
        qv.synchronize()
        qv.set_event_timing(True)
        t0 = time.perf_counter()
        qv.apply_ops(packed)
        qv.synchronize()
        wall = time.perf_counter() - t0
        ms, launches = qv.pass_timing()
        out[prec] = {"gates_per_s": c.count_gates() / wall, "ms": 1e3 * wall,
                     "kernel_ms": ms, "passes": launches}
        zz[prec] = qv.expval_pauli_multi(list(range(n)), labels)
        states[prec] = qv
    ov = states["single"].inner_product(states["double"])
    n64 = states["single"].norm()
    n128 = states["double"].norm()
    out["one_minus_fidelity"] = 1.0 - abs(ov) ** 2 / (n64 * n128)
    out["norm_c64"] = n64
    out["mean_abs_zz_error"] = float(np.mean(np.abs(zz["single"] - zz["double"])))
    out["c64_over_c128"] = out["single"]["gates_per_s"] / out["double"]["gates_per_s"]
    for qv in states.values():
        qv.close()
    return out


def run_batched(num_circuits: int, terms: int, seed: int) -> dict:
    """num_circuits hardware-efficient ansatz circuits, n ~ U{10..20} qubits,
    reps ~ U{2..8}, each with a `terms`-term random Pauli sum: circuits/s for one
    batched launch (complex128 and complex64) and circuit by circuit."""
    from qiskit_aer_amd import circuits
    from qiskit_aer_amd.batch import batch_pauli_expval
    rng = np.random.default_rng(seed)
    cs, hs = [], []
    for i in range(num_circuits):
        n = int(rng.integers(10, 21))
        c, _ = circuits.hea(n, int(rng.integers(2, 9)), seed + i)
        cs.append(c)
        hs.append(circuits.random_pauli_sum(n, terms, seed + 7919 + i))
    out = {"circuits": num_circuits, "terms": terms, "gates": sum(c.count_gates() for c in cs)}
    energies = {}
    for prec, mode in (("double", "batch"), ("single", "batch"), ("double", "sequential")):
        t0 = time.perf_counter()
        e, ms = batch_pauli_expval(cs, hs, prec, mode)
        wall = time.perf_counter() - t0
        energies[(prec, mode)] = e
        out[f"{mode}_{'c128' if prec == 'double' else 'c64'}"] = {
            "circuits_per_s": num_circuits / wall, "wall_s": wall, "kernel_ms": ms}
    out["max_abs_energy_diff_batch_vs_sequential"] = float(
        np.max(np.abs(energies[("double", "batch")] - energies[("double", "sequential")])))
    out["max_abs_energy_diff_c64_vs_c128"] = float(
        np.max(np.abs(energies[("single", "batch")] - energies[("double", "batch")])))
    out["batch_over_sequential"] = (out["batch_c128"]["circuits_per_s"]
                                    / out["sequential_c128"]["circuits_per_s"])
    return out


def cpu_baseline(n: int, layers: int, seed: int) -> dict:
    """The reference's CPU path on the first `layers` layers of the same
    brickwork: the oracle's restated driver and Aer fusion (max 5 qubits, the
    AVX2 cost table) with every fused matrix and diagonal applied by the
    reference's own compiled Transformer / TransformerAVX2 (oracle/_ref, picked
    by the reference's CPUID check), OpenMP on the host cores."""
    from oracle import pyoracle as O
    from qiskit_aer_amd import circuits
    c = circuits.brickwork_u3_cx(n, layers, seed, measure=False)
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    kind, path = "port", "restated generic transformer"
    if O.ref() is not None:
        with O.reference_transformer() as rt:
            _, _, timing = O.run_circuit(c, shots=0, seed=0, threads=threads, statevector=False)
        kind = "reference"
        path = "reference TransformerAVX2" if rt.avx2 else "reference Transformer (no AVX2 on this host)"
    else:
        _, _, timing = O.run_circuit(c, shots=0, seed=0, threads=threads, statevector=False)
    return {"value": c.count_gates() / timing[0], "unit": "gates/s", "cores": threads, "kind": kind,
            "sample": f"{n}q c128 brickwork, first {layers} of 40 layers ({c.count_gates()} gates), "
                      f"Aer fusion on (max 5), {path}, gate phase {timing[0]:.2f} s"}


def profile_summary(trace_db: str, *pmc_dbs: str) -> str:
    """Text summary of rocprofv3 output databases: per-kernel calls and average
    duration (kernel trace), and every counter of the PMC runs per dispatch
    (average over the kernel's dispatches)."""
    import collections
    import sqlite3
    out = []
    cur = sqlite3.connect(trace_db).cursor()
    # rocprofv3's top_kernels view reports durations in microseconds.
    out.append("| kernel | calls | total ms | avg ms | % |")
    out.append("|---|---|---|---|---|")
    for name, calls, total, avg, pct in cur.execute(
            "select name, total_calls, total_duration, average, percentage from top_kernels"):
        out.append(f"| `{name.split('(')[0]}` | {calls} | {total / 1e3:.1f} | {avg / 1e3:.3f} | {pct:.2f} |")
    for db in pmc_dbs:
        if not db:
            continue
        c = sqlite3.connect(db).cursor()
        agg = collections.defaultdict(lambda: collections.defaultdict(list))
        for name, counter, v in c.execute(
                "select kernel_name, counter_name, value from counters_collection"):
            agg[counter][name.split("(")[0]].append(v)
        for counter, per_kernel in sorted(agg.items()):
            out.append("")
            out.append(f"| kernel | dispatches | {counter} per dispatch (avg) |")
            out.append("|---|---|---|")
            for k, vs in per_kernel.items():
                out.append(f"| `{k}` | {len(vs)} | {sum(vs) / len(vs):.1f} |")
    return "\n".join(out)


def main() -> int:
    if len(sys.argv) > 1 and sys.argv[1] == "--profile-summary":
        print(profile_summary(sys.argv[2], *sys.argv[3:]))
        return 0
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--adjoint-batch", type=int, default=64)
    ap.add_argument("--mixed-qubits", type=int, default=32)
    ap.add_argument("--batch-circuits", type=int, default=2048)
    ap.add_argument("--big-qubits", type=int, default=33, help="0 skips the large-state leg")
    ap.add_argument("--leg", choices=("qft", "c64", "su4"), default=None,
                    help="run only this leg (for profiling it) and print its figures")
    ap.add_argument("--precision", choices=("double", "single"), default="double",
                    help="precision of the su4 leg")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the block-fusion autotuner of the brickwork legs (use --fusion-width)")
    ap.add_argument("--fusion-width", type=int, default=0,
                    help="block-fusion width of the su4 and c64 legs, and of the brickwork with "
                         "--no-tune (0: off, 2..5)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the statevector is not sharded across devices.", file=sys.stderr)
        return 2

    from qiskit_aer_amd import circuits
    n, seed = args.qubits, 2024
    if args.leg == "qft":
        q = circuits.qft(n)
        res = run_engine(n, "double", q, args.steps, args.warmup,
                         init=lambda qv: qv.initialize_random(seed))
        res["gates"] = q.count_gates()
        print(json.dumps({"leg": f"qft_{n}q_c128", **res}), flush=True)
        return 0
    if args.leg == "su4":
        from qiskit_aer_amd import QubitVector
        nq = args.mixed_qubits
        c = circuits.su4_brickwork(nq, 20, seed)
        qv = QubitVector(nq, args.precision)
        qv.set_block_fusion(args.fusion_width, cost_check=True)
        res = {}
        for step in range(args.warmup + args.steps):
            qv.initialize()
            qv.synchronize()
            t0 = time.perf_counter()
            qv.apply_ops(c)
            qv.synchronize()
            res["ms"] = 1e3 * (time.perf_counter() - t0)
        res.update({"fused_gates": qv.fused_count(), "gates": c.count_gates(), "num_qubits": nq,
                    "precision": args.precision, "fusion_width": args.fusion_width})
        qv.close()
        print(json.dumps({"leg": "su4_brickwork", **res}), flush=True)
        return 0
    circ = circuits.brickwork_u3_cx(n, args.layers, seed, measure=False)
    if args.leg == "c64":
        res = run_engine(n, "single", circ, args.steps, args.warmup, fusion_width=args.fusion_width)
        print(json.dumps({"leg": f"brickwork_{n}q_depth{args.layers}_c64", **res}), flush=True)
        return 0
    log(f"bench: brickwork {n}q depth {args.layers} c128, {circ.count_gates()} gates")
    # Block-fusion width per precision (0 = off, 2..5 qubits per fused gate,
    # the configured fusion limit being 5), picked by the autotuner on the
    # circuit itself before the timed runs.
    from qiskit_aer_amd.autotune import tune
    widths = {}
    for prec in ("double",) if args.no_extra else ("double", "single"):
        if args.no_tune:
            widths[prec] = {"best_width": args.fusion_width, "ms_by_width": {}}
            continue
        best, times = tune(circ, prec, widths=(0, 2, 3, 4, 5))
        widths[prec] = {"best_width": best, "ms_by_width": times}
        log(f"bench: fusion autotune {prec}: width {best}, " +
            ", ".join(f"{w}: {t:.0f} ms" for w, t in times.items()))
    c128 = run_engine(n, "double", circ, args.steps, args.warmup,
                      fusion_width=widths["double"]["best_width"])
    # FP64 work of the source circuit: each U3 costs 8 fused multiply-adds per
    # amplitude (a complex 2x2 on each pair), CX only moves amplitudes.  Block
    # fusion executes less: a 4x4 gate costs 16 per amplitude and replaces up to
    # four U3s with the CX between them.
    u3_per_step = n * args.layers
    fp64_tflops = u3_per_step * (1 << n) * 16 / (c128["ms_per_step"] / 1e3) / 1e12
    log(f"bench: c128 {c128['gates_per_s']:.1f} gates/s, {c128['hbm_gbs']:.0f} GB/s, "
        f"kernel {c128['kernel_gbs']:.0f} GB/s, {c128['passes_per_step']:.0f} passes, "
        f"{c128['gathers_per_step']:.0f} register layouts per step")
    extra = {"fusion_autotune": widths}
    if not args.no_extra and widths["double"]["best_width"]:
        # the same step without block fusion: more passes, each lighter
        u128 = run_engine(n, "double", circ, args.steps, args.warmup)
        extra["c128_unfused"] = {k: u128[k] for k in ("gates_per_s", "ms_per_step", "kernel_gbs",
                                                       "passes_per_step")}
        extra["c128_unfused"]["roofline_frac"] = u128["kernel_gbs"] / HBM_PEAK_GBS
    if not args.no_extra:
        w64 = widths["single"]["best_width"]
        c64 = run_engine(n, "single", circ, args.steps, args.warmup, fusion_width=w64)
        log(f"bench: c64 {c64['gates_per_s']:.1f} gates/s, {c64['hbm_gbs']:.0f} GB/s (fusion width {w64})")
        extra["c64"] = {k: c64[k] for k in ("gates_per_s", "ms_per_step", "hbm_gbs", "kernel_gbs",
                                             "passes_per_step")}
        extra["c64"]["fusion_width"] = w64
        extra["c64_over_c128"] = c64["gates_per_s"] / c128["gates_per_s"]
        if w64:
            u64 = run_engine(n, "single", circ, args.steps, args.warmup)
            extra["c64_unfused"] = {k: u64[k] for k in ("gates_per_s", "ms_per_step", "kernel_gbs",
                                                         "passes_per_step")}
        q = circuits.qft(n)

        def gaussian(qv):
            qv.initialize_random(seed)  # i.i.d. complex Gaussian, normalised, on the device
        qres = run_engine(n, "double", q, max(1, args.steps), 1, init=gaussian)
        log(f"bench: qft {qres['gates_per_s']:.1f} gates/s")
        extra["qft_c128"] = {k: qres[k] for k in ("gates_per_s", "ms_per_step", "hbm_gbs",
                                                   "passes_per_step")}
        extra["qft_c128"]["gates"] = q.count_gates()
        # Single-gate sweeps: one U3 per pass on a low and on the top qubit,
        # the HBM-bound case (one read and one write of every amplitude).
        for tq in (0, n - 1):
            sw = circuits.Circuit(n)
            sw.gate("u3", [tq], [0.3, 0.2, 0.1])
            sres = run_engine(n, "double", sw, 20, 2)
            extra[f"sweep_u3_q{tq}_c128"] = {
                "kernel_gbs": sres["kernel_gbs"], "frac_of_peak": sres["kernel_gbs"] / HBM_PEAK_GBS,
                "kernel_ms": sres["kernel_ms_avg"]}
            log(f"bench: single-gate sweep q{tq} c128 {sres['kernel_gbs']:.0f} GB/s "
                f"({100 * sres['kernel_gbs'] / HBM_PEAK_GBS:.1f} % of peak)")
        adj = run_adjoint(24, 12, args.adjoint_batch, 16, seed)
        log(f"bench: adjoint 24q {adj['grads_per_s']:.0f} grads/s, param shift "
            f"{adj['param_shift_grads_per_s']:.1f} grads/s ({adj['speedup_vs_param_shift']:.0f}x)")
        extra["adjoint_24q_c128"] = adj
        mp = run_mixed_precision(args.mixed_qubits, 20, seed)
        log(f"bench: su4 {args.mixed_qubits}q c128 {mp['double']['gates_per_s']:.0f} gates/s, c64 "
            f"{mp['single']['gates_per_s']:.0f} gates/s, 1-F {mp['one_minus_fidelity']:.2e}, "
            f"|dZZ| {mp['mean_abs_zz_error']:.2e}")
        extra[f"su4_brickwork_{args.mixed_qubits}q"] = mp
        if args.big_qubits:
            big = circuits.brickwork_u3_cx(args.big_qubits, args.layers, seed, measure=False)
            br = run_engine(args.big_qubits, "double", big, 1, 0,
                            fusion_width=widths["double"]["best_width"])
            log(f"bench: brickwork {args.big_qubits}q c128 {br['gates_per_s']:.1f} gates/s, "
                f"kernel {br['kernel_gbs']:.0f} GB/s")
            extra[f"brickwork_{args.big_qubits}q_c128"] = {
                k: br[k] for k in ("gates_per_s", "ms_per_step", "hbm_gbs", "kernel_gbs", "passes_per_step")}
            extra[f"brickwork_{args.big_qubits}q_c128"]["fusion_width"] = widths["double"]["best_width"]
        bt = run_batched(args.batch_circuits, 64, seed)
        log(f"bench: batch {bt['circuits']} circuits: one launch {bt['batch_c128']['circuits_per_s']:.0f} "
            f"circuits/s (c64 {bt['batch_c64']['circuits_per_s']:.0f}), circuit by circuit "
            f"{bt['sequential_c128']['circuits_per_s']:.0f} circuits/s")
        extra["batched_hea_circuits"] = bt

    cpu = None
    if not args.no_cpu:
        log("bench: cpu baseline")
        cpu = cpu_baseline(n, 4, seed)
        log(f"bench: cpu {cpu['value']:.2f} gates/s on {cpu['cores']} cores")

    line = {
        "metric": "gates/s + HBM GB/s at 30q (c64 vs c128); adjoint grads/s vs param-shift, 24q",
        "value": c128["gates_per_s"],
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": c128["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128",
        "data": "synthetic (seeded random U3 angles)",
        "config": {"workload": f"brickwork_u3_cx_{n}q_depth{args.layers}", "num_qubits": n,
                   "layers": args.layers, "gates": circ.count_gates(), "precision": "c128",
                   "fusion_width": widths["double"]["best_width"], "parallelism": "single GPU"},
        "hbm_gbs": c128["hbm_gbs"],
        "passes_per_step": c128["passes_per_step"],
        "roofline": {"bound": "hbm", "achieved": c128["kernel_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": c128["kernel_gbs"] / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC_30Q_C128 if (n, args.layers) == (30, 40) else None,
                     "algorithmic_bytes_per_launch": c128["sweep_bytes"]},
        # the same step against the FP64 vector peak (the pass's other bound)
        "fp64_compute": {"achieved": fp64_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": fp64_tflops / FP64_PEAK_TFLOPS,
                         "flops_per_step": u3_per_step * (1 << n) * 16},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
