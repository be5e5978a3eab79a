#!/usr/bin/env python3
"""Throughput of the state-vector simulator on one MI355X; prints ONE JSON line.

Workload: the 32-qubit depth-100 random circuit (per layer: H on every qubit, one random
Rx/Ry/Rz per qubit with angle U[0, 2pi), 16 CNOTs on random distinct pairs), complex64 state,
gates queued in a window and partitioned by the planner.  A step is five circuit layers
(400 gates), so the default `--steps 20 --warmup 5` times the full 100-layer circuit after
25 warm-up layers, with the state already resident in HBM.

Also reported:
* the same 125 layers in complex128 (window fusion, same planner), timed on the same 100 layers,
  and the infidelity 1 - |<psi_c128|psi_c64>|^2 of the complex64 result against it;
* plain complex128 path (the reference's default: one pass per gate, run() after every gate)
  timed on a few layers of the same circuit, and the speed-up of the headline over it;
* roofline of the dominant kernel (the gate pass): algorithmic bytes / summed HIP-event time;
* a CPU baseline: ProjectQ's own C++ simulator built from its sources (oracle/_ref), or the
  in-repo oracle where that build is absent, on a bounded sample of the same circuit;
* batched_variational: 512 variational ansatze with 100-term Pauli energies in one hipGraph launch;
* qft_rotations_adder_30q: the 30-qubit QFT + rotations + AddQuantum + probabilities + measurement
  workload in complex128, with bit-exact and normalisation checks;
* default_engine_list_20q: the 20-qubit random-layer circuit written as a ProjectQ program and
  compiled by the default engine list in front of the GPU simulator (host compiler time included).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "gates/s + % of 8 TB/s HBM at 32q; c64+learned fusion vs plain c128 on GPU"
HBM_PEAK_GBS = 8000.0
CNOTS_PER_LAYER = 16
LAYERS_PER_STEP = 5


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def layer_slices(cmds, n):
    """Split random_layers() output into (allocation commands, per-layer gate lists)."""
    alloc = cmds[:n]
    gates = cmds[n:]
    per = 2 * n + CNOTS_PER_LAYER
    return alloc, [gates[i : i + per] for i in range(0, len(gates), per)]


def submit(sim, layer, run_each):
    for _, m, targets, controls in layer:
        sim.apply_controlled_gate(m, list(targets), list(controls))
        if run_each:
            sim.run()


def gpu_window_run(args, cmds, n, precision, label):
    """Warm-up layers, then the timed layers, in window fusion; returns the simulator (state kept
    for the infidelity check), the timed wall seconds, the gate count and the pass statistics."""
    from projectq_amd import Simulator

    _, layers = layer_slices(cmds, n)
    warm, steps = args.warmup * LAYERS_PER_STEP, args.steps * LAYERS_PER_STEP
    sim = Simulator(args.seed, precision=precision, fusion="window", planner=args.planner, capacity_qubits=n)
    for q in range(n):
        sim.allocate_qubit(q)
    for layer in layers[:warm]:
        submit(sim, layer, run_each=False)
    sim.run()
    sim.synchronize()
    log(f"{label}: warm-up {warm} layers done")
    sim.reset_stats()
    sim.set_profiling(True)
    timed = layers[warm : warm + steps]
    t0 = time.perf_counter()
    for layer in timed:
        submit(sim, layer, run_each=False)
    sim.run()
    sim.synchronize()
    t1 = time.perf_counter()
    sim.set_profiling(False)
    stats = sim.stats()
    stats["plan"] = sim.last_plan()
    n_gates = sum(len(layer) for layer in timed)
    log(f"{label}: {len(timed)} layers, {n_gates} gates in {t1 - t0:.3f}s, {stats['passes']} passes, plan {stats['plan']}")
    return sim, t1 - t0, n_gates, stats


def gpu_plain_c128(args, cmds, n):
    from projectq_amd import Simulator

    _, layers = layer_slices(cmds, n)
    sim = Simulator(args.seed, precision="c128", fusion="reference", capacity_qubits=n)
    for q in range(n):
        sim.allocate_qubit(q)
    submit(sim, layers[0][: 2 * n], run_each=True)  # warm-up: one layer's single-qubit gates
    sim.synchronize()
    sim.reset_stats()
    sim.set_profiling(True)
    k = args.plain_layers
    t0 = time.perf_counter()
    for layer in layers[1 : 1 + k]:
        submit(sim, layer, run_each=True)
    sim.synchronize()
    t1 = time.perf_counter()
    stats = sim.stats()
    log(f"plain c128: {k} layers in {t1 - t0:.3f}s ({stats['passes']} passes)")
    del sim
    return (t1 - t0) / k, stats


def cpu_baseline(args):
    """ProjectQ's C++ simulator (oracle/_ref) on a bounded sample: `cpu_layers` layers of the same
    circuit family at `cpu_qubits` qubits, once with gate_fusion off (the reference default: run()
    after every gate) and once with gate_fusion on (its greedy min-4/max-5 blocks).  Per-gate cost
    scales with 2^n, so gates/s is also given extrapolated to the headline's qubit count; `value`
    is the faster of the two modes."""
    from oracle import ref

    from projectq_amd import circuits, commands

    mod = ref.load("intrin")
    kind = "reference"
    if mod is None:
        from oracle.cpu_sim import OracleSimulator as make

        kind = "port"
    else:
        make = mod.Simulator
    nq = args.cpu_qubits
    cmds = circuits.random_layers(nq, args.cpu_layers, CNOTS_PER_LAYER, args.seed, measure=False)
    gates = cmds[nq:]
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    scale = 2.0 ** (nq - args.qubits)
    modes = {}
    for fusion in (False, True):
        sim = make(args.seed)
        commands.execute(sim, cmds[:nq])
        t0 = time.perf_counter()
        commands.execute(sim, gates, gate_fusion=fusion)
        sim.run()
        t1 = time.perf_counter()
        del sim
        rate = len(gates) / (t1 - t0)
        key = "fusion_on" if fusion else "fusion_off"
        modes[key] = {"measured_gates_per_s": round(rate, 3), "seconds": round(t1 - t0, 3),
                      "extrapolated_gates_per_s": round(rate * scale, 4)}
        log(f"cpu baseline ({kind}, gate_fusion={fusion}): {len(gates)} gates at {nq} qubits in {t1 - t0:.2f}s")
    best = max(modes, key=lambda k: modes[k]["measured_gates_per_s"])
    return {
        "value": modes[best]["extrapolated_gates_per_s"],
        "unit": "gates/s",
        "cores": cores,
        "kind": kind,
        "sample": (
            f"{args.cpu_layers} layers ({len(gates)} gates) of the same random-circuit family at {nq} qubits, "
            f"complex128, OpenMP {cores} threads, gate_fusion off and on; value = the faster mode ({best}), "
            f"extrapolated by 2^({nq}-{args.qubits}) to {args.qubits} qubits"
        ),
        "measured_qubits": nq,
        **modes,
    }


def batched_variational(args):
    """512 hardware-efficient ansatze, qubit count uniform in {10..20}, 6 layers of Ry, Rz plus a
    linear CNOT chain; <H> of a 100-string random Pauli operator per circuit; complex64 state with
    FP64 sums; one hipGraph launch per run; tile shape autotuned.  Accuracy against the same batch
    in complex128."""
    import numpy as np

    from projectq_amd import circuits
    from projectq_amd.batch import BatchExpectation

    rng = np.random.default_rng(args.seed + 3)
    sizes = [int(x) for x in rng.integers(10, 21, size=args.batch)]
    circ, ops = [], []
    for i, n in enumerate(sizes):
        cmds = circuits.hardware_efficient_ansatz(n, 6, args.seed + 10 + i)
        circ.append((n, [c for c in cmds if c[0] == "gate"]))
        ops.append(circuits.random_pauli_operator(n, 100, args.seed + 5000 + i))
    n_gates = sum(len(c) for _, c in circ)
    t0 = time.perf_counter()
    b = BatchExpectation(circ, ops, precision="c64", autotune=True)
    t_build = time.perf_counter() - t0
    b.run()
    runs = []
    for _ in range(5):
        t0 = time.perf_counter()
        energies, gpu_ms = b.run()
        runs.append((time.perf_counter() - t0, gpu_ms))
    wall = min(r[0] for r in runs)
    gpu = min(r[1] for r in runs)
    ref, _ = BatchExpectation(circ, ops, precision="c128", autotune=False).run()
    scale = np.array([sum(abs(c) for _, c in t) for t in ops])
    log(f"batch: {len(circ)} circuits, {n_gates} gates, {gpu:.2f} ms GPU per graph launch, shape {b.shape()}")
    return {
        "workload": f"{len(circ)} hardware-efficient ansatze, n in U{{10..20}}, 6 layers Ry,Rz + CNOT chain, "
        f"<H> of 100 random Pauli strings each; complex64, FP64 sums, one hipGraph launch",
        "circuits_per_s": round(len(circ) / wall, 1),
        "gates_per_s": round(n_gates / wall, 1),
        "pauli_terms_per_s": round(100 * len(circ) / wall, 1),
        "ms_per_batch_gpu": round(gpu, 3),
        "ms_per_batch_wall": round(1e3 * wall, 3),
        "autotuned_shape": b.shape(),
        "plan_and_tune_s": round(t_build, 3),
        "max_rel_error_vs_c128": float(np.max(np.abs(energies - ref) / scale)),
    }


def qft_rotations_adder(args, n=30):
    """The 30-qubit QFT + rotations + AddQuantum workload, complex128 (the reference's precision):
    QFT (30 H + 435 controlled phases), 50 layers of a random Rx/Ry/Rz per qubit plus 15
    controlled-Rz with 1-3 random controls, b <- (a + b) mod 2^15 on the two 15-qubit halves through
    emulate_math, 8 probabilities of 8-qubit patterns, seeded measurement of all 30 qubits.
    Window fusion with the learned planner.  Checks that need no reference at this size: the
    adder moves 64 sampled amplitudes bit-exactly to (a, (a + b) mod 2^15), and P(q=0) + P(q=1) = 1.
    The plain path (run() after every gate) is timed on the first 60 gates and scaled."""
    import numpy as np

    from projectq_amd import Simulator, circuits

    cmds, queries = circuits.qft_rotations_adder(n, 50, 15, args.seed, n_probabilities=8, prob_qubits=8, measure=False)
    gates = [c for c in cmds if c[0] == "gate"]
    math = [c for c in cmds if c[0] == "math"][0]
    half = n // 2
    ids = list(range(n))

    plain = Simulator(args.seed, precision="c128", fusion="reference", capacity_qubits=n)
    for q in ids:
        plain.allocate_qubit(q)
    plain.synchronize()
    sample = gates[:60]
    t0 = time.perf_counter()
    for _, m, t, c in sample:
        plain.apply_controlled_gate(m, list(t), list(c))
        plain.run()
    plain.synchronize()
    plain_per_gate = (time.perf_counter() - t0) / len(sample)
    del plain

    sim = Simulator(args.seed, precision="c128", fusion="window", planner="learned", capacity_qubits=n)
    for q in ids:
        sim.allocate_qubit(q)
    sim.synchronize()
    t0 = time.perf_counter()
    for _, m, t, c in gates:
        sim.apply_controlled_gate(m, list(t), list(c))
    sim.run()
    sim.synchronize()
    t_gates = time.perf_counter() - t0
    plan = sim.last_plan()
    log(f"qft+rotations+adder: {len(gates)} gates in {t_gates:.3f}s ({plan})")

    rng = np.random.default_rng(args.seed + 1)
    picks = rng.integers(0, 2**n, size=64, dtype=np.int64)
    bits_of = lambda v: [(int(v) >> q) & 1 for q in ids]  # noqa: E731
    before = [sim.get_amplitude(bits_of(v), ids) for v in picks]
    t0 = time.perf_counter()
    _, fn, regs, ctrl = math
    sim.emulate_math(fn, regs, ctrl)
    sim.synchronize()
    t_add = time.perf_counter() - t0
    mask = (1 << half) - 1
    moved = [(v & mask) | ((((v & mask) + (v >> half)) & mask) << half) for v in picks]
    after = [sim.get_amplitude(bits_of(v), ids) for v in moved]
    adder_exact = all(a == b for a, b in zip(before, after))

    t0 = time.perf_counter()
    probs = [sim.get_probability(bits, qids) for bits, qids in queries]
    t_prob = time.perf_counter() - t0
    total = sim.get_probability([False], [0]) + sim.get_probability([True], [0])

    t0 = time.perf_counter()
    outcomes = [bool(sim.measure_qubits([q])[0]) for q in ids]
    t_meas = time.perf_counter() - t0
    del sim
    t_all = t_gates + t_add + t_prob + t_meas
    return {
        "workload": f"{n} qubits complex128: QFT + 50 layers (random Rx/Ry/Rz per qubit + 15 controlled-Rz, "
        f"1-3 controls) + AddQuantum(15, 15) + 8 probabilities + measure all; window fusion, learned planner",
        "gates": len(gates),
        "gates_s": round(t_gates, 4),
        "gates_per_s": round(len(gates) / t_gates, 1),
        "plan": plan,
        "add_quantum_s": round(t_add, 4),
        "probabilities_s": round(t_prob, 4),
        "measure_all_s": round(t_meas, 4),
        "total_s": round(t_all, 4),
        "plain_c128_per_gate_ms": round(1e3 * plain_per_gate, 3),
        "speedup_gates_vs_plain_c128": round(plain_per_gate * len(gates) / t_gates, 2),
        "adder_moves_amplitudes_exactly": adder_exact,
        "total_probability_minus_1": total - 1.0,
        "probabilities": [round(p, 12) for p in probs],
        "outcomes": "".join("1" if b else "0" for b in outcomes),
    }


def default_engine_list(args, n=20, layers=20):
    """The 20-qubit, 20-layer random-layer circuit (10 CNOTs per layer, measure all) written as a
    ProjectQ program and compiled by the default engine list (projectq_amd.compiler) in front of
    the GPU simulator: once as the reference runs it by default (complex128, run() after every
    gate) and once with gate_fusion=True into complex64 window fusion with the learned planner.
    Then the headline circuit (all its layers) through the same front end into complex64 window
    fusion.  Times include the host-side compiler."""
    from projectq_amd import Simulator, circuits
    from projectq_amd import compiler as C

    ops = circuits.random_layers_program(n, layers, 10, args.seed)
    n_gates = sum(1 for op in ops if op[0] == "gate")
    out = {"workload": f"{n} qubits, {layers} layers (H, random Rx/Ry/Rz, 10 CNOTs), measure all; "
           "default engine list (TagRemover, LocalOptimizer(10), AutoReplacer, TagRemover, LocalOptimizer(10))",
           "gates": n_gates}
    for label, kw, fusion in (("c128_run_each_gate", dict(precision="c128"), False),
                              ("c64_window_learned", dict(precision="c64", fusion="window", planner="learned"), True)):
        sim = Simulator(args.seed, capacity_qubits=n, **kw)
        calls = {"apply_controlled_gate": 0}

        class Counting:
            def __getattr__(self, name):
                fn = getattr(sim, name)
                if name == "apply_controlled_gate":
                    def counted(*a):
                        calls["apply_controlled_gate"] += 1
                        return fn(*a)
                    return counted
                return fn

        t0 = time.perf_counter()
        eng = C.run_program(ops, C.SimulatorEngine(Counting(), gate_fusion=fusion))
        sim.synchronize()
        dt = time.perf_counter() - t0
        out[label] = {"s": round(dt, 4), "gates_per_s": round(n_gates / dt, 1),
                      "backend_gate_calls": calls["apply_controlled_gate"],
                      "outcomes": "".join("1" if eng.results[q] else "0" for q in range(n))}
        del eng, sim

    # the headline circuit (32 qubits, 100 layers, 16 CNOTs, measure all) through the same front end
    big_n, big_layers = args.qubits, args.steps * LAYERS_PER_STEP
    ops = circuits.random_layers_program(big_n, big_layers, CNOTS_PER_LAYER, args.seed)
    big_gates = sum(1 for op in ops if op[0] == "gate")
    sim = Simulator(args.seed, precision="c64", fusion="window", planner="learned", capacity_qubits=big_n)
    backend = C.SimulatorEngine(sim, gate_fusion=True)
    t0 = time.perf_counter()
    eng = C.MainEngine(backend)
    C.run_program(ops[:-big_n], backend, engine=eng)  # all but the measurements, then a flush
    t_host = time.perf_counter() - t0  # compiler + planner + (asynchronous) pass launches
    sim.synchronize()
    t_gates = time.perf_counter() - t0
    for q in range(big_n):
        eng.measure(q)
    eng.flush()
    sim.synchronize()
    t_all = time.perf_counter() - t0
    out[f"c64_window_learned_{big_n}q_{big_layers}_layers"] = {
        "gates": big_gates, "compile_plan_launch_s": round(t_host, 3), "gates_s": round(t_gates, 3),
        "gates_per_s": round(big_gates / t_gates, 1), "with_measure_all_s": round(t_all, 3),
        "plan": sim.last_plan()}
    del eng, backend, sim
    log(f"default engine list: {out}")
    return out


def kernel_name(args, plan):
    """The gate-pass kernel the plan ran: the tile launcher uses the software-pipelined variant for
    complex64 tiles of 2^12 amplitudes with 4 register bits."""
    if not plan or plan.get("kind") != "tile":
        return "k_apply (dense fused-block gate pass)"
    if args.precision == "c64" and plan.get("tile_bits") == 12 and plan.get("reg_bits") == 4:
        return "k_tile_pass_pf (software-pipelined LDS tile pass)"
    return "k_tile_pass (LDS tile pass)"


def calibrate_copy(nbytes=1 << 31):
    """One device-to-device copy of a known size, so hardware byte counters (FETCH_SIZE,
    WRITE_SIZE) read over the same run can be scaled: the copy reads and writes nbytes."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise MemoryError("calibration buffers")
    hip.hipMemset(a, 0, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"calibration copy: {nbytes} bytes read and written")


def stream_ceiling(sim, repeats=3):
    """The practical streaming rates a pass kernel is measured against, on the headline state:
    collapse_wavefunction([], []) runs the norm reduction (a read-only pass) and the collapse
    kernel with an empty mask (an in-place read + write pass with one multiply per amplitude);
    plus a 2 GiB device-to-device copy.  Kernel times come from the profiler's trace."""
    for _ in range(repeats):
        t0 = time.perf_counter()
        sim.collapse_wavefunction([], [])
        sim.synchronize()
        log(f"stream ceiling: norm pass + in-place pass in {1e3 * (time.perf_counter() - t0):.2f} ms (wall)")
    for _ in range(repeats):
        calibrate_copy()


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_tile_pass_pf_random32_c64_traffic.json")


def measured_traffic(args, n):
    """HBM bytes per tile-pass launch from the committed FETCH_SIZE/WRITE_SIZE counter runs
    (scaled by a 2 GiB copy measured in the same runs); only for the workload they measured."""
    if n != 32 or args.precision != "c64" or args.planner == "greedy" or not os.path.exists(TRAFFIC_PROFILE):
        return None, None
    with open(TRAFFIC_PROFILE) as f:
        rec = json.load(f)
    return int(round(rec["traffic_bytes_per_launch"])), os.path.relpath(TRAFFIC_PROFILE, ROOT)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help=f"timed steps of {LAYERS_PER_STEP} circuit layers")
    ap.add_argument("--warmup", type=int, default=5, help=f"untimed steps of {LAYERS_PER_STEP} layers before the timed ones")
    ap.add_argument("--qubits", type=int, default=32)
    ap.add_argument("--precision", default="c64", choices=["c64", "c128"])
    ap.add_argument("--planner", default="learned", choices=["greedy", "tile", "learned"])
    ap.add_argument("--plain-layers", type=int, default=2)
    ap.add_argument("--cpu-qubits", type=int, default=28)
    ap.add_argument("--cpu-layers", type=int, default=3)
    ap.add_argument("--no-c128-window", action="store_true", help="skip the complex128 window run and the infidelity")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--batch", type=int, default=512, help="circuits in the batched-variational extra")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true", help="add a known-size D2D copy (counter scaling)")
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="after the headline, run read-only and in-place streaming passes and 2 GiB copies (profile them)")
    ap.add_argument("--no-qft-adder", action="store_true", help="skip the 30-qubit QFT + rotations + AddQuantum extra")
    ap.add_argument("--no-engine-list", action="store_true", help="skip the default-engine-list extra")
    args = ap.parse_args()

    if args.gpus != 1:
        print(
            "bench.py measures one GPU: this build has no multi-GPU path (the state lives in one GPU's HBM); "
            "run with --gpus 1",
            file=sys.stderr,
        )
        sys.exit(2)

    from projectq_amd import circuits

    n = args.qubits
    total_layers = (args.warmup + args.steps) * LAYERS_PER_STEP
    cmds = circuits.random_layers(n, max(total_layers, args.plain_layers + 1), CNOTS_PER_LAYER, args.seed, measure=False)
    log(f"circuit: {n} qubits, {total_layers} layers, {CNOTS_PER_LAYER} CNOTs/layer")

    sim, wall, n_gates, stats = gpu_window_run(args, cmds, n, args.precision, "headline")
    if args.calibrate_copy:
        calibrate_copy()
    if args.stream_ceiling:
        stream_ceiling(sim)
    value = n_gates / wall
    ms_per_step = 1e3 * wall / args.steps

    accuracy = None
    if not args.no_c128_window and args.precision == "c64":
        ref_sim, ref_wall, _, ref_stats = gpu_window_run(args, cmds, n, "c128", "c128 window")
        ov = ref_sim.overlap(sim)
        n_ref, n_c64 = ref_sim.overlap(ref_sim).real, sim.overlap(sim).real
        infid = 1.0 - abs(ov) ** 2
        total_p = ref_sim.get_probability([False], [0]) + ref_sim.get_probability([True], [0])
        log(f"infidelity of the complex64 state against complex128: {infid:.3e}")
        accuracy = {
            "infidelity_vs_c128": infid,
            "infidelity_vs_c128_normalized": 1.0 - abs(ov) ** 2 / (n_ref * n_c64),
            "c64_norm_minus_1": n_c64 - 1.0,
            "c128_window": {
                "gates_per_s": round(n_gates / ref_wall, 2),
                "ms_per_step": round(1e3 * ref_wall / args.steps, 3),
                "passes": ref_stats["passes"],
                "avg_pass_ms": ref_stats["gpu_ms"] / max(ref_stats["timed_passes"], 1),
                "plan": ref_stats["plan"],
                "total_probability_minus_1": total_p - 1.0,
            },
        }
        del ref_sim
    del sim

    traffic, traffic_src = measured_traffic(args, n)
    achieved = None
    if stats["timed_passes"] and stats["gpu_ms"] > 0:
        achieved = stats["bytes"] / (stats["gpu_ms"] * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
        "traffic": traffic,
        "traffic_source": traffic_src,
        "kernel": kernel_name(args, stats["plan"]),
        "passes": stats["passes"],
        "bytes_per_pass": stats["bytes"] / max(stats["passes"], 1),
        "avg_pass_ms": stats["gpu_ms"] / max(stats["timed_passes"], 1),
        "plan": stats["plan"],
    }

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision == "c64" else "fp64",
        "data": "synthetic (seeded random circuit)",
        "config": {
            "workload": (
                f"random circuit, {n} qubits, depth {args.steps * LAYERS_PER_STEP} timed "
                f"(+{args.warmup * LAYERS_PER_STEP} warm-up layers): per layer H and one "
                f"random Rx/Ry/Rz per qubit plus {CNOTS_PER_LAYER} CNOTs; {args.precision} state, window fusion, "
                f"{args.planner} planner"
            ),
            "qubits": n,
            "layers": args.steps * LAYERS_PER_STEP,
            "layers_per_step": LAYERS_PER_STEP,
            "warmup_layers": args.warmup * LAYERS_PER_STEP,
            "gates_per_layer": 2 * n + CNOTS_PER_LAYER,
            "precision": args.precision,
            "planner": args.planner,
            "parallelism": "single GPU",
        },
        "roofline": roofline,
    }
    if accuracy:
        out["infidelity_vs_c128"] = accuracy["infidelity_vs_c128"]
        out["infidelity_vs_c128_normalized"] = accuracy["infidelity_vs_c128_normalized"]
        out["c64_norm_minus_1"] = accuracy["c64_norm_minus_1"]
        out["c128_window"] = accuracy["c128_window"]
        out["speedup_vs_c128_window"] = round(value / accuracy["c128_window"]["gates_per_s"], 3)

    if not args.no_plain:
        plain_ms_layer, pstats = gpu_plain_c128(args, cmds, n)
        out["plain_c128"] = {
            "ms_per_layer": round(1e3 * plain_ms_layer, 3),
            "gates_per_s": round((2 * n + CNOTS_PER_LAYER) / plain_ms_layer, 2),
            "layers_timed": args.plain_layers,
            "achieved_gbs": round(pstats["bytes"] / (pstats["gpu_ms"] * 1e-3) / 1e9, 1) if pstats["gpu_ms"] else None,
        }
        out["speedup_vs_plain_c128"] = round((1e3 * plain_ms_layer * LAYERS_PER_STEP) / ms_per_step, 3)

    if not args.no_batch:
        out["batched_variational"] = batched_variational(args)

    if not args.no_qft_adder:
        out["qft_rotations_adder_30q"] = qft_rotations_adder(args)

    if not args.no_engine_list:
        out["default_engine_list_20q"] = default_engine_list(args)

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args)

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
