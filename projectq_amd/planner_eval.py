"""Evaluate the learned planner on depth-100 circuits of mixed 1/2/3-qubit gates (complex64).

    python -m projectq_amd.planner_eval --qubits 28 30 32 --out planner_eval.md

For each circuit the same gates run four ways on one GPU:
* unfused complex128, one pass per gate (the reference's default) -- the answer key;
* complex64 window mode with the greedy max-5 planner (the learned planner's fallback);
* complex64 window mode with the heuristic tile planner;
* complex64 window mode with the learned planner.
Reported: GPU time of the gate passes (HIP events), passes, the plan chosen, and the infidelity
1 - |<psi_c128|psi>|^2 of every complex64 result against the unfused complex128 state.
"""

import argparse
import sys
import time

from . import circuits
from ._cppsim import Simulator


def log(msg):
    print(f"[planner_eval] {msg}", file=sys.stderr, flush=True)


def run(n, gates, seed, **opts):
    sim = Simulator(seed, capacity_qubits=n, **opts)
    for q in range(n):
        sim.allocate_qubit(q)
    sim.run()
    sim.synchronize()
    sim.reset_stats()
    sim.set_profiling(True)
    window = opts.get("fusion") == "window"
    t0 = time.perf_counter()
    for _, m, targets, controls in gates:
        sim.apply_controlled_gate(m, list(targets), list(controls))
        if not window:
            sim.run()
    sim.run()
    sim.synchronize()
    wall = time.perf_counter() - t0
    return sim, wall, sim.stats()


def infidelity(ref, sim):
    ov = ref.overlap(sim)
    return 1.0 - abs(ov) ** 2 / (ref.overlap(ref).real * sim.overlap(sim).real)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--qubits", type=int, nargs="+", default=[28, 30, 32])
    ap.add_argument("--depth", type=int, default=100)
    ap.add_argument("--seed", type=int, default=777)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    rows = []
    for n in args.qubits:
        cmds = circuits.mixed_circuit(n, args.depth, args.seed + n)
        gates = [c for c in cmds if c[0] == "gate"]
        ref, ref_wall, ref_stats = run(n, gates, args.seed, precision="c128", fusion="reference")
        log(f"n={n}: {len(gates)} gates, unfused c128 {ref_stats['gpu_ms'] / 1e3:.2f}s GPU")
        rows.append((n, len(gates), "unfused c128 (reference default)", ref_stats, ref_wall, None, None))
        for planner in ("greedy", "tile", "learned"):
            sim, wall, st = run(n, gates, args.seed, precision="c64", fusion="window", planner=planner)
            inf = infidelity(ref, sim)
            plan = sim.last_plan()
            del sim
            log(f"n={n} {planner}: {st['gpu_ms'] / 1e3:.2f}s GPU, {st['passes']} passes, infidelity {inf:.2e}, plan {plan}")
            rows.append((n, len(gates), f"c64 {planner}", st, wall, inf, plan))
        del ref
    lines = [
        "| qubits | gates | path | passes | GPU s | wall s | gates/s (wall) | infidelity vs c128 | plan |",
        "|---:|---:|---|---:|---:|---:|---:|---:|---|",
    ]
    for n, g, name, st, wall, inf, plan in rows:
        plan_s = "" if plan is None else (
            f"{plan['kind']} t={plan['tile_bits']} rb={plan['reg_bits']} L={plan['low_bits']} {plan['rule']}" if plan["kind"] == "tile" else plan["kind"]
        )
        inf_s = "" if inf is None else f"{inf:.2e}"
        lines.append(
            f"| {n} | {g} | {name} | {st['passes']} | {st['gpu_ms'] / 1e3:.3f} | {wall:.3f} | {g / wall:.1f} | {inf_s} | {plan_s} |"
        )
    text = "\n".join(lines)
    print(text)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")


if __name__ == "__main__":
    main()
