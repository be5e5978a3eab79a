"""Train the learned planner's cost model on pass timings the simulator itself measures.

    python -m projectq_amd.planner_train --circuits 2000 --out projectq_amd/planner_mlp.npz

1. Generate seeded random circuits (24..30 qubits: mixed 1/2/3-qubit gates, random layers with
   CNOTs, or a QFT followed by layers of rotations and multi-controlled Rz), run each once in
   window mode with a random planner configuration (tile passes of a random shape under a random
   slot rule, with or without qubit relabelling, or greedy max-5 blocks), recording every pass:
   its features and its GPU time.
2. Train the 24 -> 64 -> 64 -> 1 MLP on log(time per 2^20 amplitudes) with the HIP training
   kernels (qs_mlp_train), holding out 10% of the circuits for validation.
3. Save the weights with the training statistics.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

from . import circuits
from ._binding import MLP, PASS_FEATURES
from ._cppsim import Simulator, library

TILE_SHAPES = {"c64": [(12, 4, 5), (12, 4, 6), (12, 4, 4), (11, 4, 5), (10, 4, 4), (12, 5, 5), (13, 4, 4), (13, 4, 5)],
               "c128": [(12, 4, 4), (12, 4, 5), (11, 4, 4), (10, 4, 4), (12, 5, 4)]}

DEFAULT_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "planner_mlp.npz")


def log(msg):
    print(f"[planner_train] {msg}", file=sys.stderr, flush=True)


def training_circuit(rng, n):
    u = rng.random()
    if u < 0.5:
        return circuits.mixed_circuit(n, int(rng.integers(3, 10)), int(rng.integers(1 << 30)))
    if u < 0.8:
        return circuits.random_layers(n, int(rng.integers(1, 4)), n // 2, int(rng.integers(1 << 30)), measure=False)
    # a QFT followed by rotation layers with multi-controlled Rz gates: long runs of controlled
    # phases, the diagonal-heavy passes the other two families rarely produce
    cmds, _ = circuits.qft_rotations_adder(n, int(rng.integers(0, 4)), n // 2, int(rng.integers(1 << 30)),
                                           n_probabilities=0, measure=False)
    return cmds


def generate(n_circuits, qmin, qmax, seed):
    """Run n_circuits circuits with random planner configurations; return features, targets,
    circuit index of every row."""
    rng = np.random.default_rng(seed)
    feats, targets, owner = [], [], []
    t0 = time.time()
    for i in range(n_circuits):
        n = int(rng.integers(qmin, qmax + 1))
        prec = "c64" if rng.random() < 0.75 else "c128"
        greedy = rng.random() < 0.25
        cmds = training_circuit(rng, n)
        sim = Simulator(int(rng.integers(1 << 31)), precision=prec, fusion="window",
                        planner="greedy" if greedy else "tile", capacity_qubits=n)
        if not greedy:
            sim.set_tile_params(*TILE_SHAPES[prec][int(rng.integers(len(TILE_SHAPES[prec])))])
            sim.set_slot_rule(Simulator.SLOT_RULES[int(rng.integers(len(Simulator.SLOT_RULES)))])
            sim.set_relabel("on" if rng.random() < 0.5 else "off")
        for q in range(n):
            sim.allocate_qubit(q)
        sim.run()
        sim.synchronize()
        sim.set_recording(True)
        for c in cmds:
            if c[0] == "gate":
                sim.apply_controlled_gate(c[1], list(c[2]), list(c[3]))
        sim.run()
        f, ms = sim.take_records()
        del sim
        keep = ms > 0
        f, ms = f[keep], ms[keep]
        feats.append(f)
        targets.append(np.log(ms.astype(np.float64)) - (n - 20) * np.log(2.0))
        owner.append(np.full(len(ms), i))
        if (i + 1) % 100 == 0:
            log(f"{i + 1}/{n_circuits} circuits, {sum(len(x) for x in targets)} passes, {time.time() - t0:.0f}s")
    return np.concatenate(feats), np.concatenate(targets).astype(np.float32), np.concatenate(owner)


def train(X, y, owner, seed, epochs, batch, lr):
    mlp = MLP(library())
    circuits_ids = np.unique(owner)
    rng = np.random.default_rng(seed)
    val_ids = set(rng.choice(circuits_ids, size=max(1, len(circuits_ids) // 10), replace=False).tolist())
    val = np.array([o in val_ids for o in owner])
    w, train_loss = mlp.train(X[~val], y[~val], epochs=epochs, batch=batch, lr=lr, seed=seed)
    pred = mlp.predict(w, X[val])
    val_rmse = float(np.sqrt(np.mean((pred - y[val]) ** 2)))
    base_rmse = float(np.sqrt(np.mean((y[val] - y[~val].mean()) ** 2)))
    return w, {"train_mse": float(train_loss), "val_rmse_log": val_rmse, "val_rmse_log_constant_model": base_rmse,
               "rows_train": int((~val).sum()), "rows_val": int(val.sum())}


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--circuits", type=int, default=2000)
    ap.add_argument("--min-qubits", type=int, default=24)
    ap.add_argument("--max-qubits", type=int, default=30)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--data", default=None, help="also save the recorded passes (npz)")
    args = ap.parse_args()
    X, y, owner = generate(args.circuits, args.min_qubits, args.max_qubits, args.seed)
    log(f"{len(y)} recorded passes from {args.circuits} circuits")
    if args.data:
        np.savez_compressed(args.data, X=X, y=y, owner=owner)
    w, stats = train(X, y, owner, args.seed, args.epochs, args.batch, args.lr)
    stats.update(circuits=args.circuits, qubits=[args.min_qubits, args.max_qubits], epochs=args.epochs,
                 batch=args.batch, lr=args.lr, seed=args.seed, features=PASS_FEATURES)
    np.savez(args.out, weights=w, stats=json.dumps(stats))
    log(f"model written to {args.out}: {json.dumps(stats)}")
    print(json.dumps(stats))


if __name__ == "__main__":
    main()
