"""Seeded workload generators: the circuits the simulator is tested and measured on.

Each generator returns a list of commands for `projectq_amd.commands.execute`.  Qubit ids are
0..n-1, allocated in order (what `MainEngine.allocate_qureg(n)` produces).  Angles go through
ProjectQ's canonicalisation (`projectq_amd.gates`), so the matrices equal the ones a ProjectQ
program would send.
"""

import math

import numpy as np

from . import gates as G
from .commands import gate


def allocate(n):
    return [("allocate", q) for q in range(n)]


def measure_all(n):
    return [("measure", q) for q in range(n)]


def random_layers(n, layers, cnots_per_layer, seed, measure=True):
    """Random circuit: per layer H on every qubit, then per qubit one of Rx/Ry/Rz (uniform) with
    angle U[0, 2pi), then `cnots_per_layer` CNOTs on uniformly random distinct (control,
    target) pairs.  Ends with Measure on every qubit when `measure`."""
    rng = np.random.default_rng(seed)
    cmds = allocate(n)
    names = ("Rx", "Ry", "Rz")
    for _ in range(layers):
        for q in range(n):
            cmds.append(gate(G.H, [q]))
        kinds = rng.integers(0, 3, size=n)
        angles = rng.uniform(0.0, 2.0 * math.pi, size=n)
        for q in range(n):
            cmds.append(gate(G.ROTATIONS[names[kinds[q]]](angles[q]), [q]))
        for _ in range(cnots_per_layer):
            c, t = rng.choice(n, size=2, replace=False)
            cmds.append(gate(G.X, [int(t)], [int(c)]))
    if measure:
        cmds += measure_all(n)
    return cmds


def random_layers_program(n, layers, cnots_per_layer, seed, measure=True):
    """`random_layers` as a program of `projectq_amd.compiler.run_program` (same seed, same
    gates): what a ProjectQ user writes, before the default engine list optimises it."""
    rng = np.random.default_rng(seed)
    ops = [["alloc", n]]
    names = ("Rx", "Ry", "Rz")
    for _ in range(layers):
        ops += [["gate", "H", [], [[q]], [], None] for q in range(n)]
        kinds = rng.integers(0, 3, size=n)
        angles = rng.uniform(0.0, 2.0 * math.pi, size=n)
        ops += [["gate", names[kinds[q]], [float(angles[q])], [[q]], [], None] for q in range(n)]
        for _ in range(cnots_per_layer):
            c, t = rng.choice(n, size=2, replace=False)
            ops.append(["gate", "X", [], [[int(t)]], [int(c)], None])
    if measure:
        ops += [["measure", q] for q in range(n)]
    return ops


def qft(qubits):
    """QFT as H and controlled-R in the order of projectq/setups/decompositions/
    qft2crandhadamard.py:35-39."""
    cmds = []
    nq = len(qubits)
    for i in range(nq):
        cmds.append(gate(G.H, [qubits[-1 - i]]))
        for j in range(nq - 1 - i):
            cmds.append(gate(G.R(math.pi / (1 << (1 + j))), [qubits[-1 - i]], [qubits[-1 - (j + i + 1)]]))
    return cmds


def qft_rotations_adder(n, layers, crz_per_layer, seed, n_probabilities=8, prob_qubits=8, measure=True):
    """QFT on all qubits, then `layers` layers of a random Rx/Ry/Rz per qubit plus
    `crz_per_layer` controlled-Rz gates with 1-3 random distinct controls; then the in-place
    modular adder b <- (a + b) mod 2^(n/2) on the two halves (AddQuantum through emulate_math).
    Returns (commands, probability queries) where each query is (bits, qubit ids)."""
    rng = np.random.default_rng(seed)
    cmds = allocate(n) + qft(list(range(n)))
    names = ("Rx", "Ry", "Rz")
    for _ in range(layers):
        kinds = rng.integers(0, 3, size=n)
        angles = rng.uniform(0.0, 2.0 * math.pi, size=n)
        for q in range(n):
            cmds.append(gate(G.ROTATIONS[names[kinds[q]]](angles[q]), [q]))
        for _ in range(crz_per_layer):
            k = int(rng.integers(1, 4))
            picks = rng.choice(n, size=k + 1, replace=False)
            cmds.append(gate(G.Rz(rng.uniform(0.0, 2.0 * math.pi)), [int(picks[0])], [int(q) for q in picks[1:]]))
    half = n // 2
    a, b = list(range(half)), list(range(half, 2 * half))
    cmds.append(("math", G.add_quantum(half), [a, b], []))
    queries = []
    for _ in range(n_probabilities):
        ids = [int(q) for q in rng.choice(n, size=prob_qubits, replace=False)]
        bits = [bool(v) for v in rng.integers(0, 2, size=prob_qubits)]
        queries.append((bits, ids))
    if measure:
        cmds += measure_all(n)
    return cmds, queries


def hardware_efficient_ansatz(n, layers, seed):
    """Variational ansatz: `layers` x (Ry, Rz on every qubit with angles U[0, 2pi), then a linear
    CNOT chain q0->q1->...->q_{n-1})."""
    rng = np.random.default_rng(seed)
    cmds = allocate(n)
    for _ in range(layers):
        th = rng.uniform(0.0, 2.0 * math.pi, size=(n, 2))
        for q in range(n):
            cmds.append(gate(G.Ry(th[q, 0]), [q]))
            cmds.append(gate(G.Rz(th[q, 1]), [q]))
        for q in range(n - 1):
            cmds.append(gate(G.X, [q + 1], [q]))
    return cmds


def random_pauli_operator(n, n_terms, seed):
    """QubitOperator-style terms: each qubit I/X/Y/Z uniform, coefficients N(0, 1).  Returned in
    the backend's format [(list[(local index, op)], coeff)]."""
    rng = np.random.default_rng(seed)
    terms = []
    for _ in range(n_terms):
        ops = rng.integers(0, 4, size=n)
        term = [(q, "IXYZ"[o]) for q, o in enumerate(ops) if o != 0]
        terms.append((term, float(rng.normal())))
    return terms


def mixed_circuit(n, depth, seed, measure=False):
    """Depth-`depth` circuit of mixed 1/2/3-qubit gates: each layer applies, on a random
    permutation of the qubits, random 1-qubit rotations, 2-qubit gates (CNOT, CZ-like
    controlled-R, Rzz, Swap) and 3-qubit gates (Toffoli, controlled-Swap)."""
    rng = np.random.default_rng(seed)
    cmds = allocate(n)
    for _ in range(depth):
        perm = [int(q) for q in rng.permutation(n)]
        i = 0
        while i < n:
            width = int(rng.choice([1, 2, 3], p=[0.5, 0.35, 0.15]))
            width = min(width, n - i)
            qs = perm[i : i + width]
            i += width
            if width == 1:
                name = ("Rx", "Ry", "Rz")[int(rng.integers(0, 3))]
                cmds.append(gate(G.ROTATIONS[name](rng.uniform(0, 2 * math.pi)), qs))
            elif width == 2:
                k = int(rng.integers(0, 4))
                if k == 0:
                    cmds.append(gate(G.X, [qs[1]], [qs[0]]))
                elif k == 1:
                    cmds.append(gate(G.R(rng.uniform(0, 2 * math.pi)), [qs[1]], [qs[0]]))
                elif k == 2:
                    cmds.append(gate(G.Rzz(rng.uniform(0, 2 * math.pi)), qs))
                else:
                    cmds.append(gate(G.Swap, qs))
            else:
                if rng.integers(0, 2) == 0:
                    cmds.append(gate(G.X, [qs[2]], qs[:2]))
                else:
                    cmds.append(gate(G.Swap, qs[1:], [qs[0]]))
    if measure:
        cmds += measure_all(n)
    return cmds
