"""Record the backend-call sequences ProjectQ's default compiler emits (run in the build container only).

Each program below is written in a small JSON op language, run through the reference's
`MainEngine` with its default engine list (projectq/setups/default.py:31-34: TagRemover,
LocalOptimizer(10), AutoReplacer, TagRemover, LocalOptimizer(10)) in front of its `Simulator`
backend, whose `_simulator` seat holds a call recorder (the swap the reference's own test fixture
makes, projectq/backends/_sim/_simulator_test.py:80-93).  `FALLBACK_TO_PYSIM` is cleared so the
backend takes the path it takes with the compiled simulator (native AddConstant / ModN calls,
projectq/backends/_sim/_simulator.py:374-395).

The fixture holds the programs and the recorded calls; nothing in it needs the reference to read.

    python tests/golden/make_engine_traces.py      # rewrites tests/golden/engine_traces.json

Op language (qubit ids are the ids the engine hands out: 0, 1, 2, ... in allocation order):
    ["alloc", k]                                     allocate_qureg(k)
    ["gate", name, params, regs, ctrls, state]       with Control(eng, ctrls, state): gate | regs
    ["measure", q] / ["dealloc", q] / ["flush"]
    ["compute"] / ["compute_end"] / ["uncompute"]    Compute / end of its with-block / Uncompute
    ["dagger"] / ["dagger_end"]                      with Dagger(eng)
    ["control", ctrls, state] / ["control_end"]      with Control(eng, ctrls, state) around ops
"""

import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from trace_recorder import Recorder  # noqa: E402

REFERENCE = "/root/reference"

FIXED = ["H", "X", "Y", "Z", "S", "Sdag", "T", "Tdag", "SqrtX", "SqrtXdag"]
ROT1 = ["Rx", "Ry", "Rz", "Ph", "R"]
ROT2 = ["Rxx", "Ryy", "Rzz"]
PI = math.pi


def g(name, regs, params=(), ctrls=(), state=None):
    return ["gate", name, list(params), [list(r) for r in regs], list(ctrls), state]


# --- hand-written programs ------------------------------------------------------------------
def prog_merge_cancel():
    ops = [["alloc", 4]]
    ops += [g("Rx", [[0]], [0.3]), g("Rx", [[0]], [0.4]), g("Rx", [[0]], [-0.7])]  # merge to ~0
    ops += [g("H", [[1]]), g("H", [[1]])]  # cancel
    ops += [g("S", [[2]]), g("Sdag", [[2]]), g("T", [[2]]), g("T", [[2]]), g("Tdag", [[2]])]
    ops += [g("X", [[3]], ctrls=[0]), g("X", [[3]], ctrls=[0])]  # CNOT pair cancels
    ops += [g("X", [[3]], ctrls=[1]), g("Z", [[1]]), g("X", [[3]], ctrls=[1])]  # blocked on qubit 1
    ops += [g("Rz", [[2]], [2 * PI]), g("Rz", [[2]], [2 * PI])]  # 4 pi -> 0 -> identity
    ops += [g("Rz", [[0]], [1.0], ctrls=[2]), g("Rz", [[0]], [2.0], ctrls=[2])]  # controlled merge
    ops += [g("R", [[1]], [PI / 8]), g("R", [[1]], [PI / 8]), g("Ph", [[1]], [0.5]), g("Ph", [[1]], [-0.5])]
    ops += [g("Rzz", [[0], [1]], [0.25]), g("Rzz", [[0], [1]], [0.5])]
    ops += [g("Rzz", [[1], [0]], [0.25])]  # same class, reversed registers: no merge
    ops += [g("Swap", [[2], [3]]), g("Swap", [[3], [2]])]  # interchangeable: cancels
    ops += [g("SqrtSwap", [[0], [2]]), g("SqrtSwap", [[2], [0]])]
    ops += [g("Ry", [[3]], [4 * PI - 1e-13]), g("Ry", [[3]], [PI]), g("Ry", [[3]], [3 * PI])]
    ops += [["measure", 0], g("H", [[0]]), ["flush"]]
    ops += [["measure", q] for q in range(4)]
    ops += [["dealloc", q] for q in range(4)]
    return ops


def prog_cache_overflow():
    ops = [["alloc", 3]]
    for k in range(25):
        ops.append(g(ROT1[k % 3], [[0]], [0.1 * (k + 1)]))
        if k % 7 == 3:
            ops.append(g("X", [[1]], ctrls=[0]))
        if k % 5 == 1:
            ops.append(g("H", [[2]]))
    for k in range(12):
        ops.append(g("Rz", [[1]], [0.01 * k]))
    ops += [g("X", [[2]], ctrls=[0, 1])]
    ops += [["measure", q] for q in range(3)] + [["dealloc", q] for q in range(3)]
    return ops


def prog_qft():
    n = 5
    ops = [["alloc", n]]
    ops += [g("H", [[q]]) for q in range(n)]
    ops += [g("QFT", [list(range(n))])]
    ops += [g("Rz", [[2]], [0.7])]
    ops += [g("QFTdag", [list(range(n))])]
    ops += [g("QFT", [[1, 3, 4]]), g("QFTdag", [[1, 3, 4]])]  # adjacent: cancels after decomposition
    ops += [g("QFT", [[0, 1, 2]], ctrls=[4])]
    ops += [g("QFTdag", [[2, 0]], ctrls=[3], state="0")]
    ops += [["measure", q] for q in range(n)] + [["dealloc", q] for q in range(n)]
    return ops


def prog_compute_uncompute():
    ops = [["alloc", 4]]
    ops += [["compute"], g("H", [[0]]), g("X", [[1]], ctrls=[0]), g("Rz", [[2]], [0.3]), ["compute_end"]]
    ops += [g("Z", [[3]], ctrls=[1, 2])]
    ops += [["uncompute"]]
    ops += [["compute"], g("S", [[0]]), g("T", [[0]]), ["compute_end"], ["uncompute"]]  # cancels entirely
    ops += [["control", [3], None], ["compute"], g("X", [[0]]), ["compute_end"]]
    ops += [g("Ry", [[1]], [1.1]), ["uncompute"], ["control_end"]]  # compute part stays uncontrolled
    ops += [["compute"], g("X", [[2]]), ["compute_end"], ["control", [2], None], g("H", [[1]])]
    ops += [["control_end"], ["uncompute"]]
    ops += [["measure", q] for q in range(4)] + [["dealloc", q] for q in range(4)]
    return ops


def prog_dagger_and_controls():
    ops = [["alloc", 4]]
    ops += [["dagger"], g("S", [[0]]), g("Rx", [[1]], [0.2]), g("X", [[2]], ctrls=[0]), g("T", [[3]]), ["dagger_end"]]
    ops += [g("X", [[1]], ctrls=[0], state="0"), g("Rz", [[2]], [0.9], ctrls=[0, 3], state="01")]
    ops += [g("H", [[3]], ctrls=[0, 1, 2], state="101")]
    ops += [["control", [0, 1], "10"], g("Ry", [[2]], [0.4]), g("Swap", [[2], [3]]), ["control_end"]]
    ops += [["dagger"], g("Rz", [[1]], [0.5], ctrls=[2], state="0"), ["dagger_end"]]
    ops += [g("Barrier", [[0, 1, 2, 3]])]
    ops += [g("Matrix", [[0]], [[[0.0, 0.0], [1.0, 0.0]], [[1.0, 0.0], [0.0, 0.0]]])]
    ops += [g("Matrix", [[0]], [[[0.0, 0.0], [1.0, 0.0]], [[1.0, 0.0], [0.0, 0.0]]])]
    ops += [["measure", q] for q in range(4)] + [["dealloc", q] for q in range(4)]
    return ops


def prog_math():
    ops = [["alloc", 9]]
    a, b = [0, 1, 2], [3, 4, 5]
    ops += [g("X", [[0]]), g("X", [[4]])]
    ops += [g("AddConstant", [a], [3]), g("SubConstant", [a], [3]), g("AddConstant", [b], [5])]
    ops += [g("AddConstantModN", [a], [2, 7]), g("SubConstantModN", [b], [1, 5])]
    ops += [g("MultiplyByConstantModN", [a], [3, 7], ctrls=[8])]
    ops += [g("AddQuantum", [a, b]), g("AddQuantum", [a, b, [6]])]
    ops += [g("AddQuantumdag", [a, b]), g("SubtractQuantum", [a, b]), g("ComparatorQuantum", [a, b, [7]])]
    ops += [g("AddConstant", [b], [1], ctrls=[6, 7])]
    ops += [["measure", q] for q in range(9)] + [["dealloc", q] for q in range(9)]
    return ops


def prog_realloc():
    ops = [["alloc", 3], g("H", [[0]]), g("X", [[1]], ctrls=[0]), ["measure", 1], ["dealloc", 1]]
    ops += [["alloc", 2], g("Rx", [[3]], [0.2]), g("Rx", [[3]], [0.3]), g("X", [[4]], ctrls=[3])]
    ops += [["measure", q] for q in (0, 2, 3, 4)] + [["dealloc", q] for q in (4, 0, 3, 2)]
    return ops


def prog_layers(n=6, layers=4, seed=11):
    """The headline random-circuit shape (H layer, random Rx/Ry/Rz, random CNOTs) on a few qubits."""
    rng = np.random.default_rng(seed)
    ops = [["alloc", n]]
    for _ in range(layers):
        ops += [g("H", [[q]]) for q in range(n)]
        kinds = rng.integers(0, 3, size=n)
        angles = rng.uniform(0.0, 2 * PI, size=n)
        ops += [g(("Rx", "Ry", "Rz")[int(kinds[q])], [[q]], [float(angles[q])]) for q in range(n)]
        for _ in range(n // 2):
            c, t = rng.choice(n, size=2, replace=False)
            ops.append(g("X", [[int(t)]], ctrls=[int(c)]))
    ops += [["measure", q] for q in range(n)] + [["dealloc", q] for q in range(n)]
    return ops


def prog_identity_multiqubit():
    """Identity commands on more than one qubit (a controlled zero rotation, Rzz(4 pi))."""
    ops = [["alloc", 3], g("H", [[0]]), g("Rz", [[1]], [0.0], ctrls=[0]), g("Rzz", [[1], [2]], [4 * PI])]
    ops += [g("H", [[2]]), g("Rx", [[2]], [0.0]), g("Y", [[2]])]
    ops += [["measure", q] for q in range(3)] + [["dealloc", q] for q in range(3)]
    return ops


def te(time, terms):
    """TimeEvolution params: terms as [[[index, "X"], ...], coefficient] pairs."""
    return [time, [[[list(lp) for lp in term], c] for term, c in terms]]


def prog_time_evolution():
    h1 = [(((0, "X"), (2, "Z")), 0.5), (((1, "Y"),), -0.25)]
    ops = [["alloc", 5]] + [g("H", [[q]]) for q in range(5)]
    ops += [g("TimeEvolution", [[0, 1, 2, 3]], te(0.3, h1))]
    ops += [g("TimeEvolution", [[0, 1, 2, 3]], te(0.3, h1))]  # equal only by identity: no merge
    ops += [g("TimeEvolution", [[1, 2]], te(0.7, [((), 1.5)]))]  # identity only: a phase gate
    ops += [g("TimeEvolution", [[2, 3]], te(0.2, [(((0, "Z"), (1, "Z")), 1.0), ((), 0.5)]), ctrls=[0])]
    ops += [g("TimeEvolution", [[4, 0, 3]], te(-0.6, [(((2, "X"), (0, "Y")), 0.75), (((1, "Z"),), 2.0)]))]
    ops += [["dagger"], g("TimeEvolution", [[3]], te(0.4, [(((0, "X"),), 1.0)])), g("H", [[3]]), ["dagger_end"]]
    ops += [["compute"], g("TimeEvolution", [[1, 4]], te(0.1, [(((0, "Y"), (1, "Y")), 0.3)])), ["compute_end"]]
    ops += [g("Z", [[2]], ctrls=[1]), ["uncompute"]]
    ops += [["measure", q] for q in range(5)] + [["dealloc", q] for q in range(5)]
    return ops


def prog_more_gates():
    """Multiply/Divide quantum and their inverses, Entangle, QubitOperator gates (identity,
    single-qubit and multi-qubit terms, a control, and two different adjacent Pauli strings on the
    same qubits, which the reference's optimizer cancels because QubitOperator gates compare equal)."""
    ops = [["alloc", 10], g("X", [[0]]), g("X", [[3]]), g("H", [[1]])]
    ops += [g("MultiplyQuantum", [[0, 1], [2, 3], [4, 5, 6, 7, 8]])]
    ops += [g("MultiplyQuantumdag", [[0, 1], [2, 3], [4, 5, 6, 7, 8]], ctrls=[9])]
    ops += [g("DivideQuantum", [[0, 1, 2], [3, 4, 5], [6, 7, 8]])]
    ops += [g("DivideQuantumdag", [[0, 1, 2], [3, 4, 5], [6, 7, 8]])]
    ops += [g("Entangle", [[0, 1, 2, 3, 4]]), g("Entangle", [[5, 6, 7]], ctrls=[9])]
    ops += [g("QubitOperator", [list(range(6))], [[[0, "X"], [2, "Z"]], [0.0, 1.0]])]
    ops += [g("QubitOperator", [list(range(6))], [[[2, "Y"]], [math.cos(0.3), math.sin(0.3)]])]
    ops += [g("QubitOperator", [[3, 4]], [[], [-1.0, 0.0]])]
    ops += [g("QubitOperator", [[4, 5, 6]], [[[0, "X"], [2, "Y"]], [0.0, -1.0]], ctrls=[9])]
    ops += [g("QubitOperator", [[7, 8]], [[[0, "X"], [1, "X"]], [1.0, 0.0]])]
    ops += [g("QubitOperator", [[7, 8]], [[[0, "Z"], [1, "Z"]], [1.0, 0.0]])]
    ops += [["dagger"], g("QubitOperator", [[1, 3, 5]], [[[0, "Y"], [1, "Z"], [2, "X"]], [0.0, 1.0]]), ["dagger_end"]]
    ops += [["measure", q] for q in range(10)] + [["dealloc", q] for q in range(10)]
    return ops


# --- seeded random programs -----------------------------------------------------------------
ANGLE_POOL = [0.0, PI / 4, PI / 2, PI, 2 * PI, 3 * PI, 4 * PI - 1e-13, 4 * PI, -PI / 2, 1e-13]


def _angle(rng, last):
    r = rng.random()
    if r < 0.35:
        return float(ANGLE_POOL[rng.integers(len(ANGLE_POOL))])
    if r < 0.5 and last is not None:
        return -last
    return float(rng.uniform(-2 * PI, 6 * PI))


def _random_gate(rng, n, last_angle, allow_neg=True):
    qs = [int(q) for q in rng.permutation(n)]
    r = rng.random()
    if r < 0.35:
        name, regs, params = FIXED[rng.integers(len(FIXED))], [[qs[0]]], []
    elif r < 0.7:
        name = ROT1[rng.integers(len(ROT1))]
        regs, params = [[qs[0]]], [_angle(rng, last_angle.get(qs[0]))]
        last_angle[qs[0]] = params[0]
    elif r < 0.82:
        name = ROT2[rng.integers(len(ROT2))]
        regs, params = [[qs[0]], [qs[1]]], [float(rng.uniform(-2 * PI, 6 * PI))]
    elif r < 0.9:
        name, regs, params = ("Swap", "SqrtSwap")[rng.integers(2)], [[qs[0]], [qs[1]]], []
    else:
        k = int(rng.integers(2, min(4, n) + 1))
        name, regs, params = ("QFT", "QFTdag")[rng.integers(2)], [qs[:k]], []
    used = {q for reg in regs for q in reg}
    free = [q for q in qs if q not in used]
    ctrls, state = [], None
    if free and rng.random() < 0.35:
        nc = int(rng.integers(1, min(2, len(free)) + 1))
        ctrls = free[:nc]
        if params and name in ROT1:
            # a controlled rotation must not reduce to the identity: the reference's optimizer
            # mishandles identity commands on several qubits (prog_identity_multiqubit pins that)
            params = [float(rng.uniform(0.1, 4 * PI - 0.1))]
        if allow_neg and rng.random() < 0.25:
            state = "".join("01"[int(b)] for b in rng.integers(0, 2, size=nc))
    return g(name, regs, params, ctrls, state)


def prog_random(seed, n=5, length=160):
    rng = np.random.default_rng(seed)
    ops = [["alloc", n]]
    last = {}
    while len(ops) < length:
        r = rng.random()
        if r < 0.04:
            ops.append(["measure", int(rng.integers(n))])
        elif r < 0.06:
            ops.append(["flush"])
        elif r < 0.1:
            body = [_random_gate(rng, n, last, allow_neg=False) for _ in range(int(rng.integers(1, 4)))]
            action = [_random_gate(rng, n, last) for _ in range(int(rng.integers(1, 3)))]
            ops += [["compute"], *body, ["compute_end"], *action, ["uncompute"]]
        elif r < 0.13:
            body = [_random_gate(rng, n, last) for _ in range(int(rng.integers(1, 5)))]
            ops += [["dagger"], *body, ["dagger_end"]]
        elif r < 0.15:
            ops.append(g("Barrier", [[int(q) for q in rng.permutation(n)[:3]]]))
        else:
            ops.append(_random_gate(rng, n, last))
            # repeat or invert the last gate now and then to provoke merges and cancellations
            if rng.random() < 0.25:
                ops.append(ops[-1])
    ops += [["measure", q] for q in range(n)] + [["dealloc", q] for q in range(n)]
    return ops


PROGRAMS = {
    "merge_cancel": prog_merge_cancel(),
    "cache_overflow": prog_cache_overflow(),
    "qft": prog_qft(),
    "compute_uncompute": prog_compute_uncompute(),
    "dagger_and_controls": prog_dagger_and_controls(),
    "math": prog_math(),
    "realloc": prog_realloc(),
    "layers6": prog_layers(),
    "identity_multiqubit": prog_identity_multiqubit(),
    "time_evolution": prog_time_evolution(),
    "more_gates": prog_more_gates(),
    "negative_control_merge": [["alloc", 3], g("H", [[0]]), g("H", [[1]]), g("Rx", [[1]], [0.3], [0], "0"),
                               g("Rx", [[1]], [0.4], [0], "0"), g("X", [[2]], [], [0], "0"), g("X", [[2]], [], [0], "1"),
                               ["measure", 0], ["measure", 1], ["measure", 2], ["dealloc", 0], ["dealloc", 1],
                               ["dealloc", 2]],
}
for _s in range(12):
    PROGRAMS[f"random{_s}"] = prog_random(100 + _s, n=4 + _s % 3)


# --- the reference driver -------------------------------------------------------------------
def run_reference(ops):
    sys.path.insert(0, REFERENCE)
    import projectq.backends._sim._simulator as simmod
    from projectq import MainEngine
    from projectq.backends import Simulator
    from projectq.libs import math as qmath
    from projectq.meta import Compute, Control, Dagger, Uncompute
    from projectq.ops import (
        QFT, Barrier, Entangle, H, MatrixGate, Measure, Ph, QubitOperator, R, Rx, Rxx, Ry, Ryy, Rz, Rzz, S, Sdag, SqrtSwap,
        SqrtX, Swap, T, Tdag, TimeEvolution, X, Y, Z, get_inverse,
    )
    from projectq.types import Qureg

    simmod.FALLBACK_TO_PYSIM = False
    backend = Simulator(rnd_seed=1)
    rec = Recorder()
    backend._simulator = rec
    eng = MainEngine(backend=backend)
    fixed = {"H": H, "X": X, "Y": Y, "Z": Z, "S": S, "Sdag": Sdag, "T": T, "Tdag": Tdag, "SqrtX": SqrtX,
             "SqrtXdag": get_inverse(SqrtX), "Swap": Swap, "SqrtSwap": SqrtSwap, "QFT": QFT,
             "QFTdag": get_inverse(QFT), "Barrier": Barrier, "AddQuantum": qmath.AddQuantum,
             "AddQuantumdag": get_inverse(qmath.AddQuantum), "SubtractQuantum": qmath.SubtractQuantum,
             "ComparatorQuantum": qmath.ComparatorQuantum, "MultiplyQuantum": qmath.MultiplyQuantum,
             "MultiplyQuantumdag": get_inverse(qmath.MultiplyQuantum), "DivideQuantum": qmath.DivideQuantum,
             "DivideQuantumdag": get_inverse(qmath.DivideQuantum), "Entangle": Entangle}
    param = {"Rx": Rx, "Ry": Ry, "Rz": Rz, "Rxx": Rxx, "Ryy": Ryy, "Rzz": Rzz, "Ph": Ph, "R": R,
             "AddConstant": qmath.AddConstant, "SubConstant": qmath.SubConstant,
             "AddConstantModN": qmath.AddConstantModN, "SubConstantModN": qmath.SubConstantModN,
             "MultiplyByConstantModN": qmath.MultiplyByConstantModN}
    qubits = {}
    stack = []

    def make_gate(name, params):
        if name in fixed:
            return fixed[name]
        if name == "Matrix":
            return MatrixGate([[complex(re, im) for re, im in row] for row in params])
        if name == "QubitOperator":
            return QubitOperator(tuple(tuple(lp) for lp in params[0]), complex(*params[1]))
        if name == "TimeEvolution":
            ham = QubitOperator()
            for term, c in params[1]:
                ham += QubitOperator(tuple(tuple(lp) for lp in term), c)
            return TimeEvolution(params[0], ham)
        return param[name](*params)

    def ctl(ctrls, state):
        return Control(eng, [qubits[c] for c in ctrls], ctrl_state=state if state is not None else "1" * len(ctrls))

    error = None
    try:
        for op in ops:
            kind = op[0]
            if kind == "alloc":
                for qb in eng.allocate_qureg(op[1]):
                    qubits[qb.id] = qb
            elif kind == "gate":
                _, name, params, regs, ctrls, state = op
                gate = make_gate(name, params)
                target = tuple(Qureg([qubits[q] for q in reg]) for reg in regs)
                if ctrls:
                    with ctl(ctrls, state):
                        gate | target
                else:
                    gate | target
            elif kind == "measure":
                Measure | qubits[op[1]]
            elif kind == "dealloc":
                eng.deallocate_qubit(qubits.pop(op[1]))
            elif kind == "flush":
                eng.flush()
            elif kind in ("compute", "dagger", "control"):
                cm = {"compute": lambda: Compute(eng), "dagger": lambda: Dagger(eng),
                      "control": lambda: ctl(op[1], op[2])}[kind]()
                cm.__enter__()
                stack.append(cm)
            elif kind in ("compute_end", "dagger_end", "control_end"):
                stack.pop().__exit__(None, None, None)
            elif kind == "uncompute":
                Uncompute(eng)
            else:
                raise ValueError(kind)
        eng.flush()
    except Exception as err:  # noqa: BLE001 -- the fixture records the reference's error
        error = type(err).__name__
    calls = list(rec.calls)
    backend._simulator = Recorder()  # absorb the engine's exit-time flush
    return calls, error


def main():
    out = {"engine_list": "TagRemover, LocalOptimizer(10), AutoReplacer(default rules), TagRemover, LocalOptimizer(10)",
           "programs": {}}
    for name, ops in PROGRAMS.items():
        calls, error = run_reference(ops)
        out["programs"][name] = {"ops": ops, "calls": calls, "error": error}
        n_gates = sum(1 for c in calls if c[0] == "apply_controlled_gate")
        print(f"{name}: {len(ops)} ops -> {len(calls)} calls ({n_gates} gates) error={error}")
    with open(os.path.join(HERE, "engine_traces.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
