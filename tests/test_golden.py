"""Golden vectors written by ProjectQ's own C++ simulator (tests/golden/make_golden.py, compiled
from the reference sources): every backend must reproduce them.  The CPU oracle is checked
here on every run; the HIP library (complex128, and the tile planner) under the `gpu` marker."""

import glob
import os

import numpy as np
import pytest

from projectq_amd import commands
from tests.backends import make

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

BACKENDS = [
    pytest.param("oracle", {}, id="oracle"),
    pytest.param("hip_c128", {}, id="hip_c128", marks=pytest.mark.gpu),
    pytest.param("hip_c128", {"fusion": "window", "planner": "tile"}, id="hip_c128_tile", marks=pytest.mark.gpu),
    pytest.param("hip_c128", {"fusion": "window", "planner": "greedy"}, id="hip_c128_greedy", marks=pytest.mark.gpu),
]


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name)))


def unpack_gates(rec, prefix=""):
    k, tg, nc, ct, mats = (rec[prefix + x] for x in ("k", "targets", "n_ctrl", "controls", "matrices"))
    cmds, ti, ci, mi = [], 0, 0, 0
    for i in range(len(k)):
        kk = int(k[i])
        d = 1 << kk
        m = mats[mi : mi + d * d].reshape(d, d)
        mi += d * d
        targets = [int(x) for x in tg[ti : ti + kk]]
        ti += kk
        controls = [int(x) for x in ct[ci : ci + int(nc[i])]]
        ci += int(nc[i])
        cmds.append(("gate", m, tuple(targets), tuple(sorted(controls))))
    return cmds


def unpack_terms(rec, prefix, real):
    idx, ops, off, co = (rec[f"{prefix}_{x}"] for x in ("idx", "ops", "off", "coef"))
    terms = []
    for t in range(len(co)):
        term = [(int(idx[j]), chr(int(ops[j]))) for j in range(off[t], off[t + 1])]
        terms.append((term, float(co[t].real) if real else complex(co[t])))
    return terms


def fresh(kind, opts, n, seed=1):
    sim = make(kind, seed, **opts)
    for q in range(n):
        sim.allocate_qubit(q)
    return sim


STATE_FILES = sorted(os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN, "state_*.npz")))


@pytest.mark.parametrize("kind,opts", BACKENDS)
@pytest.mark.parametrize("fname", STATE_FILES)
def test_state_vectors(kind, opts, fname):
    rec = load(fname)
    n = int(rec["n"])
    sim = fresh(kind, opts, n, 7)
    commands.execute(sim, unpack_gates(rec), gate_fusion=fname.endswith("fused.npz"))
    got = sim.cheat_numpy()[1]
    assert np.max(np.abs(got - rec["state"])) <= 1e-12


@pytest.mark.parametrize("kind,opts", BACKENDS)
def test_random20_amplitudes_probabilities_measurements(kind, opts):
    rec = load("random20_measure.npz")
    n = int(rec["n"])
    sim = fresh(kind, opts, n, int(rec["seed"]))
    commands.execute(sim, unpack_gates(rec), gate_fusion=bool(opts))
    state = sim.cheat_numpy()[1]
    assert np.max(np.abs(state[rec["sample_idx"]] - rec["sample_amp"])) <= 1e-12
    assert abs(np.vdot(state, state).real - rec["norm"]) <= 1e-12
    for bits, ids, p in zip(rec["prob_bits"], rec["prob_ids"], rec["probs"]):
        assert abs(sim.get_probability([bool(b) for b in bits], [int(q) for q in ids]) - p) <= 1e-12
    outcomes = [int(sim.measure_qubits([q])[0]) for q in range(n)]
    assert outcomes == [int(x) for x in rec["outcomes"]]


@pytest.mark.parametrize("kind,opts", BACKENDS)
def test_expectation_values(kind, opts):
    rec = load("expectation_ansatz.npz")
    for i in range(2):
        n = int(rec[f"a{i}_n"])
        sim = fresh(kind, opts, n)
        commands.execute(sim, unpack_gates(rec, f"a{i}_"), gate_fusion=True)
        val = sim.get_expectation_value(unpack_terms(rec, f"a{i}_h", real=True), list(range(n)))
        assert abs(val - rec[f"a{i}_value"]) <= 1e-11


MATH = ["add_quantum", "add_quantum_carry", "subtract", "comparator", "multiply", "divide"]


@pytest.mark.parametrize("kind,opts", BACKENDS[:2])
def test_emulate_math_bit_exact(kind, opts):
    from projectq_amd import gates as G

    fns = {
        "add_quantum": (G.add_quantum(4), [[0, 1, 2, 3], [4, 5, 6, 7]], []),
        "add_quantum_carry": (G.add_quantum(3), [[0, 1, 2], [3, 4, 5], [6]], [7]),
        "subtract": (G.subtract_quantum(), [[0, 1, 2], [3, 4, 5]], [6]),
        "comparator": (G.comparator_quantum(), [[0, 1, 2], [3, 4, 5], [6]], []),
        "multiply": (G.multiply_quantum(), [[0, 1], [2, 3], [4, 5, 6, 7, 8]], []),
        "divide": (G.divide_quantum(), [[0, 1, 2], [3, 4, 5], [6, 7, 8]], []),
    }
    rec = load("emulate_math.npz")
    n = int(rec["n"])
    for name in MATH:
        fn, regs, ctrl = fns[name]
        sim = fresh(kind, opts, n)
        sim.set_wavefunction(rec["wf"], list(range(n)))
        sim.emulate_math(fn, regs, ctrl)
        assert np.array_equal(sim.cheat_numpy()[1], rec[f"{name}_out"]), name
    for which, a, mod in (("add", 5, 0), ("add_mod", 4, 13), ("mul_mod", 7, 15)):
        sim = fresh(kind, opts, n)
        sim.set_wavefunction(rec["wf"], list(range(n)))
        commands.execute(sim, [("math_const", which, a, mod, [[0, 1, 2, 3]], [7])])
        assert np.array_equal(sim.cheat_numpy()[1], rec[f"const_{which}_out"]), which


@pytest.mark.parametrize("kind,opts", BACKENDS[:2])
def test_time_evolution_and_qubit_operator(kind, opts):
    rec = load("pauli_ops.npz")
    n = int(rec["n"])
    sim = fresh(kind, opts, n)
    sim.set_wavefunction(rec["wf"], list(range(n)))
    sim.emulate_time_evolution(unpack_terms(rec, "h", real=True), float(rec["time"]), list(range(6)), [7])
    assert np.max(np.abs(sim.cheat_numpy()[1] - rec["te_out"])) <= 1e-11
    sim = fresh(kind, opts, n)
    sim.set_wavefunction(rec["wf"], list(range(n)))
    sim.apply_qubit_operator(unpack_terms(rec, "op", real=False), list(range(n)))
    assert np.max(np.abs(sim.cheat_numpy()[1] - rec["ao_out"])) <= 1e-12


@pytest.mark.parametrize("kind,opts", BACKENDS)
def test_qft_rotations_adder14(kind, opts):
    """The QFT + rotations + AddQuantum workload at 14 qubits (the 30-qubit one's shape): state
    after b <- (a + b) mod 2^7, probabilities of 6-qubit patterns, seeded measurement of all."""
    from projectq_amd import gates as G

    rec = load("qft_rotations_adder14.npz")
    n = int(rec["n"])
    half = n // 2
    sim = fresh(kind, opts, n, int(rec["seed"]))
    cmds = unpack_gates(rec) + [("math", G.add_quantum(half), [list(range(half)), list(range(half, n))], [])]
    commands.execute(sim, cmds, gate_fusion=bool(opts))
    state = sim.cheat_numpy()[1]
    assert np.max(np.abs(state - rec["state"])) <= 1e-12
    for bits, ids, p in zip(rec["prob_bits"], rec["prob_ids"], rec["probs"]):
        assert abs(sim.get_probability([bool(b) for b in bits], [int(q) for q in ids]) - p) <= 1e-12
    outcomes = [int(sim.measure_qubits([q])[0]) for q in range(n)]
    assert outcomes == [int(x) for x in rec["outcomes"]]
