"""Write golden vectors with ProjectQ's own C++ simulator (run in the build container only).

The reference simulator is compiled from its sources by `make -C oracle ref` into oracle/_ref/
and driven directly through its pybind11 surface, the way the reference's own test fixture
swaps `Simulator._simulator` (projectq/backends/_sim/_simulator_test.py:80-93).  Every fixture
stores its inputs (gate lists, states, operators) next to the reference's outputs, so the tests
never need the reference or the circuit generators to reproduce them.

    python tests/golden/make_golden.py        # rewrites tests/golden/*.npz
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402
from projectq_amd import circuits, commands  # noqa: E402
from projectq_amd import gates as G  # noqa: E402


def pack_gates(cmds):
    """Gate commands -> flat arrays (k, targets, n_ctrl, controls, matrices)."""
    ks, tg, nc, ct, mats = [], [], [], [], []
    for c in cmds:
        if c[0] != "gate":
            continue
        _, m, targets, controls = c
        ks.append(len(targets))
        tg.extend(targets)
        nc.append(len(controls))
        ct.extend(controls)
        mats.append(np.asarray(m, dtype=np.complex128).reshape(-1))
    return {
        "k": np.asarray(ks, np.int64),
        "targets": np.asarray(tg, np.int64),
        "n_ctrl": np.asarray(nc, np.int64),
        "controls": np.asarray(ct, np.int64),
        "matrices": np.concatenate(mats) if mats else np.zeros(0, np.complex128),
    }


def pack_terms(terms, prefix):
    idx, ops, off, co = [], [], [0], []
    for term, c in terms:
        for q, p in term:
            idx.append(q)
            ops.append(ord(p))
        off.append(len(idx))
        co.append(c)
    return {
        f"{prefix}_idx": np.asarray(idx, np.int64),
        f"{prefix}_ops": np.asarray(ops, np.int64),
        f"{prefix}_off": np.asarray(off, np.int64),
        f"{prefix}_coef": np.asarray(co, np.complex128),
    }


def main():
    cpp = ref.load("intrin")
    if cpp is None:
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` first")
    out = {}

    # 1. state vectors of seeded circuits, gate_fusion off and on
    cases = {
        "random10": circuits.random_layers(10, 10, 5, 101, measure=False),
        "random14": circuits.random_layers(14, 12, 7, 102, measure=False),
        "mixed12": circuits.mixed_circuit(12, 10, 103),
        "qft12": circuits.allocate(12)
        + [commands.gate(G.Ry(0.1 * q + 0.3), [q]) for q in range(12)]
        + circuits.qft(list(range(12))),
    }
    for name, cmds in cases.items():
        n = sum(1 for c in cmds if c[0] == "allocate")
        for fusion in (False, True):
            sim = cpp.Simulator(7)
            commands.execute(sim, cmds, gate_fusion=fusion)
            mapping, state = sim.cheat()
            rec = pack_gates(cmds)
            rec["n"] = np.int64(n)
            rec["state"] = np.asarray(state, np.complex128)
            np.savez_compressed(os.path.join(HERE, f"state_{name}_{'fused' if fusion else 'plain'}.npz"), **rec)

    # 2. a 20-qubit random circuit (the CPU-runnable workload): sampled amplitudes, seeded
    # measurement outcomes of all qubits, probabilities
    n = 20
    cmds = circuits.random_layers(n, 20, 10, 2020, measure=False)
    sim = cpp.Simulator(2020)
    commands.execute(sim, cmds, gate_fusion=False)
    _, state = sim.cheat()
    state = np.asarray(state, np.complex128)
    rng = np.random.default_rng(5)
    sample_idx = rng.choice(2**n, size=256, replace=False)
    prob_ids = [[int(q) for q in rng.choice(n, size=6, replace=False)] for _ in range(8)]
    prob_bits = [[int(b) for b in rng.integers(0, 2, size=6)] for _ in range(8)]
    probs = [sim.get_probability([bool(b) for b in bits], ids) for bits, ids in zip(prob_bits, prob_ids)]
    outcomes = [sim.measure_qubits([q])[0] for q in range(n)]
    rec = pack_gates(cmds)
    rec.update(
        n=np.int64(n),
        seed=np.int64(2020),
        sample_idx=np.asarray(sample_idx, np.int64),
        sample_amp=state[sample_idx],
        norm=np.float64(np.vdot(state, state).real),
        prob_ids=np.asarray(prob_ids, np.int64),
        prob_bits=np.asarray(prob_bits, np.int64),
        probs=np.asarray(probs, np.float64),
        outcomes=np.asarray(outcomes, np.int64),
    )
    np.savez_compressed(os.path.join(HERE, "random20_measure.npz"), **rec)

    # 3. expectation values of a 100-term random Pauli operator on variational ansatz states
    exp_rec = {}
    for i, nq in enumerate((10, 12)):
        cmds = circuits.hardware_efficient_ansatz(nq, 6, 300 + i)
        terms = circuits.random_pauli_operator(nq, 100, 400 + i)
        sim = cpp.Simulator(1)
        commands.execute(sim, cmds, gate_fusion=True)
        val = sim.get_expectation_value(terms, list(range(nq)))
        g = pack_gates(cmds)
        exp_rec.update({f"a{i}_{k}": v for k, v in g.items()})
        exp_rec.update(pack_terms(terms, f"a{i}_h"))
        exp_rec[f"a{i}_n"] = np.int64(nq)
        exp_rec[f"a{i}_value"] = np.float64(val)
    np.savez_compressed(os.path.join(HERE, "expectation_ansatz.npz"), **exp_rec)

    # 4. emulate_math through the reference's generic callback path, on a seeded state
    wf_rng = np.random.default_rng(77)
    nm = 9
    wf = wf_rng.normal(size=2**nm) + 1j * wf_rng.normal(size=2**nm)
    wf /= np.linalg.norm(wf)
    math_cases = [
        ("add_quantum", G.add_quantum(4), [[0, 1, 2, 3], [4, 5, 6, 7]], []),
        ("add_quantum_carry", G.add_quantum(3), [[0, 1, 2], [3, 4, 5], [6]], [7]),
        ("subtract", G.subtract_quantum(), [[0, 1, 2], [3, 4, 5]], [6]),
        ("comparator", G.comparator_quantum(), [[0, 1, 2], [3, 4, 5], [6]], []),
        ("multiply", G.multiply_quantum(), [[0, 1], [2, 3], [4, 5, 6, 7, 8]], []),
        ("divide", G.divide_quantum(), [[0, 1, 2], [3, 4, 5], [6, 7, 8]], []),
    ]
    mrec = {"wf": wf, "n": np.int64(nm)}
    for name, fn, regs, ctrl in math_cases:
        sim = cpp.Simulator(1)
        for q in range(nm):
            sim.allocate_qubit(q)
        sim.set_wavefunction(list(wf), list(range(nm)))
        sim.emulate_math(fn, regs, ctrl)
        mrec[f"{name}_out"] = np.asarray(sim.cheat()[1], np.complex128)
    for which, a, mod in (("add", 5, 0), ("add_mod", 4, 13), ("mul_mod", 7, 15)):
        sim = cpp.Simulator(1)
        for q in range(nm):
            sim.allocate_qubit(q)
        sim.set_wavefunction(list(wf), list(range(nm)))
        commands.execute(sim, [("math_const", which, a, mod, [[0, 1, 2, 3]], [7])])
        mrec[f"const_{which}_out"] = np.asarray(sim.cheat()[1], np.complex128)
    np.savez_compressed(os.path.join(HERE, "emulate_math.npz"), **mrec)

    # 5. time evolution (controlled) and apply_qubit_operator on a seeded 8-qubit state
    n8 = 8
    wf8 = wf_rng.normal(size=2**n8) + 1j * wf_rng.normal(size=2**n8)
    wf8 /= np.linalg.norm(wf8)
    h_terms = circuits.random_pauli_operator(6, 6, 55) + [([], 0.35)]
    sim = cpp.Simulator(1)
    for q in range(n8):
        sim.allocate_qubit(q)
    sim.set_wavefunction(list(wf8), list(range(n8)))
    sim.emulate_time_evolution(h_terms, 0.7, list(range(6)), [7])
    te_out = np.asarray(sim.cheat()[1], np.complex128)
    op_terms = [(t, c * (0.5 - 0.25j)) for t, c in circuits.random_pauli_operator(8, 12, 56)]
    sim2 = cpp.Simulator(1)
    for q in range(n8):
        sim2.allocate_qubit(q)
    sim2.set_wavefunction(list(wf8), list(range(n8)))
    sim2.apply_qubit_operator(op_terms, list(range(n8)))
    ao_out = np.asarray(sim2.cheat()[1], np.complex128)
    trec = {"wf": wf8, "n": np.int64(n8), "time": np.float64(0.7), "te_out": te_out, "ao_out": ao_out}
    trec.update(pack_terms(h_terms, "h"))
    trec.update(pack_terms(op_terms, "op"))
    np.savez_compressed(os.path.join(HERE, "pauli_ops.npz"), **trec)

    # 6. the QFT + rotations + AddQuantum workload at 14 qubits: QFT, 6 layers of random
    # rotations and controlled-Rz (1-3 controls), then b <- (a + b) mod 2^7 on the two halves
    # through the generic emulate_math callback; probabilities of 8 random 6-qubit patterns and
    # seeded measurement outcomes of every qubit
    n = 14
    cmds, queries = circuits.qft_rotations_adder(n, 6, 7, 1401, n_probabilities=8, prob_qubits=6, measure=False)
    sim = cpp.Simulator(1401)
    commands.execute(sim, cmds, gate_fusion=False)
    _, state = sim.cheat()
    probs = [sim.get_probability(bits, ids) for bits, ids in queries]
    outcomes = [sim.measure_qubits([q])[0] for q in range(n)]
    rec = pack_gates(cmds)
    rec.update(
        n=np.int64(n),
        seed=np.int64(1401),
        state=np.asarray(state, np.complex128),
        prob_ids=np.asarray([ids for _, ids in queries], np.int64),
        prob_bits=np.asarray([[int(b) for b in bits] for bits, _ in queries], np.int64),
        probs=np.asarray(probs, np.float64),
        outcomes=np.asarray(outcomes, np.int64),
    )
    np.savez_compressed(os.path.join(HERE, "qft_rotations_adder14.npz"), **rec)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
