"""Parity at the measured configurations' full sizes and on the large-state code paths (all `gpu`).

* get_expectation_value on states above the 64 MiB cache threshold (one streaming pass per
  string, identity terms through the norm) against the oracle, complex128 and complex64;
* get_classical_value, is_classical and the full state after every deallocate_qubit against the
  oracle;
* the 30-qubit QFT + rotations + AddQuantum workload: window fusion with the learned planner
  against the plain path (run() after every gate), both complex128, amplitude by amplitude on the
  GPU; AddQuantum moves amplitudes bit-exactly; probabilities and seeded outcomes agree;
* the 32-qubit depth-100 random circuit in complex64 (learned planner, pipelined tile pass)
  against complex128 (tile passes) and against complex64 greedy max-5 blocks (the dense-block
  kernel, a different planner and kernel): infidelity <= 1e-6 both ways;
* a batch shaped like the 512-ansatz variational workload (16 of its circuits, including n = 20,
  100 Pauli strings each) against the oracle's single-circuit energies.

The oracle (oracle/cpu_sim.cpp) is the checker where it fits the host; at 30 and 32 qubits the
complex128 GPU paths pinned to it at smaller sizes are the reference.
"""

import numpy as np
import pytest

from oracle.cpu_sim import OracleSimulator
from projectq_amd import Simulator, circuits, commands

pytestmark = pytest.mark.gpu


def _gates(cmds):
    return [c for c in cmds if c[0] == "gate"]


def _submit(sim, gates, run_each):
    for _, m, t, c in gates:
        sim.apply_controlled_gate(m, list(t), list(c))
        if run_each:
            sim.run()
    sim.run()
    sim.synchronize()


@pytest.mark.parametrize("precision,n", [("c128", 24), ("c64", 25)])
def test_expectation_on_large_states(precision, n):
    """States of 256 MiB take the per-string streaming path (k_pauli_expect) and the norm pass for
    identity terms; the reference's loop is simulator.hpp:292-322."""
    print(f"expectation {precision} n={n}", flush=True)
    cmds = circuits.random_layers(n, 2, n // 2, 61, measure=False)
    o = OracleSimulator(3)
    commands.execute(o, cmds, gate_fusion=False)
    h = Simulator(3, precision=precision, capacity_qubits=n)
    commands.execute(h, cmds, gate_fusion=False)
    terms = circuits.random_pauli_operator(n, 10, 62)
    terms += [([], 0.75), ([], -0.25), ([(0, "Y"), (n - 1, "Y")], 0.5), ([(3, "Z")], 1.25),
              ([(n - 1, "X")], 0.6), ([(0, "X")], -0.4), ([(5, "X"), (5, "Y")], 0.8)]
    ids = list(range(n))
    rng = np.random.default_rng(63)
    perm = [int(q) for q in rng.permutation(n)]  # local index -> qubit id not the identity
    for order in (ids, perm):
        want = o.get_expectation_value(terms, order)
        got = h.get_expectation_value(terms, order)
        scale = sum(abs(c) for _, c in terms)
        tol = 1e-12 if precision == "c128" else 2e-5
        assert abs(got - want) <= tol * scale, (got, want)


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_classical_values_and_states_after_each_deallocation(precision):
    """Measure some qubits of an entangled state, query is_classical / get_classical_value on every
    qubit, then deallocate the classical qubits one by one in a random order (bottom, middle and
    top positions) and compare the whole state after each deallocation (simulator.hpp:76-108,
    :110-143, :194-202)."""
    n, seed = 11, 71
    cmds = circuits.random_layers(n, 5, 5, seed, measure=False)
    o = OracleSimulator(seed)
    h = Simulator(seed, precision=precision)
    for sim in (o, h):
        commands.execute(sim, cmds, gate_fusion=False)
    tol = 1e-12 if precision == "c128" else 2e-6
    measured = [0, 4, 10, 7, 2]
    for q in measured:
        assert o.measure_qubits([q]) == h.measure_qubits([q])
    # a qubit put into |1> after measurement: get_classical_value must say True
    from projectq_amd import gates as G

    for sim in (o, h):
        commands.execute(sim, [commands.gate(G.X, [4])])
    for q in range(n):
        for t in (1e-12, 1e-6):
            assert o.is_classical(q, t) == h.is_classical(q, t), (q, t)
        if o.is_classical(q, 1e-12):
            assert o.get_classical_value(q, 1e-12) == h.get_classical_value(q, 1e-12), q
    rng = np.random.default_rng(seed)
    for q in [int(x) for x in rng.permutation(measured)]:
        o.deallocate_qubit(q)
        h.deallocate_qubit(q)
        mo, vo = o.cheat_numpy()
        mh, vh = h.cheat_numpy()
        assert mo == mh
        assert np.max(np.abs(vo - vh)) <= tol, q
    # a superposed qubit cannot be deallocated (RuntimeError, as the reference throws)
    left = sorted(set(range(n)) - set(measured))
    if not o.is_classical(left[0], 1e-12):
        with pytest.raises(RuntimeError):
            h.deallocate_qubit(left[0])


def test_qft_rotations_adder_30q_learned_vs_plain():
    """30 qubits, complex128: QFT + 50 layers of rotations and controlled-Rz (2715 gates), then
    AddQuantum on the two 15-qubit halves, 8 probabilities and a seeded measurement of every
    qubit.  Window fusion with the learned planner against the plain path (run() after every gate,
    one dense pass per gate) at <= 1e-12 per amplitude."""
    n, seed = 30, 2024
    print("30-qubit learned vs plain", flush=True)
    cmds, queries = circuits.qft_rotations_adder(n, 50, 15, seed, n_probabilities=8, prob_qubits=8, measure=False)
    gates = _gates(cmds)
    _, fn, regs, ctrl = [c for c in cmds if c[0] == "math"][0]
    plain = Simulator(seed, precision="c128", fusion="reference", capacity_qubits=n)
    learned = Simulator(seed, precision="c128", fusion="window", planner="learned", capacity_qubits=n)
    for sim in (plain, learned):
        for q in range(n):
            sim.allocate_qubit(q)
    _submit(plain, gates, run_each=True)
    _submit(learned, gates, run_each=False)
    assert learned.last_plan()["kind"] == "tile"
    diff, unequal = learned.compare(plain)
    print(f"gates: max |diff| {diff:.3e}, {unequal} amplitudes differ in their last bits", flush=True)
    assert diff <= 1e-12
    # AddQuantum is a permutation: both paths move their amplitudes bit for bit, so the set of
    # differences is permuted, not changed
    ids = list(range(n))
    rng = np.random.default_rng(5)
    picks = [int(v) for v in rng.integers(0, 2**n, size=32)]
    bits_of = lambda v: [(v >> q) & 1 for q in ids]  # noqa: E731
    before = [learned.get_amplitude(bits_of(v), ids) for v in picks]
    for sim in (plain, learned):
        sim.emulate_math(fn, regs, ctrl)
    assert learned.compare(plain) == (diff, unequal)
    mask = (1 << 15) - 1
    moved = [(v & mask) | ((((v & mask) + (v >> 15)) & mask) << 15) for v in picks]
    assert [learned.get_amplitude(bits_of(v), ids) for v in moved] == before
    for bits, qids in queries:
        assert abs(learned.get_probability(bits, qids) - plain.get_probability(bits, qids)) <= 1e-13
    assert [learned.measure_qubits([q])[0] for q in ids] == [plain.measure_qubits([q])[0] for q in ids]
    del plain, learned


def test_headline_32q_depth100_infidelity():
    """The measured configuration: 32 qubits, depth 100 (H, a random Rx/Ry/Rz per qubit and 16
    CNOTs per layer, 8000 gates), complex64 with the learned planner.  Against complex128 tile
    passes and against complex64 greedy max-5 blocks (k_apply: another planner and another
    kernel, so an addressing mistake shared by the tile paths cannot cancel): 1 - |<a|b>|^2 <= 1e-6."""
    n, seed = 32, 2024
    print("32-qubit depth-100 infidelity", flush=True)
    gates = _gates(circuits.random_layers(n, 100, 16, seed, measure=False))

    def run(precision, planner):
        sim = Simulator(seed, precision=precision, fusion="window", planner=planner, capacity_qubits=n)
        for q in range(n):
            sim.allocate_qubit(q)
        _submit(sim, gates, run_each=False)
        print(f"{precision} {planner}: plan {sim.last_plan()}", flush=True)
        return sim

    def infid(a, b):
        return 1.0 - abs(a.overlap(b)) ** 2 / (a.overlap(a).real * b.overlap(b).real)

    learned = run("c64", "learned")
    plan = learned.last_plan()
    assert plan["kind"] == "tile" and plan["tile_bits"] == 12 and plan["reg_bits"] == 4
    ref = run("c128", "tile")
    a = infid(ref, learned)
    raw = 1.0 - abs(ref.overlap(learned)) ** 2  # norm error included
    print(f"c64 learned vs c128 tile: {a:.3e} ({raw:.3e} with the norm error)", flush=True)
    assert a <= 1e-6 and raw <= 1e-6
    greedy = run("c64", "greedy")
    assert greedy.last_plan()["kind"] == "greedy_blocks"
    b = infid(ref, greedy)
    c = infid(learned, greedy)
    print(f"c64 greedy vs c128 tile: {b:.3e}; vs c64 learned: {c:.3e}", flush=True)
    assert b <= 1e-6 and c <= 1e-6
    del learned, ref, greedy


def test_batch_shaped_like_the_variational_workload():
    """16 circuits drawn from the 512-ansatz workload's generator (same seeds: n in U{10..20},
    6 layers of Ry, Rz and a CNOT chain, 100 random Pauli strings with N(0, 1) coefficients),
    forced to include n = 20, in complex64 with FP64 sums against the oracle's single-circuit
    energies; graph replays are bit-identical."""
    from projectq_amd.batch import BatchExpectation

    seed = 2024
    rng = np.random.default_rng(seed + 3)
    sizes = [int(x) for x in rng.integers(10, 21, size=512)]
    pick = sorted(set([sizes.index(20)] + [int(i) for i in np.random.default_rng(9).choice(512, size=15, replace=False)]))
    circ, ops = [], []
    for i in pick:
        n = sizes[i]
        circ.append((n, _gates(circuits.hardware_efficient_ansatz(n, 6, seed + 10 + i))))
        ops.append(circuits.random_pauli_operator(n, 100, seed + 5000 + i))
    assert max(n for n, _ in circ) == 20
    b = BatchExpectation(circ, ops, precision="c64", autotune=True)
    got, _ = b.run()
    for _ in range(3):
        again, _ = b.run()
        assert np.array_equal(again, got)
    for (n, gl), terms, e in zip(circ, ops, got):
        sim = OracleSimulator(1)
        commands.execute(sim, circuits.allocate(n) + gl)
        want = sim.get_expectation_value(terms, list(range(n)))
        scale = sum(abs(c) for _, c in terms)
        assert abs(e - want) <= 2e-5 * scale, (n, e, want)
