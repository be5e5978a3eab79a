"""Pins of the CPU oracle (oracle/cpu_sim.cpp) beyond the golden vectors of test_golden.py.

What each check pins:
* compiled reference (oracle/_ref, when built here): oracle == reference on random circuits,
  controlled k-qubit unitaries, fused and unfused, including seeded measurements;
* gates: brute-force dense unitaries (numpy kron of the embedded gate, controls as a projector
  sum) on up to 7 qubits -- catches a transposed matrix, a wrong bit order or control sense;
* measurement: mt19937 restated in numpy (pinned by the C++ standard's 10000th output) plus
  libstdc++'s generate_canonical<double,53> and the reference's serial pick rule
  (simulator.hpp:145-186) predict every seeded outcome;
* expectation values: numpy <psi|P|psi> with dense Pauli matrices;
* math functions: the known answers of projectq/libs/math/_gates_math_test.py.
"""

import math

import numpy as np
import pytest

from oracle import ref
from oracle.cpu_sim import OracleSimulator
from projectq_amd import circuits, commands
from projectq_amd import gates as G

PAULI = {
    "X": np.array([[0, 1], [1, 0]], dtype=complex),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=complex),
    "Z": np.array([[1, 0], [0, -1]], dtype=complex),
}


def dense_gate(n, m, targets, controls):
    """Full 2^n unitary of a controlled k-qubit gate by explicit index arithmetic."""
    dim = 2**n
    k = len(targets)
    u = np.zeros((dim, dim), dtype=complex)
    for col in range(dim):
        if any(((col >> c) & 1) == 0 for c in controls):
            u[col, col] = 1.0
            continue
        jc = sum(((col >> targets[l]) & 1) << l for l in range(k))
        for jr in range(2**k):
            row = col
            for l in range(k):
                row = (row & ~(1 << targets[l])) | (((jr >> l) & 1) << targets[l])
            u[row, col] += m[jr, jc]
    return u


def test_dense_gates_brute_force():
    rng = np.random.default_rng(11)
    n = 7
    psi = np.zeros(2**n, dtype=complex)
    psi[0] = 1.0
    sim = OracleSimulator(1)
    for q in range(n):
        sim.allocate_qubit(q)
    for _ in range(30):
        k = int(rng.integers(1, 5))
        picks = rng.choice(n, size=k + int(rng.integers(0, 3)), replace=False)
        a = rng.normal(size=(2**k, 2**k)) + 1j * rng.normal(size=(2**k, 2**k))
        u, _ = np.linalg.qr(a)
        t, c = [int(x) for x in picks[:k]], [int(x) for x in picks[k:]]
        sim.apply_controlled_gate(u, t, c)
        sim.run()
        psi = dense_gate(n, u, t, c) @ psi
    assert np.max(np.abs(sim.cheat_numpy()[1] - psi)) <= 1e-12


def test_kron_convention():
    """Matrix bit l belongs to ids[l]: kron(m2, m1) on [q0, q1] == m1 on q0 then m2 on q1
    (the convention of projectq/backends/_sim/_simulator_test.py:287-314)."""
    m1, m2 = G.Rx(0.3), G.Ry(0.8)
    a, b = OracleSimulator(1), OracleSimulator(1)
    for s in (a, b):
        for q in range(3):
            s.allocate_qubit(q)
        s.apply_controlled_gate(G.H, [2], [])
        s.run()
    a.apply_controlled_gate(np.kron(m2, m1), [0, 1], [])
    a.run()
    b.apply_controlled_gate(m1, [0], [])
    b.apply_controlled_gate(m2, [1], [])
    b.run()
    assert np.max(np.abs(a.cheat_numpy()[1] - b.cheat_numpy()[1])) <= 1e-15


class MT19937:
    """std::mt19937 (C++ [rand.predef]) restated; 10000th output of the default seed is 4123659995."""

    def __init__(self, seed):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xFFFFFFFF
        self.i = 624

    def __call__(self):
        if self.i >= 624:
            for k in range(624):
                y = (self.mt[k] & 0x80000000) | (self.mt[(k + 1) % 624] & 0x7FFFFFFF)
                self.mt[k] = self.mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF

    def uniform(self):
        """libstdc++ generate_canonical<double, 53>: two 32-bit draws, low word first."""
        lo, hi = self(), self()
        r = (lo + hi * 2.0**32) / 2.0**64
        return r if r < 1.0 else math.nextafter(1.0, 0.0)


def test_mt19937_restatement():
    g = MT19937(5489)
    for _ in range(9999):
        g()
    assert g() == 4123659995


@pytest.mark.parametrize("seed", [1, 7, 2024, 4294967295])
def test_measurement_rule(seed):
    rng = np.random.default_rng(seed % 1000)
    n = 6
    sim = OracleSimulator(seed)
    for q in range(n):
        sim.allocate_qubit(q)
    wf = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    wf /= np.linalg.norm(wf)
    sim.set_wavefunction(wf, list(range(n)))
    g = MT19937(seed)
    state = wf.copy()
    for q in [3, 0, 5, 1, 4, 2]:
        r = g.uniform()
        cum = np.cumsum(np.abs(state) ** 2)
        pick = int(np.argmax(cum >= r)) if cum[-1] >= r else len(state) - 1
        bit = (pick >> q) & 1
        assert sim.measure_qubits([q])[0] == bool(bit)
        keep = np.array([((i >> q) & 1) == bit for i in range(len(state))])
        state = np.where(keep, state, 0)
        state /= math.sqrt(np.sum(np.abs(state) ** 2))
    assert np.max(np.abs(sim.cheat_numpy()[1] - state)) <= 1e-12


def test_expectation_dense():
    rng = np.random.default_rng(3)
    n = 5
    sim = OracleSimulator(1)
    for q in range(n):
        sim.allocate_qubit(q)
    wf = rng.normal(size=2**n) + 1j * rng.normal(size=2**n)
    wf /= np.linalg.norm(wf)
    sim.set_wavefunction(wf, list(range(n)))
    terms = circuits.random_pauli_operator(n, 20, 8)
    ids = [2, 0, 4, 1, 3]  # local index -> qubit id
    want = 0.0
    for term, c in terms:
        mats = [np.eye(2, dtype=complex)] * n
        for loc, p in term:
            mats[ids[loc]] = PAULI[p]
        full = mats[n - 1]
        for q in range(n - 2, -1, -1):
            full = np.kron(full, mats[q])
        want += c * np.vdot(wf, full @ wf).real
    assert abs(sim.get_expectation_value(terms, ids) - want) <= 1e-12


def prob_of(sim, bits, ids):
    return sim.get_probability([bool(b) for b in bits], ids)


def test_math_known_answers():
    """projectq/libs/math/_gates_math_test.py known answers, run through the oracle."""
    # test_addition: 4 + 8 = 12 with carry 0
    s = OracleSimulator(1)
    for q in range(11):
        s.allocate_qubit(q)
    a, b, c = list(range(5)), list(range(5, 10)), [10]
    commands.execute(s, [commands.gate(G.X, [a[2]]), commands.gate(G.X, [b[3]])])
    s.emulate_math(G.add_quantum(5), [a, b, c], [])
    assert prob_of(s, [0, 0, 1, 0, 0], a) == pytest.approx(1.0)
    assert prob_of(s, [0, 0, 1, 1, 0], b) == pytest.approx(1.0)
    assert prob_of(s, [0], c) == pytest.approx(1.0)
    # test_addition_with_control_carry: 2 + 15 = 17 -> b = 1, carry register = 1
    s = OracleSimulator(1)
    for q in range(11):
        s.allocate_qubit(q)
    a, b, ctl, cc = [0, 1, 2, 3], [4, 5, 6, 7], [8], [9, 10]
    commands.execute(s, [commands.gate(G.X, [a[1]])] + [commands.gate(G.X, [q]) for q in b] + [commands.gate(G.X, ctl)])
    s.emulate_math(G.add_quantum(4), [a, b, cc], ctl)
    assert prob_of(s, [0, 1, 0, 0], a) == pytest.approx(1.0)
    assert prob_of(s, [1, 0, 0, 0], b) == pytest.approx(1.0)
    assert prob_of(s, [1, 0], cc) == pytest.approx(1.0)
    # comparator (docstring of ComparatorQuantumGate): a=16 > b=8 flips c
    s = OracleSimulator(1)
    for q in range(11):
        s.allocate_qubit(q)
    a, b, c = list(range(5)), list(range(5, 10)), [10]
    commands.execute(s, [commands.gate(G.X, [a[4]]), commands.gate(G.X, [b[3]])])
    s.emulate_math(G.comparator_quantum(), [a, b, c], [])
    assert prob_of(s, [1], c) == pytest.approx(1.0)
    # multiply (docstring of MultiplyQuantumGate): 4 * 8 = 32 into a 9-qubit register
    s = OracleSimulator(1)
    for q in range(17):
        s.allocate_qubit(q)
    a, b, c = list(range(4)), list(range(4, 8)), list(range(8, 17))
    commands.execute(s, [commands.gate(G.X, [a[2]]), commands.gate(G.X, [b[3]])])
    s.emulate_math(G.multiply_quantum(), [a, b, c], [])
    assert prob_of(s, [0, 0, 0, 0, 0, 1, 0, 0, 0], c) == pytest.approx(1.0)
    # divide (docstring of DivideQuantumGate): 9 / 4 -> remainder 1, quotient 2, divisor 4
    s = OracleSimulator(1)
    for q in range(15):
        s.allocate_qubit(q)
    a, b, c = list(range(5)), list(range(5, 10)), list(range(10, 15))
    commands.execute(s, [commands.gate(G.X, [a[0]]), commands.gate(G.X, [a[3]]), commands.gate(G.X, [c[2]])])
    s.emulate_math(G.divide_quantum(), [a, b, c], [])
    assert prob_of(s, [1, 0, 0, 0, 0], a) == pytest.approx(1.0)
    assert prob_of(s, [0, 1, 0, 0, 0], b) == pytest.approx(1.0)
    assert prob_of(s, [0, 0, 1, 0, 0], c) == pytest.approx(1.0)


@pytest.mark.skipif(not ref.available("intrin"), reason="reference simulator not built here (make -C oracle ref)")
@pytest.mark.parametrize("fusion", [False, True])
def test_against_compiled_reference(fusion):
    REF = ref.load("intrin")  # loaded here, not at import: GPU test processes never map it
    rng = np.random.default_rng(21)
    n = 12
    cmds = circuits.random_layers(n, 6, 6, 22, measure=False)
    for _ in range(25):
        k = int(rng.integers(1, 6))
        picks = rng.choice(n, size=k + int(rng.integers(0, 4)), replace=False)
        a = rng.normal(size=(2**k, 2**k)) + 1j * rng.normal(size=(2**k, 2**k))
        u, _ = np.linalg.qr(a)
        cmds.append(commands.gate(u, [int(q) for q in picks[:k]], [int(q) for q in picks[k:]]))
    r, o = REF.Simulator(9), OracleSimulator(9)
    commands.execute(r, cmds, fusion)
    commands.execute(o, cmds, fusion)
    assert np.max(np.abs(np.asarray(r.cheat()[1]) - o.cheat_numpy()[1])) <= 1e-12
    meas = circuits.measure_all(n)
    assert commands.execute(r, meas) == commands.execute(o, meas)
