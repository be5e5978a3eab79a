"""Batched variational circuits: per-circuit energies must equal the oracle's single-circuit
results (complex128 to 1e-10, complex64 to 2e-5 relative to sum |c_t|); the hipGraph launch,
eager launches and autotuned shapes must all give the same answers."""

import os

import numpy as np
import pytest

from oracle.cpu_sim import OracleBatch, OracleSimulator
from projectq_amd import circuits, commands


def make_batch(sizes, seed):
    circ, ops = [], []
    for i, n in enumerate(sizes):
        cmds = circuits.hardware_efficient_ansatz(n, 6, seed + i)
        circ.append((n, [c for c in cmds if c[0] == "gate"]))
        ops.append(circuits.random_pauli_operator(n, 30, seed + 1000 + i))
    return circ, ops


def test_oracle_batch_matches_single_circuits():
    circ, ops = make_batch([3, 5, 7, 4], 11)
    energies, _ = OracleBatch(circ, ops).run()
    for (n, cmds), terms, e in zip(circ, ops, energies):
        sim = OracleSimulator(1)
        commands.execute(sim, circuits.allocate(n) + cmds)
        assert abs(sim.get_expectation_value(terms, list(range(n))) - e) <= 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("autotune", [False, True])
def test_batch_c128_matches_oracle(autotune):
    from projectq_amd.batch import BatchExpectation

    rng = np.random.default_rng(5)
    sizes = [int(x) for x in rng.integers(3, 15, size=24)]
    circ, ops = make_batch(sizes, 21)
    want, _ = OracleBatch(circ, ops).run()
    b = BatchExpectation(circ, ops, precision="c128", autotune=autotune)
    got, ms = b.run()
    print("shape", b.shape(), "gpu ms", ms, flush=True)
    assert np.max(np.abs(got - want)) <= 1e-10


@pytest.mark.gpu
def test_batch_c64_accuracy_and_graph_replays():
    from projectq_amd import Simulator
    from projectq_amd.batch import BatchExpectation

    rng = np.random.default_rng(6)
    sizes = [int(x) for x in rng.integers(4, 17, size=32)]
    circ, ops = make_batch(sizes, 31)
    want, _ = OracleBatch(circ, ops).run()
    scale = np.array([sum(abs(c) for _, c in terms) for terms in ops])
    b = BatchExpectation(circ, ops, precision="c64", autotune=True)
    eager, _ = b.run(graph=False)
    assert np.max(np.abs(eager - want) / scale) <= 2e-5
    for _ in range(4):  # graph replays, with unrelated eager GPU work in between
        got, _ = b.run(graph=True)
        assert np.array_equal(got, eager)
        s = Simulator(1)
        for q in range(12):
            s.allocate_qubit(q)
        s.run()
        s.synchronize()


@pytest.mark.gpu
def test_batch_shapes_agree():
    """Every candidate tile shape gives the same energies (the autotuner only changes speed)."""
    from projectq_amd.batch import BatchExpectation

    circ, ops = make_batch([12, 13, 14, 9, 16], 41)
    a, _ = BatchExpectation(circ, ops, precision="c128", autotune=False).run()
    b, _ = BatchExpectation(circ, ops, precision="c128", autotune=True).run()
    assert np.max(np.abs(a - b)) <= 1e-12
    # every candidate on its own, including the shapes whose tile stores relabel qubits (the
    # Pauli strings then act on each circuit's final layout)
    relabelled = 0
    for k in range(9):
        os.environ["PROJECTQ_AMD_BATCH_SHAPE"] = str(k)
        try:
            for prec, tol in (("c128", 1e-12), ("c64", 2e-5)):
                try:
                    batch = BatchExpectation(circ, ops, precision=prec, autotune=False)
                except ValueError:
                    continue  # fewer candidates for this precision
                e, _ = batch.run()
                scale = np.array([sum(abs(c) for _, c in op) for op in ops])
                assert np.all(np.abs(e - a) <= tol * scale), (k, prec, batch.shape())
                relabelled += batch.shape()["relabel"]
        finally:
            del os.environ["PROJECTQ_AMD_BATCH_SHAPE"]
    assert relabelled >= 2
