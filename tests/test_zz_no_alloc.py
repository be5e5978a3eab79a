"""qs_expectation_value_noalloc (the entry point that allocates nothing) against
qs_get_expectation_value, through the C ABI: two calls in a row on one GPU, the second with a
smaller input, on the cached single-pass path (states <= 64 MiB) and on the per-string streaming
path, in complex128 and complex64.  Same kernels in the same order, so the results agree to
1e-12 relative to sum |c_t| (in practice bit for bit)."""

import numpy as np
import pytest

from projectq_amd import Simulator, circuits, commands
from projectq_amd._binding import expectation_value_noalloc

pytestmark = pytest.mark.gpu


def _prepared(n, precision, seed):
    sim = Simulator(seed, precision=precision, capacity_qubits=n)
    commands.execute(sim, circuits.random_layers(n, 2, n // 2, seed, measure=False), gate_fusion=False)
    return sim


def _terms(n, k, seed):
    terms = circuits.random_pauli_operator(n, k, seed)
    return terms + [([], 0.5), ([(0, "Y"), (n - 1, "X")], -0.75)]


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("sizes", [(12, 10), (24, 23)], ids=["cached", "streaming"])
def test_noalloc_matches_get_expectation_value(precision, sizes):
    for call, n in enumerate(sizes):
        print(f"noalloc {precision} call {call} n={n}", flush=True)
        sim = _prepared(n, precision, 40 + n)
        ids = [int(q) for q in np.random.default_rng(n).permutation(n)]
        terms = _terms(n, 40 if call == 0 else 12, 50 + n)
        want = sim.get_expectation_value(terms, ids)
        mapping, state = sim.device_state()
        got = expectation_value_noalloc(sim._lib, state, [mapping[q] for q in ids], terms)
        scale = sum(abs(c) for _, c in terms)
        print(f"  want {want!r} got {got!r} bit-equal {got == want}", flush=True)
        assert abs(got - want) <= 1e-12 * scale
        del sim


def test_noalloc_rejects_oversized_and_bad_input():
    sim = _prepared(8, "c128", 3)
    mapping, state = sim.device_state()
    pos = [mapping[q] for q in range(8)]
    with pytest.raises(ValueError):
        expectation_value_noalloc(sim._lib, state, pos, [([(0, "Z")], 1.0)] * 4097)
    with pytest.raises(ValueError):
        expectation_value_noalloc(sim._lib, state, pos + [9], [([(8, "Z")], 1.0)])
    with pytest.raises(ValueError):
        expectation_value_noalloc(sim._lib, state, pos, [([(9, "Z")], 1.0)])
    # still usable afterwards
    terms = [([(1, "Z")], 1.0), ([], 2.0)]
    assert abs(expectation_value_noalloc(sim._lib, state, pos, terms) - sim.get_expectation_value(terms, list(range(8)))) <= 1e-12
