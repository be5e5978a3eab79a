"""A backend with the `_cppsim.Simulator` method surface that records every call it receives.

ProjectQ's own test fixture swaps `Simulator._simulator` for another object with this surface
(projectq/backends/_sim/_simulator_test.py:80-93); a recorder in that seat captures the exact
backend-call sequence a compiler front end emits.  The calls are stored as JSON-ready lists so a
sequence recorded from ProjectQ's MainEngine can be committed as a fixture and compared, element
by element, with the one `projectq_amd.compiler` emits for the same program:

    ["allocate_qubit", id]                  ["deallocate_qubit", id]
    ["measure_qubits", ids]                 ["run"]
    ["apply_controlled_gate", rows, ids, ctrl]        rows: [[[re, im], ...], ...]
    ["emulate_math", table, quregs, ctrl]             table: f over every register value
    ["emulate_math_addConstant", a, quregs, ctrl]     (and the ModN / multiply variants)
    ["emulate_time_evolution", terms, time, ids, ctrl]
"""

import itertools


def _cplx(z):
    z = complex(z)
    return [z.real, z.imag]


def _regs(quregs):
    return [[int(q) for q in reg] for reg in quregs]


class Recorder:
    """Records calls; measurements return False (the compiler never reads them back)."""

    # a generic math function is tabulated over every register value up to this many bits
    MAX_TABLE_BITS = 14

    def __init__(self):
        self.calls = []

    def allocate_qubit(self, qubit_id):
        self.calls.append(["allocate_qubit", int(qubit_id)])

    def deallocate_qubit(self, qubit_id):
        self.calls.append(["deallocate_qubit", int(qubit_id)])

    def measure_qubits(self, ids):
        ids = [int(i) for i in ids]
        self.calls.append(["measure_qubits", ids])
        return [False] * len(ids)

    def apply_controlled_gate(self, matrix, ids, ctrl):
        rows = [[_cplx(v) for v in row] for row in matrix]
        self.calls.append(["apply_controlled_gate", rows, [int(i) for i in ids], [int(c) for c in ctrl]])

    def run(self):
        self.calls.append(["run"])

    def emulate_math(self, f, quregs, ctrl):
        regs = _regs(quregs)
        bits = sum(len(r) for r in regs)
        if bits > self.MAX_TABLE_BITS:
            raise ValueError("register too wide to tabulate")
        table = []
        for values in itertools.product(*[range(1 << len(r)) for r in regs]):
            table.append([int(v) for v in f(list(values))])
        self.calls.append(["emulate_math", table, regs, [int(c) for c in ctrl]])

    def emulate_math_addConstant(self, a, quregs, ctrl):  # noqa: N802 -- reference surface
        self.calls.append(["emulate_math_addConstant", int(a), _regs(quregs), [int(c) for c in ctrl]])

    def emulate_math_addConstantModN(self, a, N, quregs, ctrl):  # noqa: N802,N803
        self.calls.append(["emulate_math_addConstantModN", int(a), int(N), _regs(quregs), [int(c) for c in ctrl]])

    def emulate_math_multiplyByConstantModN(self, a, N, quregs, ctrl):  # noqa: N802,N803
        self.calls.append(
            ["emulate_math_multiplyByConstantModN", int(a), int(N), _regs(quregs), [int(c) for c in ctrl]]
        )

    def emulate_time_evolution(self, terms, time, ids, ctrl):
        out = [[[[int(q), str(p)] for q, p in term], float(c)] for term, c in terms]
        self.calls.append(["emulate_time_evolution", out, float(time), [int(i) for i in ids], [int(c) for c in ctrl]])
