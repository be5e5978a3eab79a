"""Command streams and their execution against a `_cppsim.Simulator`-surface backend.

`execute` restates the command -> backend-call mapping of ProjectQ's `Simulator._handle` and
`receive` (projectq/backends/_sim/_simulator.py:324-438) for the commands a circuit with an
empty engine list produces:

* ("allocate", id)                              -> allocate_qubit            (:355-357)
* ("deallocate", id)                            -> deallocate_qubit          (:358-360)
* ("measure", id)                               -> measure_qubits([id])      (:338-354)
* ("gate", matrix, targets, controls)           -> apply_controlled_gate, then run() unless
                                                   gate_fusion               (:404-415)
* ("math_const", kind, a, N, quregs, controls)  -> emulate_math_addConstant / ModN /
                                                   multiplyByConstantModN    (:374-395)
* ("math", f, quregs, controls)                 -> emulate_math(f, ...)      (:396-397)
* ("time_evolution", terms, time, ids, controls)-> emulate_time_evolution    (:398-403)
* ("flush",)                                    -> run()                     (:435-436)

Controls are sorted by qubit id, as ProjectQ's Command does (projectq/ops/_command.py:231-232).
"""

import numpy as np


def gate(matrix, targets, controls=()):
    return ("gate", np.asarray(matrix, dtype=np.complex128), tuple(int(t) for t in targets), tuple(sorted(int(c) for c in controls)))


def execute(backend, commands, gate_fusion=False):
    """Run `commands` on `backend`; return {qubit id: measured bool} in measurement order."""
    results = {}
    for cmd in commands:
        kind = cmd[0]
        if kind == "gate":
            _, m, targets, controls = cmd
            backend.apply_controlled_gate(m, list(targets), list(controls))
            if not gate_fusion:
                backend.run()
        elif kind == "allocate":
            backend.allocate_qubit(cmd[1])
        elif kind == "deallocate":
            backend.deallocate_qubit(cmd[1])
        elif kind == "measure":
            out = backend.measure_qubits([cmd[1]])
            results[cmd[1]] = bool(out[0])
        elif kind == "math_const":
            _, which, a, n_mod, quregs, controls = cmd
            ctrl = sorted(controls)
            if which == "add":
                backend.emulate_math_addConstant(a, quregs, ctrl)
            elif which == "add_mod":
                backend.emulate_math_addConstantModN(a, n_mod, quregs, ctrl)
            elif which == "mul_mod":
                backend.emulate_math_multiplyByConstantModN(a, n_mod, quregs, ctrl)
            else:
                raise ValueError(which)
        elif kind == "math":
            _, fn, quregs, controls = cmd
            backend.emulate_math(fn, quregs, sorted(controls))
        elif kind == "time_evolution":
            _, terms, time, ids, controls = cmd
            backend.emulate_time_evolution(terms, time, list(ids), sorted(controls))
        elif kind == "flush":
            backend.run()
        else:
            raise ValueError(f"unknown command {kind!r}")
    return results


def count_gates(commands):
    """Circuit gates (matrix gates and math gates), the unit of the gates/s metric."""
    return sum(1 for c in commands if c[0] in ("gate", "math", "math_const"))
