"""Integer functions behind `emulate_math`, and recognition of known Python math functions.

The reference evaluates every BasicMathGate other than AddConstant / AddConstantModN /
MultiplyByConstantModN through a Python callback, once per basis index
(projectq/backends/_sim/_simulator.py:361-397 -> simulator.hpp:224-269).  The functions of
projectq/libs/math/_gates.py are pure integer maps, so the simulator runs them natively as a
basis-state permutation instead.  `recognise(f, quregs)` maps such a function to its opcode,
but only after evaluating `f` itself on a sample of register tuples and finding it equal to the
native definition on every one; otherwise the caller falls back to the callback path.  The
native definition therefore only changes speed, never answers.
"""

import random

# opcodes: keep equal to the QS_MATH_* constants of include/projectq_amd.h
ADD_CONSTANT = 0
ADD_CONSTANT_MOD_N = 1
MULTIPLY_BY_CONSTANT_MOD_N = 2
ADD_QUANTUM = 3
SUBTRACT_QUANTUM = 4
INVERSE_ADD_QUANTUM = 5
COMPARATOR = 6
MULTIPLY_QUANTUM = 7
INVERSE_MULTIPLY_QUANTUM = 8
DIVIDE_QUANTUM = 9
INVERSE_DIVIDE_QUANTUM = 10

_INT32_MIN, _INT32_MAX = -(2**31), 2**31 - 1


def _wrap32(x):
    return ((x + 2**31) % 2**32) - 2**31


def _cmod(a, b):
    """C++ `%` on ints (truncates toward zero)."""
    r = abs(a) % abs(b)
    return -r if a < 0 else r


def apply_op(opcode, params, widths, values):
    """Reference semantics of each opcode on one tuple of register values (ints).

    The constant forms follow the C++ lambdas of simulator.hpp:271-290 (int arithmetic); the
    quantum forms follow the Python functions of projectq/libs/math/_gates.py (exact integers).
    """
    v = list(values)
    if opcode == ADD_CONSTANT:
        return [_wrap32(x + _wrap32(params[0])) for x in v]
    if opcode == ADD_CONSTANT_MOD_N:
        return [_wrap32(_cmod(_wrap32(x + _wrap32(params[0])), _wrap32(params[1]))) for x in v]
    if opcode == MULTIPLY_BY_CONSTANT_MOD_N:
        return [_wrap32(_cmod(_wrap32(x * _wrap32(params[0])), _wrap32(params[1]))) for x in v]
    if opcode == ADD_QUANTUM:  # _gates.py:255-263
        n = widths[0]
        v[1] = v[0] + v[1]
        if len(bin(v[1])[2:]) > n:
            v[1] = v[1] % (2**n)
            if len(v) == 3:
                v[2] ^= 1
        return v
    if opcode == INVERSE_ADD_QUANTUM:  # _gates.py:287-293
        if len(v) == 3:
            v[2] ^= 1
        v[1] -= v[0]
        return v
    if opcode == SUBTRACT_QUANTUM:  # _gates.py:322-323
        return [v[0], v[1] - v[0]]
    if opcode == COMPARATOR:  # _gates.py:372-379
        a, b, c = v
        if b < a:
            c = 1 if c == 0 else 0
        return [a, b, c]
    if opcode == MULTIPLY_QUANTUM:  # _gates.py:516-517
        return [v[0], v[1], v[2] + v[0] * v[1]]
    if opcode == INVERSE_MULTIPLY_QUANTUM:  # _gates.py:547-548
        return [v[0], v[1], v[2] - v[0] * v[1]]
    if opcode == DIVIDE_QUANTUM:  # _gates.py:438-442
        dividend, remainder, divisor = v
        if divisor == 0 or divisor > dividend:
            return [remainder, dividend, divisor]
        quotient = remainder + dividend // divisor
        return [dividend - quotient * divisor, quotient, divisor]
    if opcode == INVERSE_DIVIDE_QUANTUM:  # _gates.py:473-478
        remainder, quotient, divisor = v
        if divisor == 0:
            return [quotient, remainder, divisor]
        return [remainder + quotient * divisor, 0, divisor]
    raise ValueError(f"unknown math opcode {opcode}")


# qualified names of the functions projectq/libs/math/_gates.py hands to emulate_math, and
# the number of registers each takes (None = 2 or 3)
_DIRECT = {
    "AddQuantumGate.get_math_function.<locals>.math_fun": (ADD_QUANTUM, None),
    "_InverseAddQuantumGate.get_math_function.<locals>.math_fun": (INVERSE_ADD_QUANTUM, None),
}
_WRAPPED = {
    "SubtractQuantumGate.__init__.<locals>.subtract": (SUBTRACT_QUANTUM, 2),
    "ComparatorQuantumGate.__init__.<locals>.compare": (COMPARATOR, 3),
    "MultiplyQuantumGate.__init__.<locals>.multiply": (MULTIPLY_QUANTUM, 3),
    "_InverseMultiplyQuantumGate.__init__.<locals>.inverse_multiplication": (INVERSE_MULTIPLY_QUANTUM, 3),
    "DivideQuantumGate.__init__.<locals>.division": (DIVIDE_QUANTUM, 3),
    "_InverseDivideQuantumGate.__init__.<locals>.inverse_division": (INVERSE_DIVIDE_QUANTUM, 3),
}


def _candidate(f, n_regs):
    tagged = getattr(f, "_qs_math", None)
    if tagged is not None:
        return tagged
    qual = getattr(f, "__qualname__", "")
    if qual in _DIRECT:
        op, _ = _DIRECT[qual]
        return (op, []) if n_regs in (2, 3) else None
    if qual == "BasicMathGate.__init__.<locals>.math_function" and getattr(f, "__closure__", None):
        for cell in f.__closure__:
            inner = cell.cell_contents
            if callable(inner) and getattr(inner, "__qualname__", "") in _WRAPPED:
                op, nr = _WRAPPED[inner.__qualname__]
                return (op, []) if n_regs == nr else None
    return None


def _samples(widths, count, seed=12345):
    rng = random.Random(seed)
    tuples = [[0] * len(widths), [(1 << w) - 1 for w in widths], [1 << max(w - 1, 0) for w in widths]]
    for _ in range(count):
        tuples.append([rng.randrange(1 << w) if w else 0 for w in widths])
    return tuples


def recognise(f, quregs, n_samples=48):
    """Return (opcode, params) if `f` is a known math function on this register layout, else None."""
    widths = [len(r) for r in quregs]
    if not widths or max(widths) > 30:
        return None
    cand = _candidate(f, len(widths))
    if cand is None:
        return None
    opcode, params = cand
    try:
        for vals in _samples(widths, n_samples):
            want = apply_op(opcode, params, widths, vals)
            got = f(list(vals))
            got = [int(x) for x in got][: len(widths)]
            if got != want:
                return None
            if any(x < _INT32_MIN or x > _INT32_MAX for x in want):
                return None
    except Exception:  # noqa: BLE001 -- any surprise means "not recognised"
        return None
    return opcode, list(params)
