"""Gate matrices and angle conventions that reach the simulator boundary.

These restate what ProjectQ's operator layer hands to `apply_controlled_gate`, so that circuits
built here give the backend exactly the matrices a ProjectQ program would give it:

* matrices of projectq/ops/_gates.py:62-366 (H, X, Y, Z, S, T, SqrtX, Swap, SqrtSwap, Ph, Rx,
  Ry, Rz, Rxx, Ryy, Rzz, R), evaluated with the same numpy/cmath expressions;
* the angle canonicalisation of projectq/ops/_basics.py:321-441: rotation angles are stored as
  round(angle mod 4*pi, 12) (mod 2*pi for phase gates), and a value within 1e-12 of the period
  becomes 0;
* daggered gates use the conjugate transpose (projectq/ops/_metagates.py DaggeredGate.matrix);
* the math functions of projectq/libs/math/_gates.py, tagged with their native opcode so that
  `emulate_math` can run them as integer permutations.
"""

import cmath
import math

import numpy as np

from . import _mathfuncs

ANGLE_PRECISION = 12  # projectq/ops/_basics.py:44
ANGLE_TOLERANCE = 10**-ANGLE_PRECISION  # projectq/ops/_basics.py:45


def rotation_angle(angle):
    """BasicRotationGate.__init__ (projectq/ops/_basics.py:331-341)."""
    rounded = round(float(angle) % (4.0 * math.pi), ANGLE_PRECISION)
    if rounded > 4 * math.pi - ANGLE_TOLERANCE:
        rounded = 0.0
    return rounded


def phase_angle(angle):
    """BasicPhaseGate.__init__ (projectq/ops/_basics.py:431-441)."""
    rounded = round(float(angle) % (2.0 * math.pi), ANGLE_PRECISION)
    if rounded > 2 * math.pi - ANGLE_TOLERANCE:
        rounded = 0.0
    return rounded


def _m(rows):
    return np.array(rows, dtype=np.complex128)


# fixed gates -- projectq/ops/_gates.py:62-200
H = 1.0 / cmath.sqrt(2.0) * np.array([[1, 1], [1, -1]], dtype=np.complex128)
X = _m([[0, 1], [1, 0]])
Y = _m([[0, -1j], [1j, 0]])
Z = _m([[1, 0], [0, -1]])
S = _m([[1, 0], [0, 1j]])
T = _m([[1, 0], [0, cmath.exp(1j * cmath.pi / 4)]])
Sdag = S.conj().T.copy()
Tdag = T.conj().T.copy()
SqrtX = 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]], dtype=np.complex128)
Swap = _m([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])
SqrtSwap = _m([[1, 0, 0, 0], [0, 0.5 + 0.5j, 0.5 - 0.5j, 0], [0, 0.5 - 0.5j, 0.5 + 0.5j, 0], [0, 0, 0, 1]])


def Rx(angle):  # noqa: N802 -- reference names; projectq/ops/_gates.py:273-284
    a = rotation_angle(angle)
    return _m(
        [
            [math.cos(0.5 * a), -1j * math.sin(0.5 * a)],
            [-1j * math.sin(0.5 * a), math.cos(0.5 * a)],
        ]
    )


def Ry(angle):  # noqa: N802 -- projectq/ops/_gates.py:287-298
    a = rotation_angle(angle)
    return _m(
        [
            [math.cos(0.5 * a), -math.sin(0.5 * a)],
            [math.sin(0.5 * a), math.cos(0.5 * a)],
        ]
    )


def Rz(angle):  # noqa: N802 -- projectq/ops/_gates.py:301-312
    a = rotation_angle(angle)
    return _m([[cmath.exp(-0.5 * 1j * a), 0], [0, cmath.exp(0.5 * 1j * a)]])


def Rxx(angle):  # noqa: N802 -- projectq/ops/_gates.py:315-328
    a = rotation_angle(angle)
    c, s = cmath.cos(0.5 * a), cmath.sin(0.5 * a)
    return _m([[c, 0, 0, -1j * s], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [-1j * s, 0, 0, c]])


def Ryy(angle):  # noqa: N802 -- projectq/ops/_gates.py:331-344
    a = rotation_angle(angle)
    c, s = cmath.cos(0.5 * a), cmath.sin(0.5 * a)
    return _m([[c, 0, 0, 1j * s], [0, c, -1j * s, 0], [0, -1j * s, c, 0], [1j * s, 0, 0, c]])


def Rzz(angle):  # noqa: N802 -- projectq/ops/_gates.py:347-357
    a = rotation_angle(angle)
    m, p = cmath.exp(-0.5 * 1j * a), cmath.exp(0.5 * 1j * a)
    return _m([[m, 0, 0, 0], [0, p, 0, 0], [0, 0, p, 0], [0, 0, 0, m]])


def Ph(angle):  # noqa: N802 -- projectq/ops/_gates.py:263-270
    a = phase_angle(angle)
    return _m([[cmath.exp(1j * a), 0], [0, cmath.exp(1j * a)]])


def R(angle):  # noqa: N802 -- projectq/ops/_gates.py:360-366
    a = phase_angle(angle)
    return _m([[1, 0], [0, cmath.exp(1j * a)]])


def dagger(matrix):
    """DaggeredGate.matrix: conjugate transpose."""
    return np.ascontiguousarray(np.asarray(matrix).conj().T)


ROTATIONS = {"Rx": Rx, "Ry": Ry, "Rz": Rz}


# --- math functions (projectq/libs/math/_gates.py), tagged for native emulation -----------------
def _tag(fn, opcode, params=()):
    fn._qs_math = (opcode, list(params))
    return fn


def add_constant(a):
    """AddConstant(a) math function: x -> x + a (_gates.py:20-62)."""
    return _tag(lambda x: [x[0] + a], _mathfuncs.ADD_CONSTANT, [a])


def add_quantum(n_qubits):
    """AddQuantum math function for registers of n_qubits (_gates.py:251-265)."""

    def math_fun(a):
        a = list(a)
        a[1] = a[0] + a[1]
        if len(bin(a[1])[2:]) > n_qubits:
            a[1] = a[1] % (2**n_qubits)
            if len(a) == 3:
                a[2] ^= 1
        return a

    return _tag(math_fun, _mathfuncs.ADD_QUANTUM)


def subtract_quantum():
    """SubtractQuantum math function (_gates.py:322-325)."""
    return _tag(lambda v: [v[0], v[1] - v[0]], _mathfuncs.SUBTRACT_QUANTUM)


def comparator_quantum():
    """ComparatorQuantum math function (_gates.py:372-379)."""

    def compare(v):
        a, b, c = v
        if b < a:
            c = 1 if c == 0 else 0
        return [a, b, c]

    return _tag(compare, _mathfuncs.COMPARATOR)


def multiply_quantum():
    """MultiplyQuantum math function (_gates.py:516-517)."""
    return _tag(lambda v: [v[0], v[1], v[2] + v[0] * v[1]], _mathfuncs.MULTIPLY_QUANTUM)


def divide_quantum():
    """DivideQuantum math function (_gates.py:438-442)."""

    def division(v):
        dividend, remainder, divisor = v
        if divisor == 0 or divisor > dividend:
            return [remainder, dividend, divisor]
        quotient = remainder + dividend // divisor
        return [dividend - quotient * divisor, quotient, divisor]

    return _tag(division, _mathfuncs.DIVIDE_QUANTUM)
