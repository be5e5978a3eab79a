"""ProjectQ's default compiler engine list in front of a `_cppsim.Simulator`-surface backend.

A ProjectQ program run with `MainEngine()` does not reach the simulator gate by gate: the default
engine list (projectq/setups/default.py:31-34)

    TagRemover -> LocalOptimizer(10) -> AutoReplacer(default rules) -> TagRemover -> LocalOptimizer(10)

merges rotations, cancels gates against their inverses, drops identities, expands the QFT and
negatively controlled gates, and only then hands commands to the `Simulator` backend, whose
`_handle` turns each one into calls on the `_cppsim` object (projectq/backends/_sim/_simulator.py
:86-116, 324-438).  This module restates that pipeline on a symbolic command model so a program
written against it produces the same backend-call sequence as the reference -- the same gates in
the same order with bit-identical matrices, the same `run()` calls and measurements -- in front of
`projectq_amd.Simulator` or any object with that method surface.  The call sequences are pinned by
fixtures recorded from the reference's own MainEngine.

Restated pieces and the reference code each follows:

* gates and their algebra -- equality, inverse, merge, identity test, angle canonicalisation
  (projectq/ops/_basics.py:66-504, _gates.py:62-486, _metagates.py:38-133, _qftgate.py,
  libs/math/_gates.py:20-400);
* `Command` -- qubit ordering of interchangeable registers, sorted controls with their control
  state, equality, inverse and merge (projectq/ops/_command.py:92-340);
* `TagRemover` (cengines/_tagremover.py), `LocalOptimizer` (cengines/_optimize.py:116-255),
  `AutoReplacer` with the rules this gate set needs: QFT -> H + controlled R
  (setups/decompositions/qft2crandhadamard.py:31-39), its inverse through Dagger
  (cengines/_replacer/_decomposition_rule_set.py get_inverse_decomposition), negative controls
  (setups/decompositions/controlstate.py) and barriers (setups/decompositions/barrier.py);
* the meta sections Control, Compute/Uncompute (without qubit allocation inside) and Dagger
  (projectq/meta/_control.py, _compute.py, _dagger.py);
* `TimeEvolution` and single-term `QubitOperator` gates with the rewriting of their `|` operators
  (projectq/ops/_time_evolution.py, _qubit_operator.py:224-343), the Entangle and QubitOperator
  rules (setups/decompositions/entangle.py, qubitop2onequbit.py);
* `SimulatorEngine` -- the backend's `is_available`, `_handle` and `receive`.

The optimizer's list operations are kept step for step, including two behaviours of the reference
that a user can observe: an identity command on several qubits is removed from the pipeline of its
last qubit only, and the resulting broken pipeline is reported by printing "Invalid qubit pipeline
encountered" rather than raising.  Programs whose reference run raises reproduce the same
exception type.
"""

import cmath
import math
from contextlib import contextmanager

import numpy as np

from . import _mathfuncs
from . import gates as G


class NotMergeable(Exception):
    pass


class NotInvertible(Exception):
    pass


class NoGateDecompositionError(Exception):
    pass


class IncompatibleControlState(Exception):
    pass


class QubitManagementError(Exception):
    pass


# ============================================================================================
# gates
# ============================================================================================
class BasicGate:
    """Gate algebra of projectq/ops/_basics.py:66-235: equal by class, no inverse, no merge."""

    interchangeable = ()  # groups of register indices whose order does not matter
    name = "Gate"

    @property
    def matrix(self):
        return None

    def get_inverse(self):
        raise NotInvertible(f"{self.name}: no inverse")

    def get_merged(self, other):
        raise NotMergeable(f"{self.name}: no merge")

    def is_identity(self):
        return False

    def __eq__(self, other):
        if isinstance(other, self.__class__):
            return True
        if isinstance(other, MatrixGate):
            return NotImplemented
        return False

    def __ne__(self, other):
        eq = self.__eq__(other)
        return eq if eq is NotImplemented else not eq

    __hash__ = None

    def __repr__(self):
        return self.name


def get_inverse(gate):
    """_metagates.py:98-115: the gate's inverse, or a DaggeredGate when it has none."""
    try:
        return gate.get_inverse()
    except NotInvertible:
        return DaggeredGate(gate)


class SelfInverseGate(BasicGate):
    def get_inverse(self):
        return self.__class__()


class _Fixed:
    """Mixin: a gate with a constant matrix from projectq_amd.gates."""

    _m = None

    @property
    def matrix(self):
        return self._m


class HGate(_Fixed, SelfInverseGate):
    name, _m = "H", G.H


class XGate(_Fixed, SelfInverseGate):
    name, _m = "X", G.X


class YGate(_Fixed, SelfInverseGate):
    name, _m = "Y", G.Y


class ZGate(_Fixed, SelfInverseGate):
    name, _m = "Z", G.Z


class SwapGate(_Fixed, SelfInverseGate):
    name, _m, interchangeable = "Swap", G.Swap, ((0, 1),)


class SGate(_Fixed, BasicGate):
    name, _m = "S", G.S


class TGate(_Fixed, BasicGate):
    name, _m = "T", G.T


class SqrtXGate(_Fixed, BasicGate):
    name, _m = "SqrtX", G.SqrtX


class SqrtSwapGate(_Fixed, BasicGate):
    name, _m, interchangeable = "SqrtSwap", G.SqrtSwap, ((0, 1),)


class QFTGate(BasicGate):
    name = "QFT"


class BarrierGate(BasicGate):
    name = "Barrier"

    def get_inverse(self):
        return BarrierGate()


class DaggeredGate(BasicGate):
    """_metagates.py:38-95: inverse of a gate without its own inverse; matrix = conjugate transpose."""

    def __init__(self, gate):
        self.gate = gate
        self.name = f"{gate.name}^dagger"

    @property
    def matrix(self):
        m = self.gate.matrix
        return None if m is None else G.dagger(m)

    def get_inverse(self):
        return self.gate

    def __eq__(self, other):
        return isinstance(other, DaggeredGate) and self.gate == other.gate


class RotationGate(BasicGate):
    """BasicRotationGate (_basics.py:321-419): angle kept as round(angle mod 4 pi, 12)."""

    period = 4 * math.pi
    _canon = staticmethod(G.rotation_angle)

    def __init__(self, angle):
        self.angle = self._canon(angle)

    @property
    def name(self):
        return f"{self.__class__.__name__}({self.angle!r})"

    def get_inverse(self):
        if self.angle == 0:
            return self.__class__(0)
        return self.__class__(-self.angle + self.period)

    def get_merged(self, other):
        if isinstance(other, self.__class__):
            return self.__class__(self.angle + other.angle)
        raise NotMergeable("different rotation classes")

    def __eq__(self, other):
        if isinstance(other, self.__class__):
            return self.angle == other.angle
        return False

    def is_identity(self):
        return self.angle in (0.0, 4 * math.pi)

    @property
    def matrix(self):
        return self._matrix_fn(self.angle)


class PhaseGate(RotationGate):
    """BasicPhaseGate (_basics.py:421-502): mod 2 pi, never reported as an identity."""

    period = 2 * math.pi
    _canon = staticmethod(G.phase_angle)

    def is_identity(self):
        return False


class Rx(RotationGate):
    _matrix_fn = staticmethod(G.Rx)


class Ry(RotationGate):
    _matrix_fn = staticmethod(G.Ry)


class Rz(RotationGate):
    _matrix_fn = staticmethod(G.Rz)


class Rxx(RotationGate):
    _matrix_fn = staticmethod(G.Rxx)


class Ryy(RotationGate):
    _matrix_fn = staticmethod(G.Ryy)


class Rzz(RotationGate):
    _matrix_fn = staticmethod(G.Rzz)


class Ph(PhaseGate):
    _matrix_fn = staticmethod(G.Ph)


class R(PhaseGate):
    _matrix_fn = staticmethod(G.R)


class MatrixGate(BasicGate):
    """_basics.py:238-300: equal to any gate whose matrix is allclose (rtol 1e-10, atol 1e-12)."""

    name = "MatrixGate"

    def __init__(self, matrix):
        self._matrix = np.array(matrix, dtype=np.complex128)

    @property
    def matrix(self):
        return self._matrix

    def __eq__(self, other):
        m = getattr(other, "matrix", None)
        if m is None:
            return False
        return self._matrix.shape == m.shape and bool(np.allclose(self._matrix, m, rtol=1e-10, atol=1e-12))

    def get_inverse(self):
        return MatrixGate(np.linalg.inv(self._matrix))


class ClassicalInstructionGate(BasicGate):
    pass


class FastForwardingGate(ClassicalInstructionGate):
    pass


class MeasureGate(FastForwardingGate):
    name = "Measure"


class AllocateQubitGate(ClassicalInstructionGate):
    name = "Allocate"

    def get_inverse(self):
        return DeallocateQubitGate()


class DeallocateQubitGate(FastForwardingGate):
    name = "Deallocate"

    def get_inverse(self):
        return AllocateQubitGate()


class FlushGate(FastForwardingGate):
    name = "Flush"


# --- math gates (projectq/libs/math/_gates.py) ---------------------------------------------
class BasicMathGate(BasicGate):
    """A classical reversible function of register values, emulated by the backend."""

    def math_function(self, quregs):
        raise NotImplementedError


class AddConstant(BasicMathGate):
    def __init__(self, a):
        self.a = a
        self.name = f"AddConstant({a})"

    def get_inverse(self):
        return SubConstant(self.a)

    def __eq__(self, other):
        return isinstance(other, AddConstant) and self.a == other.a

    def math_function(self, quregs):
        return G.add_constant(self.a)


def SubConstant(a):  # noqa: N802 -- reference name (_gates.py:66-79)
    return AddConstant(-a)


class AddConstantModN(BasicMathGate):
    def __init__(self, a, N):  # noqa: N803
        self.a, self.N = a, N
        self.name = f"AddConstantModN({a}, {N})"

    def get_inverse(self):
        return SubConstantModN(self.a, self.N)

    def __eq__(self, other):
        return isinstance(other, AddConstantModN) and self.a == other.a and self.N == other.N


def SubConstantModN(a, N):  # noqa: N802,N803 -- reference name (_gates.py:143-161)
    return AddConstantModN(N - a, N)


class MultiplyByConstantModN(BasicMathGate):
    def __init__(self, a, N):  # noqa: N803
        self.a, self.N = a, N
        self.name = f"MultiplyByConstantModN({a}, {N})"

    def __eq__(self, other):
        return isinstance(other, MultiplyByConstantModN) and self.a == other.a and self.N == other.N


class AddQuantumGate(BasicMathGate):
    name = "AddQuantum"

    def get_inverse(self):
        return InverseAddQuantumGate()

    def math_function(self, quregs):
        return G.add_quantum(len(quregs[0]))


class InverseAddQuantumGate(BasicMathGate):
    name = "_InverseAddQuantum"

    def math_function(self, quregs):
        def math_fun(v):  # _gates.py:286-295
            if len(v) == 3:
                v[2] ^= 1
            v[1] -= v[0]
            return v

        return G._tag(math_fun, _mathfuncs.INVERSE_ADD_QUANTUM)


class SubtractQuantumGate(BasicMathGate):
    name = "SubtractQuantum"

    def get_inverse(self):
        return AddQuantumGate()

    def math_function(self, quregs):
        return G.subtract_quantum()


class ComparatorQuantumGate(BasicMathGate):
    name = "Comparator"

    def get_inverse(self):
        # the reference returns AddQuantum here (_gates.py:395-397); kept so cancellations agree
        return AddQuantumGate()

    def math_function(self, quregs):
        return G.comparator_quantum()


class MultiplyQuantumGate(BasicMathGate):
    name = "MultiplyQuantum"

    def get_inverse(self):
        return InverseMultiplyQuantumGate()

    def math_function(self, quregs):
        return G.multiply_quantum()


class InverseMultiplyQuantumGate(BasicMathGate):
    name = "_InverseMultiplyQuantum"

    def math_function(self, quregs):  # _gates.py:541-551
        return G._tag(lambda v: [v[0], v[1], v[2] - v[0] * v[1]], _mathfuncs.INVERSE_MULTIPLY_QUANTUM)


class DivideQuantumGate(BasicMathGate):
    name = "DivideQuantum"

    def get_inverse(self):
        return InverseDivideQuantumGate()

    def math_function(self, quregs):
        return G.divide_quantum()


class InverseDivideQuantumGate(BasicMathGate):
    name = "_InverseDivideQuantum"

    def math_function(self, quregs):  # _gates.py:467-486
        def inverse_division(v):
            remainder, quotient, divisor = v
            if divisor == 0:
                return [quotient, remainder, divisor]
            return [remainder + quotient * divisor, 0, divisor]

        return G._tag(inverse_division, _mathfuncs.INVERSE_DIVIDE_QUANTUM)


class EntangleGate(BasicGate):
    """H on the first qubit, then X on every other one controlled by it (_gates.py:245-255)."""

    name = "Entangle"


class QubitOperatorGate(BasicGate):
    """A single Pauli string with a unit coefficient used as a gate (projectq/ops/_qubit_operator.py
    __or__, get_inverse).  Like the reference's QubitOperator it takes BasicGate's equality: any
    two compare equal, so the optimizer cancels two adjacent ones on the same qubits whatever
    their Pauli strings."""

    name = "QubitOperator"

    def __init__(self, term, coefficient):
        self.term = tuple(sorted(((int(q), str(p)) for q, p in term), key=lambda lp: lp[0]))
        self.coefficient = complex(coefficient)

    def get_inverse(self):
        if not abs(self.coefficient) < 1 - 1e-12 and not abs(self.coefficient) > 1 + 1e-12:
            return QubitOperatorGate(self.term, self.coefficient ** (-1))
        raise NotInvertible("BasicGate: No get_inverse() implemented.")

    def commands_for(self, qureg):
        """What `QubitOperator(...) | qureg` emits: Ph for the identity, the Pauli and Ph for one
        qubit, else the gate on the qubits it acts on, renumbered."""
        c = self.coefficient
        if abs(c) < 1 - 1e-12 or abs(c) > 1 + 1e-12:
            raise TypeError("abs(coefficient) != 1.")
        phase = cmath.phase(c)
        if self.term == ():
            return [(Ph(phase), [qureg[0]])]
        used = sorted({q for q, _ in self.term})
        if used[-1] >= len(qureg):
            raise ValueError("QubitOperator acts on more qubits than the gate is applied to.")
        if len(self.term) == 1:
            q, p = self.term[0]
            return [({"X": XGate, "Y": YGate, "Z": ZGate}[p](), [qureg[q]]), (Ph(phase), [qureg[q]])]
        pos = {q: i for i, q in enumerate(used)}
        return [(QubitOperatorGate(tuple((pos[q], p) for q, p in self.term), c), [qureg[q] for q in used])]


class NotHermitianOperatorError(ValueError):
    pass


class TimeEvolution(BasicGate):
    """exp(-i t H) for a Pauli-sum H with real coefficients (projectq/ops/_time_evolution.py).

    `hamiltonian` is a dict or a list of (term, coefficient) pairs, a term being a tuple of
    (qubit index, 'X' | 'Y' | 'Z') sorted by index; the order of the pairs is the order the
    backend receives the terms in.  Two TimeEvolution gates compare equal only when they are the
    same object (the reference's __eq__ returns NotImplemented), so the optimizer never merges or
    cancels them."""

    name = "TimeEvolution"

    def __init__(self, time, hamiltonian):
        if not isinstance(time, (float, int)):
            raise TypeError("time needs to be a (real) numeric type.")
        items = hamiltonian.items() if isinstance(hamiltonian, dict) else hamiltonian
        self.time = time
        self.terms = {}
        for term, coeff in items:
            term = tuple(sorted(((int(q), str(p)) for q, p in term), key=lambda lp: lp[0]))
            coeff = complex(coeff)
            if coeff.imag != 0:
                raise NotHermitianOperatorError("hamiltonian must be hermitian and hence only have real coefficients.")
            self.terms[term] = float(coeff.real)

    def get_inverse(self):
        return TimeEvolution(self.time * -1.0, self.terms)

    def get_merged(self, other):
        if isinstance(other, TimeEvolution) and set(self.terms) == set(other.terms):
            factor = None
            for term in self.terms:
                ratio = self.terms[term] / float(other.terms[term])
                if factor is None:
                    factor = ratio
                elif not abs(factor - ratio) <= 1e-9 * max(abs(factor), abs(ratio)):
                    raise NotMergeable("Cannot merge these two gates.")
            return TimeEvolution(self.time + other.time / factor, self.terms)
        raise NotMergeable("Cannot merge these two gates.")

    def __eq__(self, other):
        return NotImplemented

    def commands_for(self, qureg):
        """What `TimeEvolution(...) | qureg` emits (_time_evolution.py __or__): a phase gate on
        the first qubit for an identity-only H, else the gate on the qubits H acts on, with
        the term indices renumbered to positions in that list."""
        if len(self.terms) == 1 and () in self.terms:
            return [(Ph(-1 * self.time * self.terms[()]), [qureg[0]])]
        used = sorted({q for term in self.terms for q, _ in term})
        if used[-1] >= len(qureg):
            raise ValueError("hamiltonian acts on more qubits than the gate is applied to.")
        pos = {q: i for i, q in enumerate(used)}
        terms = [(tuple((pos[q], p) for q, p in term), c) for term, c in self.terms.items()]
        return [(TimeEvolution(self.time, terms), [qureg[q] for q in used])]


# shortcuts, as projectq.ops / projectq.libs.math export them
H, X, Y, Z, S, T, SqrtX, Swap, SqrtSwap = HGate(), XGate(), YGate(), ZGate(), SGate(), TGate(), SqrtXGate(), SwapGate(), SqrtSwapGate()
Sdag, Tdag = get_inverse(S), get_inverse(T)
QFT, Barrier, Measure = QFTGate(), BarrierGate(), MeasureGate()
Allocate, Deallocate = AllocateQubitGate(), DeallocateQubitGate()
AddQuantum, SubtractQuantum, ComparatorQuantum = AddQuantumGate(), SubtractQuantumGate(), ComparatorQuantumGate()


# ============================================================================================
# commands
# ============================================================================================
COMPUTE_TAG, UNCOMPUTE_TAG = "ComputeTag", "UncomputeTag"


def canonical_ctrl_state(state, n):
    """projectq/meta/_control.py canonical_ctrl_state: None = all ones; ints are LSB-first."""
    if not n:
        return ""
    if state is None:
        return "1" * n
    if isinstance(state, int):
        s = f"{state:b}".zfill(n)[::-1]
        if len(s) != n:
            raise ValueError(f"control state {state} needs more than {n} qubits")
        return s
    if isinstance(state, str):
        if len(state) != n:
            raise ValueError(f"control state {state} has a different length than the {n} control qubits")
        if not set(state) <= {"0", "1"}:
            raise ValueError(f"control state {state} has characters other than 0 and 1")
        return state
    raise TypeError("control state must be None, an int or a string")


class Command:
    """A gate on registers of qubit ids, with sorted control ids, control state and tags
    (projectq/ops/_command.py:92-340).  Inverse, merge and copy keep the controls but reset the
    control state to all ones, as the reference's constructors do."""

    def __init__(self, gate, qubits, controls=(), tags=(), control_state=None):
        self.gate = gate
        self.tags = list(tags)
        self.qubits = self._order(tuple([int(q) for q in reg] for reg in qubits))
        self.control_qubits = sorted(int(c) for c in controls)
        self.control_state = canonical_ctrl_state(control_state, len(self.control_qubits))

    def _order(self, qubits):
        ordered = list(qubits)
        for group in self.gate.interchangeable:
            by_first = sorted(group, key=lambda i: ordered[i][0])
            regs = [ordered[i] for i in by_first]
            for pos, reg in zip(group, regs):
                ordered[pos] = reg
        return tuple(ordered)

    @property
    def all_qubits(self):
        return (self.control_qubits,) + self.qubits

    def copy(self):
        return Command(self.gate, self.qubits, list(self.control_qubits), list(self.tags))

    def get_inverse(self):
        return Command(get_inverse(self.gate), self.qubits, list(self.control_qubits), list(self.tags))

    def is_identity(self):
        return self.gate.is_identity()

    def get_merged(self, other):
        if self.tags == other.tags and self.all_qubits == other.all_qubits:
            return Command(self.gate.get_merged(other.gate), self.qubits, self.control_qubits, list(self.tags))
        raise NotMergeable("commands not mergeable")

    def add_control_qubits(self, qubits, state=None):
        qubits = [int(q) for q in qubits]
        ids = self.control_qubits + qubits
        states = self.control_state + canonical_ctrl_state(state, len(qubits))
        pairs = sorted(zip(ids, states), key=lambda p: p[0])
        by_id = {}
        for q, s in pairs:
            if by_id.setdefault(q, s) != s:
                raise IncompatibleControlState(f"control qubit {q} has conflicting control states")
        self.control_qubits = [q for q, _ in pairs]
        self.control_state = "".join(s for _, s in pairs)

    def __eq__(self, other):
        return (
            isinstance(other, Command)
            and self.gate == other.gate
            and self.tags == other.tags
            and self.all_qubits == other.all_qubits
        )

    def __ne__(self, other):
        return not self.__eq__(other)

    __hash__ = None

    def __repr__(self):
        ctrl = f" ctrl={self.control_qubits}/{self.control_state}" if self.control_qubits else ""
        return f"{self.gate!r} | {self.qubits}{ctrl}"


def has_negative_control(cmd):
    return len(cmd.control_qubits) > 0 and "0" in cmd.control_state


# ============================================================================================
# engines
# ============================================================================================
class Engine:
    """A compiler stage: receives command lists and sends them to `next_engine`."""

    next_engine = None

    def send(self, cmds):
        self.next_engine.receive(cmds)

    def is_available(self, cmd):
        return self.next_engine.is_available(cmd)

    def receive(self, cmds):  # pragma: no cover -- every stage overrides it
        raise NotImplementedError


class TagRemover(Engine):
    """cengines/_tagremover.py: drops Compute/Uncompute tags so optimisation can cross them."""

    def __init__(self, tags=(COMPUTE_TAG, UNCOMPUTE_TAG)):
        self._tags = tuple(tags)

    def receive(self, cmds):
        for cmd in cmds:
            cmd.tags = [t for t in cmd.tags if t not in self._tags]
            self.send([cmd])


class LocalOptimizer(Engine):
    """cengines/_optimize.py:116-255: a per-qubit pipeline of at most `cache_size` commands in
    which identities are removed, inverse pairs cancel and mergeable neighbours merge."""

    def __init__(self, cache_size=5):
        self._l = {}
        self._cache_size = cache_size

    def _send_qubit_pipeline(self, idx, n_gates):
        pipeline = self._l[idx]
        for i in range(min(n_gates, len(pipeline))):
            cmd = pipeline[i]
            others = [q for reg in cmd.all_qubits for q in reg if q != idx]
            for q in others:
                try:
                    loc = 0
                    while self._l[q][loc] != cmd:
                        loc += 1
                    loc = self._optimize(q, loc)
                    self._send_qubit_pipeline(q, loc)
                    self._l[q] = self._l[q][1:]
                except IndexError:
                    print("Invalid qubit pipeline encountered (in the process of shutting down?).")
            self.send([cmd])
        self._l[idx] = self._l[idx][n_gates:]

    def _gate_indices(self, idx, i, qubit_ids):
        """Position of command `i` of qubit `idx` in the pipeline of every qubit it involves."""
        if len(qubit_ids) == 1:
            return [i]
        cmd = self._l[idx][i]
        skip = sum(1 for prev in self._l[idx][:i] if prev == cmd)
        out = []
        for q in qubit_ids:
            matches = [k for k, c in enumerate(self._l[q]) if c == cmd]
            out.append(matches[skip])
        return out

    def _optimize(self, idx, lim=None):
        i = 0
        limit = len(self._l[idx]) if lim is None else lim
        while i < limit - 1:
            cmd = self._l[idx][i]
            ids = [q for reg in cmd.all_qubits for q in reg]
            if cmd.is_identity():
                gid = self._gate_indices(idx, i, ids)
                new_list = None
                for j, q in enumerate(ids):
                    new_list = self._l[q][0 : gid[j]] + self._l[q][gid[j] + 1 :]
                # only the last qubit's pipeline is replaced (_optimize.py:150-156)
                self._l[ids[-1]] = new_list
                i = 0
                limit -= 1
                continue

            inv = cmd.get_inverse()
            if inv == self._l[idx][i + 1]:
                gid = self._gate_indices(idx, i, ids)
                erase = True
                for j, q in enumerate(ids):  # every pipeline is read, as `erase *= ...` does
                    same = inv == self._l[q][gid[j] + 1]
                    erase = erase and same
                if erase:
                    for j, q in enumerate(ids):
                        self._l[q] = self._l[q][0 : gid[j]] + self._l[q][gid[j] + 2 :]
                    i = 0
                    limit -= 2
                    continue

            try:
                merged_cmd = cmd.get_merged(self._l[idx][i + 1])
                gid = self._gate_indices(idx, i, ids)
                merge = True
                for j, q in enumerate(ids):
                    same = self._l[q][gid[j]].get_merged(self._l[q][gid[j] + 1]) == merged_cmd
                    merge = merge and same
                if merge:
                    for j, q in enumerate(ids):
                        self._l[q][gid[j]] = merged_cmd
                        self._l[q] = self._l[q][0 : gid[j] + 1] + self._l[q][gid[j] + 2 :]
                    i = 0
                    limit -= 1
                    continue
            except NotMergeable:
                pass
            i += 1
        return limit

    def _check_and_send(self):
        cache = self._cache_size
        for idx in self._l:
            pipe = self._l[idx]
            fast = len(pipe) > 0 and isinstance(pipe[-1].gate, FastForwardingGate)
            if len(pipe) >= cache or fast:
                self._optimize(idx)
                pipe = self._l[idx]
                if len(pipe) >= cache and not isinstance(pipe[-1].gate, FastForwardingGate):
                    self._send_qubit_pipeline(idx, len(pipe) - cache + 1)
                elif len(pipe) > 0 and isinstance(pipe[-1].gate, FastForwardingGate):
                    self._send_qubit_pipeline(idx, len(pipe))
        self._l = {q: p for q, p in self._l.items() if len(p) > 0}

    def _cache_cmd(self, cmd):
        for q in [q for reg in cmd.all_qubits for q in reg]:
            self._l.setdefault(q, [])
            self._l[q] += [cmd]
        self._check_and_send()

    def receive(self, cmds):
        for cmd in cmds:
            if isinstance(cmd.gate, FlushGate):
                for idx in self._l:
                    self._optimize(idx)
                    self._send_qubit_pipeline(idx, len(self._l[idx]))
                self._l = {q: p for q, p in self._l.items() if len(p) > 0}
                if self._l:
                    raise RuntimeError("Internal compiler error: qubits remaining in LocalOptimizer after a flush!")
                self.send([cmd])
            else:
                self._cache_cmd(cmd)


class AutoReplacer(Engine):
    """cengines/_replacer/_replacer.py: commands the backend cannot take are decomposed, and the
    decomposition's commands (carrying the original's tags first) come back through this stage."""

    def receive(self, cmds):
        for cmd in cmds:
            if isinstance(cmd.gate, FlushGate):
                self.send([cmd])
            else:
                self._process(cmd)

    def _process(self, cmd):
        if self.is_available(cmd):
            self.send([cmd])
            return
        if has_negative_control(cmd):
            rule = _decompose_controlstate
        elif isinstance(cmd.gate, QFTGate):
            rule = _decompose_qft
        elif isinstance(cmd.gate, DaggeredGate) and isinstance(cmd.gate.gate, QFTGate):
            rule = _inverse_rule(_decompose_qft)
        elif isinstance(cmd.gate, BarrierGate):
            rule = _decompose_barrier
        elif isinstance(cmd.gate, EntangleGate):
            rule = _decompose_entangle
        elif isinstance(cmd.gate, QubitOperatorGate) and len(cmd.control_qubits) <= 1:
            rule = _decompose_qubitop
        else:
            raise NoGateDecompositionError(f"No replacement found for {cmd!r}!")
        old_tags = list(cmd.tags)

        def retag(cmds):
            for c in cmds:
                c.tags = old_tags + c.tags
                self.receive([c])

        rule(cmd, _Frontend(retag))


# --- decomposition rules --------------------------------------------------------------------
def _decompose_qft(cmd, fe):
    """qft2crandhadamard.py:31-39."""
    qb = cmd.qubits[0]
    with fe.control(cmd.control_qubits):
        for i in range(len(qb)):
            fe.apply(HGate(), [qb[-1 - i]])
            for j in range(len(qb) - 1 - i):
                with fe.control([qb[-1 - (j + i + 1)]]):
                    fe.apply(R(math.pi / (1 << (1 + j))), [qb[-1 - i]])


def _inverse_rule(rule):
    """The inverse of a rule: the rule on the command's inverse inside a Dagger section."""

    def decompose(cmd, fe):
        with fe.dagger():
            rule(cmd.get_inverse(), fe)

    return decompose


def _decompose_controlstate(cmd, fe):
    """controlstate.py: X on the 0-controls around the command with all-ones controls."""
    fe.begin_compute()
    for state, ctrl in zip(cmd.control_state, cmd.control_qubits):
        if state == "0":
            fe.apply(XGate(), [ctrl])
    fe.end_compute()
    fe.receive([cmd.copy()])
    fe.uncompute()


def _decompose_barrier(cmd, fe):
    """barrier.py: a barrier the backend does not take is dropped."""


def _decompose_entangle(cmd, fe):
    """entangle.py: H on the first qubit, then X on each other qubit controlled by it."""
    qureg = cmd.qubits[0]
    with fe.control(cmd.control_qubits):
        fe.apply(HGate(), [qureg[0]])
        with fe.control([qureg[0]]):
            for q in qureg[1:]:
                fe.apply(XGate(), [q])


def _decompose_qubitop(cmd, fe):
    """qubitop2onequbit.py: Ph(phase of the coefficient) on the first qubit, then the Paulis."""
    qureg = cmd.qubits[0]
    gate = cmd.gate
    with fe.control(cmd.control_qubits):
        fe.apply(Ph(cmath.phase(gate.coefficient)), [qureg[0]])
        for index, action in gate.term:
            fe.apply({"X": XGate, "Y": YGate, "Z": ZGate}[action](), [qureg[index]])


# --- meta sections --------------------------------------------------------------------------
class _ControlStage:
    def __init__(self, qubits, state):
        self.qubits, self.state = list(qubits), state
        self.next = None

    def receive(self, cmds):
        for cmd in cmds:
            tagged = COMPUTE_TAG in cmd.tags or UNCOMPUTE_TAG in cmd.tags
            if not tagged and not isinstance(cmd.gate, ClassicalInstructionGate):
                cmd.add_control_qubits(self.qubits, self.state)
            self.next.receive([cmd])


class _ComputeStage:
    def __init__(self):
        self.stored = []
        self.computing = True
        self.next = None

    def receive(self, cmds):
        if self.computing:
            for cmd in cmds:
                if isinstance(cmd.gate, (AllocateQubitGate, DeallocateQubitGate)):
                    raise QubitManagementError("qubit allocation inside a Compute section is not restated here")
                self.stored.append(cmd.copy())
                cmd.tags.append(COMPUTE_TAG)
        self.next.receive(cmds)

    def run_uncompute(self):
        out = []
        for cmd in reversed(self.stored):
            inv = cmd.get_inverse()
            inv.tags.append(UNCOMPUTE_TAG)
            out.append(inv)
        self.next.receive(out)


class _DaggerStage:
    def __init__(self):
        self.stored = []
        self.next = None

    def receive(self, cmds):
        for cmd in cmds:
            if isinstance(cmd.gate, (AllocateQubitGate, DeallocateQubitGate)):
                raise QubitManagementError("qubit allocation inside a Dagger section is not restated here")
        self.stored.extend(cmds)

    def run(self):
        for cmd in reversed(self.stored):
            self.next.receive([cmd.get_inverse()])


class _Sink:
    def __init__(self, fn):
        self.receive = fn


class _Frontend:
    """A command source with a stack of meta stages in front of a sink (the MainEngine of a
    program, or the forwarder a decomposition rule writes to)."""

    def __init__(self, sink):
        self._head = _Sink(sink) if callable(sink) else sink

    def receive(self, cmds):
        self._head.receive(cmds)

    def _push(self, stage):
        stage.next = self._head
        self._head = stage
        return stage

    def _pop(self):
        stage = self._head
        self._head = stage.next
        return stage

    def apply(self, gate, *regs, controls=(), control_state=None):
        if isinstance(gate, (TimeEvolution, QubitOperatorGate)):
            # these gates' | operator rewrites them first (onto the qubits they act on)
            if len(regs) != 1:
                raise TypeError("Only one qubit or qureg allowed.")
            for g, reg in gate.commands_for(list(regs[0])):
                self._apply_one(g, [reg], controls, control_state)
            return
        self._apply_one(gate, regs, controls, control_state)

    def _apply_one(self, gate, regs, controls, control_state):
        cmd = Command(gate, [list(r) for r in regs])
        if len(controls):
            cmd.add_control_qubits(list(controls), control_state)
        self.receive([cmd])

    @contextmanager
    def control(self, qubits, state=None):
        qubits = list(qubits)
        if not qubits:
            yield
            return
        self._push(_ControlStage(qubits, state))
        try:
            yield
        finally:
            self._pop()

    def begin_compute(self):
        self._push(_ComputeStage())

    def end_compute(self):
        stage = self._head
        if not isinstance(stage, _ComputeStage):
            raise QubitManagementError("end of a Compute section without its start")
        stage.computing = False

    def uncompute(self):
        stage = self._head
        if not isinstance(stage, _ComputeStage):
            raise QubitManagementError("Uncompute without a preceding Compute section")
        stage.run_uncompute()
        self._pop()

    @contextmanager
    def compute(self):
        self.begin_compute()
        try:
            yield
        finally:
            self.end_compute()

    @contextmanager
    def dagger(self):
        stage = self._push(_DaggerStage())
        try:
            yield
        finally:
            if self._head is stage:
                self._pop()
        stage.run()


# ============================================================================================
# backend and main engine
# ============================================================================================
class SimulatorEngine(Engine):
    """The `Simulator` backend engine (projectq/backends/_sim/_simulator.py:86-116, 324-438)
    over an object with the `_cppsim.Simulator` method surface."""

    def __init__(self, simulator, gate_fusion=False):
        self.sim = simulator
        self.gate_fusion = gate_fusion
        self.main = None

    def is_available(self, cmd):
        if has_negative_control(cmd):
            return False
        if isinstance(cmd.gate, (MeasureGate, AllocateQubitGate, DeallocateQubitGate, BasicMathGate, TimeEvolution)):
            return True
        m = cmd.gate.matrix
        return m is not None and len(m) <= 2**5

    def receive(self, cmds):
        for cmd in cmds:
            if isinstance(cmd.gate, FlushGate):
                self.sim.run()
            else:
                self._handle(cmd)

    def _handle(self, cmd):
        gate = cmd.gate
        ctrl = list(cmd.control_qubits)
        if isinstance(gate, MeasureGate):
            if ctrl:
                raise ValueError("Cannot have control qubits with a measurement gate!")
            ids = [q for reg in cmd.qubits for q in reg]
            out = self.sim.measure_qubits(ids)
            if self.main is not None:
                for q, bit in zip(ids, out):
                    self.main.results[q] = bool(bit)
        elif isinstance(gate, AllocateQubitGate):
            self.sim.allocate_qubit(cmd.qubits[0][0])
        elif isinstance(gate, DeallocateQubitGate):
            self.sim.deallocate_qubit(cmd.qubits[0][0])
        elif isinstance(gate, BasicMathGate):
            quregs = [list(reg) for reg in cmd.qubits]
            if isinstance(gate, AddConstant):
                self.sim.emulate_math_addConstant(gate.a, quregs, ctrl)
            elif isinstance(gate, AddConstantModN):
                self.sim.emulate_math_addConstantModN(gate.a, gate.N, quregs, ctrl)
            elif isinstance(gate, MultiplyByConstantModN):
                self.sim.emulate_math_multiplyByConstantModN(gate.a, gate.N, quregs, ctrl)
            else:
                self.sim.emulate_math(gate.math_function(quregs), quregs, ctrl)
        elif isinstance(gate, TimeEvolution):
            terms = [(list(term), coeff) for term, coeff in gate.terms.items()]
            self.sim.emulate_time_evolution(terms, gate.time, list(cmd.qubits[0]), ctrl)
        elif gate.matrix is not None and len(gate.matrix) <= 2**5:
            m = gate.matrix
            ids = [q for reg in cmd.qubits for q in reg]
            if 2 ** len(ids) != len(m):
                raise Exception(f"Simulator: Error applying {gate!r} gate: {len(m)}x{len(m)} matrix on {len(ids)} qubits.")
            self.sim.apply_controlled_gate(m, ids, ctrl)
            if not self.gate_fusion:
                self.sim.run()
        else:
            raise Exception("This simulator only supports controlled k-qubit gates with k < 6!")


def get_engine_list():
    """projectq/setups/default.py:31-34."""
    return [TagRemover(), LocalOptimizer(10), AutoReplacer(), TagRemover(), LocalOptimizer(10)]


class MainEngine(_Frontend):
    """Program front end: qubit ids in allocation order, measurement results, meta sections and
    the compiler engine list in front of `backend` (projectq/cengines/_main.py)."""

    def __init__(self, backend, engine_list=None):
        engines = get_engine_list() if engine_list is None else list(engine_list)
        chain = engines + [backend]
        for a, b in zip(chain, chain[1:]):
            a.next_engine = b
        backend.main = self
        self.backend = backend
        self.engines = engines
        self.results = {}
        self._next_id = 0
        super().__init__(chain[0])

    def allocate_qubit(self):
        q = self._next_id
        self._next_id += 1
        self.receive([Command(Allocate, [[q]])])
        return q

    def allocate_qureg(self, n):
        return [self.allocate_qubit() for _ in range(n)]

    def deallocate_qubit(self, q):
        self.receive([Command(Deallocate, [[q]])])

    def measure(self, q):
        self.receive([Command(Measure, [[q]])])

    def flush(self):
        self.receive([Command(FlushGate(), [[-1]])])


# ============================================================================================
# the JSON op language of the recorded programs
# ============================================================================================
_FIXED_GATES = {
    "H": HGate, "X": XGate, "Y": YGate, "Z": ZGate, "S": SGate, "T": TGate, "SqrtX": SqrtXGate,
    "Swap": SwapGate, "SqrtSwap": SqrtSwapGate, "QFT": QFTGate, "Barrier": BarrierGate,
    "AddQuantum": AddQuantumGate, "SubtractQuantum": SubtractQuantumGate,
    "ComparatorQuantum": ComparatorQuantumGate, "MultiplyQuantum": MultiplyQuantumGate,
    "DivideQuantum": DivideQuantumGate, "Entangle": EntangleGate,
}
_PARAM_GATES = {
    "Rx": Rx, "Ry": Ry, "Rz": Rz, "Rxx": Rxx, "Ryy": Ryy, "Rzz": Rzz, "Ph": Ph, "R": R,
    "AddConstant": AddConstant, "SubConstant": SubConstant, "AddConstantModN": AddConstantModN,
    "SubConstantModN": SubConstantModN, "MultiplyByConstantModN": MultiplyByConstantModN,
}


def make_gate(name, params=()):
    """Gate from its name: the fixed gates, '<name>dag' for the inverse, parametrised gates and
    'Matrix' (params = rows of [re, im] pairs)."""
    if name in _FIXED_GATES:
        return _FIXED_GATES[name]()
    if name.endswith("dag") and name[:-3] in _FIXED_GATES:
        return get_inverse(_FIXED_GATES[name[:-3]]())
    if name == "Matrix":
        return MatrixGate([[complex(re, im) for re, im in row] for row in params])
    if name == "QubitOperator":  # params = [[[index, "X"], ...], [re, im]]
        return QubitOperatorGate([tuple(lp) for lp in params[0]], complex(*params[1]))
    if name == "TimeEvolution":  # params = [time, [[term, coeff], ...]], term = [[index, "X"], ...]
        return TimeEvolution(params[0], [(tuple(tuple(lp) for lp in term), c) for term, c in params[1]])
    if name in _PARAM_GATES:
        return _PARAM_GATES[name](*params)
    raise ValueError(f"unknown gate {name!r}")


def run_program(ops, backend, engine_list=None, engine=None, flush=True):
    """Run a program of the op language on `backend` (a SimulatorEngine) through the default
    engine list (or through `engine`, a MainEngine to continue), flushing at the end unless
    `flush` is false; returns the MainEngine (its `results` hold the measurement outcomes).

        ["alloc", k]  ["gate", name, params, regs, ctrls, state]  ["measure", q]  ["dealloc", q]
        ["flush"]  ["compute"]  ["compute_end"]  ["uncompute"]  ["dagger"]  ["dagger_end"]
        ["control", ctrls, state]  ["control_end"]
    """
    eng = MainEngine(backend, engine_list) if engine is None else engine
    open_sections = []
    for op in ops:
        kind = op[0]
        if kind == "alloc":
            eng.allocate_qureg(op[1])
        elif kind == "gate":
            _, name, params, regs, ctrls, state = op
            eng.apply(make_gate(name, params), *regs, controls=ctrls, control_state=state)
        elif kind == "measure":
            eng.measure(op[1])
        elif kind == "dealloc":
            eng.deallocate_qubit(op[1])
        elif kind == "flush":
            eng.flush()
        elif kind == "compute":
            eng.begin_compute()
            open_sections.append(None)
        elif kind == "compute_end":
            open_sections.pop()
            eng.end_compute()
        elif kind == "uncompute":
            eng.uncompute()
        elif kind in ("dagger", "control"):
            cm = eng.dagger() if kind == "dagger" else eng.control(op[1], op[2])
            cm.__enter__()
            open_sections.append(cm)
        elif kind in ("dagger_end", "control_end"):
            open_sections.pop().__exit__(None, None, None)
        else:
            raise ValueError(f"unknown op {kind!r}")
    if flush:
        eng.flush()
    return eng
