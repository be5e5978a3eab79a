"""ctypes binding over the simulator C ABI (include/projectq_amd.h).

`CAbiSimulator` reproduces the method surface, argument conventions and exception types of the
pybind11 class `projectq.backends._sim._cppsim.Simulator` (projectq/backends/_sim/_cppsim.cpp:43-67):
lists in, lists/dicts/complex out, `RuntimeError` / `ValueError` / `MemoryError` raised where the
reference's C++ throws `runtime_error` / `length_error`+`invalid_argument` / `bad_alloc`.

The class is bound to one shared library at construction.  `projectq_amd.Simulator` binds it to
the HIP library; nothing in the package binds anything else.
"""

import ctypes
import os
import math

import numpy as np

from . import _mathfuncs

QS_OK = 0
QS_ERR_RUNTIME = 1
QS_ERR_VALUE = 2
QS_ERR_MEMORY = 3
QS_ERR_DEVICE = 4
QS_ERR_CALLBACK = 5

PRECISIONS = {"c128": 0, "c64": 1}
FUSION_MODES = {"reference": 0, "window": 1}
PLANNERS = {"greedy": 0, "learned": 1, "tile": 2}

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_szp = ctypes.POINTER(ctypes.c_size_t)
_i32p = ctypes.POINTER(ctypes.c_int32)

MATH_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, _i32p, ctypes.c_size_t)

_SIGNATURES = {
    "qs_last_error": (ctypes.c_char_p, []),
    "qs_abi_version": (ctypes.c_int, []),
    "qs_create": (
        ctypes.c_int,
        [ctypes.c_uint32, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)],
    ),
    "qs_destroy": (None, [ctypes.c_void_p]),
    "qs_allocate_qubit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "qs_deallocate_qubit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "qs_get_classical_value": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double, ctypes.POINTER(ctypes.c_int)],
    ),
    "qs_is_classical": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]),
    "qs_measure_qubits": (ctypes.c_int, [ctypes.c_void_p, _u32p, ctypes.c_size_t, _u8p]),
    "qs_apply_controlled_gate": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t],
    ),
    "qs_emulate_math": (
        ctypes.c_int,
        [ctypes.c_void_p, MATH_FN, ctypes.c_void_p, _u32p, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t],
    ),
    "qs_emulate_math_op": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_int, _i64p, ctypes.c_size_t, _u32p, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t],
    ),
    "qs_get_expectation_value": (
        ctypes.c_int,
        [ctypes.c_void_p, _u32p, _u8p, _szp, ctypes.c_size_t, _f64p, _u32p, ctypes.c_size_t, _f64p],
    ),
    "qs_apply_qubit_operator": (
        ctypes.c_int,
        [ctypes.c_void_p, _u32p, _u8p, _szp, ctypes.c_size_t, _f64p, _u32p, ctypes.c_size_t],
    ),
    "qs_emulate_time_evolution": (
        ctypes.c_int,
        [
            ctypes.c_void_p,
            _u32p,
            _u8p,
            _szp,
            ctypes.c_size_t,
            _f64p,
            ctypes.c_double,
            _u32p,
            ctypes.c_size_t,
            _u32p,
            ctypes.c_size_t,
        ],
    ),
    "qs_get_probability": (ctypes.c_int, [ctypes.c_void_p, _u8p, _u32p, ctypes.c_size_t, _f64p]),
    "qs_get_amplitude": (ctypes.c_int, [ctypes.c_void_p, _u8p, _u32p, ctypes.c_size_t, _f64p]),
    "qs_set_wavefunction": (ctypes.c_int, [ctypes.c_void_p, _f64p, ctypes.c_size_t, _u32p, ctypes.c_size_t]),
    "qs_collapse_wavefunction": (ctypes.c_int, [ctypes.c_void_p, _u32p, ctypes.c_size_t, _u8p, ctypes.c_size_t]),
    "qs_run": (ctypes.c_int, [ctypes.c_void_p]),
    "qs_cheat_sizes": (ctypes.c_int, [ctypes.c_void_p, _szp, _szp]),
    "qs_cheat": (ctypes.c_int, [ctypes.c_void_p, _u32p, _u32p, _f64p]),
    "qs_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "qs_set_profiling": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "qs_get_stats": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), _f64p, _f64p, ctypes.POINTER(ctypes.c_uint64)],
    ),
    "qs_reset_stats": (ctypes.c_int, [ctypes.c_void_p]),
    "qs_overlap": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _f64p]),
    "qs_expectation_value_noalloc": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, _u32p, _u8p, _szp, ctypes.c_size_t, _f64p, _u32p,
         ctypes.c_size_t, _f64p],
    ),
    "qs_compare_states": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _f64p, ctypes.POINTER(ctypes.c_uint64)]),
    "qs_set_tile_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    "qs_set_slot_rule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "qs_set_relabel": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "qs_last_plan_relabels": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32)]),
    "qs_last_plan_rule": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "qs_set_recording": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "qs_take_records": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_size_t, _szp],
    ),
    "qs_set_planner_model": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t]),
    "qs_last_plan": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _u32p, _u32p, _u32p, _u32p, _u32p, _f64p],
    ),
    "qs_mlp_num_weights": (ctypes.c_size_t, []),
    "qs_mlp_predict": (
        ctypes.c_int,
        [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_size_t, ctypes.POINTER(ctypes.c_float)],
    ),
    "qs_mlp_loss_grad": (
        ctypes.c_int,
        [
            ctypes.POINTER(ctypes.c_float),
            ctypes.POINTER(ctypes.c_float),
            ctypes.POINTER(ctypes.c_float),
            ctypes.c_size_t,
            ctypes.POINTER(ctypes.c_float),
            ctypes.POINTER(ctypes.c_float),
        ],
    ),
    "qs_mlp_train": (
        ctypes.c_int,
        [
            ctypes.POINTER(ctypes.c_float),
            ctypes.POINTER(ctypes.c_float),
            ctypes.c_size_t,
            ctypes.c_int,
            ctypes.c_int,
            ctypes.c_float,
            ctypes.c_uint32,
            ctypes.POINTER(ctypes.c_float),
            ctypes.POINTER(ctypes.c_float),
            ctypes.POINTER(ctypes.c_float),
        ],
    ),
    "qs_batch_create": (
        ctypes.c_int,
        [
            ctypes.c_int,
            ctypes.c_size_t,
            _u32p,
            _szp,
            _u32p,
            _u32p,
            _u32p,
            _f64p,
            _szp,
            _szp,
            _u32p,
            _u8p,
            _f64p,
            ctypes.c_int,
            ctypes.POINTER(ctypes.c_void_p),
        ],
    ),
    "qs_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _f64p, _f64p]),
    "qs_batch_shape": (
        ctypes.c_int,
        [ctypes.c_void_p, _u32p, _u32p, _u32p, _u32p],
    ),
    "qs_batch_relabel": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "qs_batch_destroy": (None, [ctypes.c_void_p]),
    "qs_device_state": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int)],
    ),
    "qs_qubit_map": (ctypes.c_int, [ctypes.c_void_p, _u32p, _u32p, ctypes.c_size_t, _szp]),
    "qs_save_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "qs_load_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p]),
    "qs_set_wavefunction_device": (
        ctypes.c_int,
        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, _u32p, ctypes.c_size_t],
    ),
}

PASS_FEATURES = 24

#: every symbol the C ABI declares (include/projectq_amd.h)
ABI_SYMBOLS = tuple(_SIGNATURES)


def bind_library(lib):
    """Set argtypes/restype for every C-ABI symbol of `lib` (a ctypes.CDLL); returns `lib`."""
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _ptr(arr, ctype):
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


def _u32(seq):
    arr = np.ascontiguousarray(np.asarray(seq, dtype=np.int64).reshape(-1))
    if arr.size and (arr.min() < 0 or arr.max() > 0xFFFFFFFF):
        raise TypeError("qubit ids must be unsigned 32-bit integers")
    return arr.astype(np.uint32)


def _u32_array(seq):
    """Qubit ids as a ctypes uint32 array; ids outside [0, 2^32) raise TypeError, as _u32 does."""
    if not isinstance(seq, (list, tuple)):
        arr = _u32(seq)
        return (ctypes.c_uint32 * arr.size)(*arr.tolist())
    for v in seq:
        if not 0 <= v <= 0xFFFFFFFF:
            raise TypeError("qubit ids must be unsigned 32-bit integers")
    return (ctypes.c_uint32 * len(seq))(*seq)


def _bits(seq):
    """Booleans as uint8; '0'/'1' characters count as their digit (_simulator.py:221, :247)."""
    return np.ascontiguousarray(
        np.array([1 if (bool(int(b)) if isinstance(b, str) else bool(b)) else 0 for b in seq], dtype=np.uint8)
    )


def _pauli_terms(terms, complex_coeffs):
    """Flatten [(list[(idx, 'X'|'Y'|'Z')], coeff)] into the C-ABI arrays."""
    idx, ops, offsets, coeffs = [], [], [0], []
    for term, coeff in terms:
        for local, op in term:
            op = str(op)
            if op not in ("X", "Y", "Z"):
                raise ValueError(f"Pauli operator must be 'X', 'Y' or 'Z', got {op!r}")
            idx.append(int(local))
            ops.append(ord(op))
        offsets.append(len(idx))
        if complex_coeffs:
            c = complex(coeff)
            coeffs.extend([c.real, c.imag])
        else:
            if isinstance(coeff, complex):
                raise TypeError("coefficients must be real for this method")
            coeffs.append(float(coeff))
    return (
        _u32(idx),
        np.asarray(ops, dtype=np.uint8),
        np.asarray(offsets, dtype=np.uintp),
        np.asarray(coeffs, dtype=np.float64),
    )


def expectation_value_noalloc(lib, state, positions, terms):
    """qs_expectation_value_noalloc: sum_t c_t Re<psi|P_t|psi> on a device state (a DeviceState or
    any 1-d contiguous complex64/complex128 array with `__cuda_array_interface__` on the current
    GPU) with no allocation; local index l of a string is the qubit at bit position positions[l].
    `terms` in the backend's format [(list[(local index, 'X'|'Y'|'Z')], real coefficient)]."""
    iface = state.__cuda_array_interface__
    n_amps, typestr = int(iface["shape"][0]), iface["typestr"]
    if typestr not in ("<c8", "<c16"):
        raise ValueError("expectation_value_noalloc: complex64 or complex128 amplitudes")
    n_qubits = n_amps.bit_length() - 1
    if n_amps != 1 << n_qubits:
        raise ValueError("expectation_value_noalloc: the state must hold 2^n amplitudes")
    idx, ops, offs, coeffs = _pauli_terms(terms, complex_coeffs=False)
    pos = _u32(positions)
    out = ctypes.c_double()
    rc = lib.qs_expectation_value_noalloc(
        ctypes.c_void_p(int(iface["data"][0])), n_qubits, 1 if typestr == "<c8" else 0, _ptr(idx, ctypes.c_uint32),
        _ptr(ops, ctypes.c_uint8), _ptr(offs, ctypes.c_size_t), len(offs) - 1, _ptr(coeffs, ctypes.c_double),
        _ptr(pos, ctypes.c_uint32), len(pos), ctypes.byref(out))
    if rc != QS_OK:
        msg = (lib.qs_last_error() or b"").decode(errors="replace")
        raise {QS_ERR_VALUE: ValueError, QS_ERR_MEMORY: MemoryError}.get(rc, RuntimeError)(msg)
    return out.value


def _regs(quregs):
    flat, sizes = [], []
    for reg in quregs:
        reg = list(reg)
        flat.extend(reg)
        sizes.append(len(reg))
    return _u32(flat), np.asarray(sizes, dtype=np.uint32)


class CAbiSimulator:
    """The `_cppsim.Simulator` method surface over a C-ABI library."""

    def __init__(self, lib, seed, precision="c128", fusion="reference", planner="greedy", capacity_qubits=0):
        self._lib = lib
        self._handle = None
        self.precision = precision
        self.fusion = fusion
        self.planner = planner
        handle = ctypes.c_void_p()
        try:
            prec = PRECISIONS[precision]
            fmode = FUSION_MODES[fusion]
            plan = PLANNERS[planner]
        except KeyError as err:
            raise ValueError(f"unknown option {err}") from None
        seed = int(seed)
        if seed < 0 or seed > 0xFFFFFFFF:
            raise TypeError("seed must be an unsigned 32-bit integer")
        self._check(lib.qs_create(seed, prec, fmode, plan, int(capacity_qubits), ctypes.byref(handle)))
        self._handle = handle
        self._cb_error = None

    # ------------------------------------------------------------------ plumbing
    def _check(self, status):
        if status == QS_OK:
            return
        msg = (self._lib.qs_last_error() or b"").decode(errors="replace")
        if status == QS_ERR_CALLBACK and getattr(self, "_cb_error", None) is not None:
            err, self._cb_error = self._cb_error, None
            raise err
        if status == QS_ERR_VALUE:
            raise ValueError(msg)
        if status == QS_ERR_MEMORY:
            raise MemoryError(msg)
        raise RuntimeError(msg)

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle is not None and handle.value:
            self._lib.qs_destroy(handle)
            self._handle = None

    # ------------------------------------------------------------------ _cppsim surface
    def allocate_qubit(self, qubit_id):
        self._check(self._lib.qs_allocate_qubit(self._handle, int(qubit_id)))

    def deallocate_qubit(self, qubit_id):
        self._check(self._lib.qs_deallocate_qubit(self._handle, int(qubit_id)))

    def get_classical_value(self, qubit_id, tol):
        out = ctypes.c_int()
        self._check(self._lib.qs_get_classical_value(self._handle, int(qubit_id), float(tol), ctypes.byref(out)))
        return bool(out.value)

    def is_classical(self, qubit_id, tol):
        out = ctypes.c_int()
        self._check(self._lib.qs_is_classical(self._handle, int(qubit_id), float(tol), ctypes.byref(out)))
        return bool(out.value)

    def measure_qubits(self, ids):
        ids = _u32(ids)
        out = np.zeros(len(ids), dtype=np.uint8)
        self._check(self._lib.qs_measure_qubits(self._handle, _ptr(ids, ctypes.c_uint32), len(ids), _ptr(out, ctypes.c_uint8)))
        return [bool(v) for v in out]

    def apply_controlled_gate(self, m, ids, ctrl):
        # the per-gate path of every circuit: short id lists go straight into ctypes arrays and a
        # complex128 C-contiguous matrix is passed by address, without intermediate numpy copies
        ids = _u32_array(ids)
        ctrl = _u32_array(ctrl)
        mat = m
        if not (type(mat) is np.ndarray and mat.dtype == np.complex128 and mat.flags.c_contiguous):
            mat = np.ascontiguousarray(np.asarray(m, dtype=np.complex128))
        dim = 1 << len(ids)
        if mat.shape != (dim, dim):
            raise ValueError(f"apply_controlled_gate: a gate on {len(ids)} qubits needs a {dim}x{dim} matrix, got {mat.shape}")
        self._check(self._lib.qs_apply_controlled_gate(self._handle, mat.ctypes.data, ids, len(ids), ctrl, len(ctrl)))

    def _math_op(self, opcode, params, quregs, ctrl):
        flat, sizes = _regs(quregs)
        ctrl = _u32(ctrl)
        prm = np.asarray(params, dtype=np.int64) if params else np.zeros(1, dtype=np.int64)
        self._check(
            self._lib.qs_emulate_math_op(
                self._handle,
                int(opcode),
                _ptr(prm, ctypes.c_int64),
                len(params),
                _ptr(flat, ctypes.c_uint32),
                _ptr(sizes, ctypes.c_uint32),
                len(sizes),
                _ptr(ctrl, ctypes.c_uint32),
                len(ctrl),
            )
        )

    def emulate_math(self, f, quregs, ctrl):
        """simulator.hpp:224-269 through _cppsim.cpp:33-41.

        Functions of projectq.libs.math (and of projectq_amd.gates) are recognised and run as
        native integer permutations after a sampled cross-check against `f` itself; anything
        else is called back once per distinct tuple of register values.
        """
        quregs = [list(r) for r in quregs]
        recognised = _mathfuncs.recognise(f, quregs)
        if recognised is not None:
            opcode, params = recognised
            self._math_op(opcode, params, quregs, ctrl)
            return
        flat, sizes = _regs(quregs)
        ctrl = _u32(ctrl)
        self._cb_error = None

        def trampoline(_user, values, n_regs):
            try:
                res = f([int(values[i]) for i in range(n_regs)])
                res = list(res)
                if len(res) < n_regs:
                    raise ValueError("emulate_math: function returned fewer values than registers")
                for i in range(n_regs):
                    v = int(res[i])
                    if v < -(2**31) or v > 2**31 - 1:
                        raise RuntimeError("Unable to cast Python instance to C++ type int")
                    values[i] = v
                return 0
            except BaseException as err:  # noqa: BLE001 -- re-raised by _check
                self._cb_error = err
                return 1

        cb = MATH_FN(trampoline)
        self._check(
            self._lib.qs_emulate_math(
                self._handle,
                cb,
                None,
                _ptr(flat, ctypes.c_uint32),
                _ptr(sizes, ctypes.c_uint32),
                len(sizes),
                _ptr(ctrl, ctypes.c_uint32),
                len(ctrl),
            )
        )

    def emulate_math_addConstant(self, a, quregs, ctrl):  # noqa: N802 (reference name)
        self._math_op(_mathfuncs.ADD_CONSTANT, [int(a)], quregs, ctrl)

    def emulate_math_addConstantModN(self, a, N, quregs, ctrl):  # noqa: N802,N803
        self._math_op(_mathfuncs.ADD_CONSTANT_MOD_N, [int(a), int(N)], quregs, ctrl)

    def emulate_math_multiplyByConstantModN(self, a, N, quregs, ctrl):  # noqa: N802,N803
        self._math_op(_mathfuncs.MULTIPLY_BY_CONSTANT_MOD_N, [int(a), int(N)], quregs, ctrl)

    def get_expectation_value(self, terms, ids):
        idx, ops, offs, coeffs = _pauli_terms(terms, complex_coeffs=False)
        ids = _u32(ids)
        out = ctypes.c_double()
        self._check(
            self._lib.qs_get_expectation_value(
                self._handle,
                _ptr(idx, ctypes.c_uint32),
                _ptr(ops, ctypes.c_uint8),
                _ptr(offs, ctypes.c_size_t),
                len(offs) - 1,
                _ptr(coeffs, ctypes.c_double),
                _ptr(ids, ctypes.c_uint32),
                len(ids),
                ctypes.byref(out),
            )
        )
        return out.value

    def apply_qubit_operator(self, terms, ids):
        idx, ops, offs, coeffs = _pauli_terms(terms, complex_coeffs=True)
        ids = _u32(ids)
        self._check(
            self._lib.qs_apply_qubit_operator(
                self._handle,
                _ptr(idx, ctypes.c_uint32),
                _ptr(ops, ctypes.c_uint8),
                _ptr(offs, ctypes.c_size_t),
                len(offs) - 1,
                _ptr(coeffs, ctypes.c_double),
                _ptr(ids, ctypes.c_uint32),
                len(ids),
            )
        )

    def emulate_time_evolution(self, terms, time, ids, ctrl):
        idx, ops, offs, coeffs = _pauli_terms(terms, complex_coeffs=False)
        ids = _u32(ids)
        ctrl = _u32(ctrl)
        self._check(
            self._lib.qs_emulate_time_evolution(
                self._handle,
                _ptr(idx, ctypes.c_uint32),
                _ptr(ops, ctypes.c_uint8),
                _ptr(offs, ctypes.c_size_t),
                len(offs) - 1,
                _ptr(coeffs, ctypes.c_double),
                float(time),
                _ptr(ids, ctypes.c_uint32),
                len(ids),
                _ptr(ctrl, ctypes.c_uint32),
                len(ctrl),
            )
        )

    def get_probability(self, bit_string, ids):
        bits = _bits(bit_string)
        ids = _u32(ids)
        if len(bits) < len(ids):
            raise ValueError("get_probability: fewer bits than qubit ids")
        out = ctypes.c_double()
        self._check(
            self._lib.qs_get_probability(
                self._handle, _ptr(bits, ctypes.c_uint8), _ptr(ids, ctypes.c_uint32), len(ids), ctypes.byref(out)
            )
        )
        return out.value

    def get_amplitude(self, bit_string, ids):
        bits = _bits(bit_string)
        ids = _u32(ids)
        if len(bits) < len(ids):
            raise ValueError("get_amplitude: fewer bits than qubit ids")
        out = np.zeros(2, dtype=np.float64)
        self._check(
            self._lib.qs_get_amplitude(
                self._handle, _ptr(bits, ctypes.c_uint8), _ptr(ids, ctypes.c_uint32), len(ids), _ptr(out, ctypes.c_double)
            )
        )
        return complex(out[0], out[1])

    def set_wavefunction(self, wavefunction, ordering):
        wf = np.ascontiguousarray(np.asarray(wavefunction, dtype=np.complex128).reshape(-1))
        ordering = _u32(ordering)
        self._check(
            self._lib.qs_set_wavefunction(
                self._handle, _ptr(wf.view(np.float64), ctypes.c_double), len(wf), _ptr(ordering, ctypes.c_uint32), len(ordering)
            )
        )

    def collapse_wavefunction(self, ids, values):
        ids = _u32(ids)
        vals = _bits(values)
        self._check(
            self._lib.qs_collapse_wavefunction(
                self._handle, _ptr(ids, ctypes.c_uint32), len(ids), _ptr(vals, ctypes.c_uint8), len(vals)
            )
        )

    def run(self):
        self._check(self._lib.qs_run(self._handle))

    # ------------------------------------------------------------------ extensions
    def synchronize(self):
        """Wait for every queued device operation (gate passes run asynchronously)."""
        self._check(self._lib.qs_synchronize(self._handle))

    def set_profiling(self, enabled=True):
        self._check(self._lib.qs_set_profiling(self._handle, 1 if enabled else 0))

    def stats(self):
        """dict(passes, bytes, gpu_ms, timed_passes) of the gate passes since the last reset."""
        passes, timed = ctypes.c_uint64(), ctypes.c_uint64()
        nbytes, ms = ctypes.c_double(), ctypes.c_double()
        self._check(
            self._lib.qs_get_stats(self._handle, ctypes.byref(passes), ctypes.byref(nbytes), ctypes.byref(ms), ctypes.byref(timed))
        )
        return {"passes": passes.value, "bytes": nbytes.value, "gpu_ms": ms.value, "timed_passes": timed.value}

    def reset_stats(self):
        self._check(self._lib.qs_reset_stats(self._handle))

    def overlap(self, other):
        """<self|other> computed on the GPU (same qubits at the same positions)."""
        if other._lib is not self._lib:
            raise ValueError("overlap needs two simulators of the same library")
        out = np.zeros(2, dtype=np.float64)
        self._check(self._lib.qs_overlap(self._handle, other._handle, _ptr(out, ctypes.c_double)))
        return complex(out[0], out[1])

    def compare(self, other):
        """(max |a_i - b_i|, number of amplitudes whose stored values differ) between this
        simulator's state and `other`'s, computed on the GPU (same qubits at the same positions)."""
        if other._lib is not self._lib:
            raise ValueError("compare needs two simulators of the same library")
        mx = ctypes.c_double()
        cnt = ctypes.c_uint64()
        self._check(self._lib.qs_compare_states(self._handle, other._handle, ctypes.byref(mx), ctypes.byref(cnt)))
        return mx.value, cnt.value

    # ------------------------------------------------------------------ planner controls
    def set_tile_params(self, tile_bits, reg_bits, low_bits):
        self._check(self._lib.qs_set_tile_params(self._handle, int(tile_bits), int(reg_bits), int(low_bits)))

    SLOT_RULES = ("lookahead", "earliest")

    def set_slot_rule(self, rule):
        """Slot rule of the heuristic tile planner: "lookahead" (default) or "earliest"."""
        self._check(self._lib.qs_set_slot_rule(self._handle, self.SLOT_RULES.index(rule)))

    RELABEL_MODES = {"default": -1, "off": 0, "on": 1}

    def set_relabel(self, mode):
        """Qubit relabelling at tile stores: "default" (on for the learned planner, off for the
        heuristic tile planner), "off" or "on"."""
        self._check(self._lib.qs_set_relabel(self._handle, self.RELABEL_MODES[mode]))

    def set_recording(self, enabled=True):
        self._check(self._lib.qs_set_recording(self._handle, 1 if enabled else 0))

    def take_records(self, cap=1 << 16):
        """(features [rows, PASS_FEATURES] float32, pass times in ms) of the recorded passes."""
        feats = np.zeros((cap, PASS_FEATURES), dtype=np.float32)
        ms = np.zeros(cap, dtype=np.float32)
        rows = ctypes.c_size_t()
        self._check(
            self._lib.qs_take_records(
                self._handle, _ptr(feats, ctypes.c_float), _ptr(ms, ctypes.c_float), cap, ctypes.byref(rows)
            )
        )
        return feats[: rows.value].copy(), ms[: rows.value].copy()

    def set_planner_model(self, weights):
        if weights is None:
            self._check(self._lib.qs_set_planner_model(self._handle, None, 0))
            return
        w = np.ascontiguousarray(np.asarray(weights, dtype=np.float32).reshape(-1))
        self._check(self._lib.qs_set_planner_model(self._handle, _ptr(w, ctypes.c_float), len(w)))

    def last_plan(self):
        kind = ctypes.c_int()
        t, rb, low, passes, cands = (ctypes.c_uint32() for _ in range(5))
        pred = ctypes.c_double()
        self._check(
            self._lib.qs_last_plan(
                self._handle,
                ctypes.byref(kind),
                ctypes.byref(t),
                ctypes.byref(rb),
                ctypes.byref(low),
                ctypes.byref(passes),
                ctypes.byref(cands),
                ctypes.byref(pred),
            )
        )
        rule = ctypes.c_int()
        self._check(self._lib.qs_last_plan_rule(self._handle, ctypes.byref(rule)))
        relabels = ctypes.c_uint32()
        self._check(self._lib.qs_last_plan_relabels(self._handle, ctypes.byref(relabels)))
        return {
            "kind": {0: "greedy_blocks", 1: "tile"}.get(kind.value, "none"),
            "rule": self.SLOT_RULES[rule.value] if kind.value == 1 else None,
            "relabels": relabels.value,
            "tile_bits": t.value,
            "reg_bits": rb.value,
            "low_bits": low.value,
            "passes": passes.value,
            "candidates": cands.value,
            "predicted_ms": pred.value,
        }

    def cheat_numpy(self):
        """cheat() without the Python-list conversion: (dict id->position, complex128 ndarray)."""
        nq = ctypes.c_size_t()
        na = ctypes.c_size_t()
        self._check(self._lib.qs_cheat_sizes(self._handle, ctypes.byref(nq), ctypes.byref(na)))
        ids = np.zeros(max(nq.value, 1), dtype=np.uint32)
        pos = np.zeros(max(nq.value, 1), dtype=np.uint32)
        state = np.zeros(na.value, dtype=np.complex128)
        self._check(
            self._lib.qs_cheat(
                self._handle, _ptr(ids, ctypes.c_uint32), _ptr(pos, ctypes.c_uint32), _ptr(state.view(np.float64), ctypes.c_double)
            )
        )
        mapping = {int(ids[i]): int(pos[i]) for i in range(nq.value)}
        return mapping, state

    def qubit_map(self):
        """dict qubit id -> bit position, without copying the state."""
        n = ctypes.c_size_t()
        self._check(self._lib.qs_qubit_map(self._handle, None, None, ctypes.c_size_t(0), ctypes.byref(n)))
        ids = np.zeros(max(n.value, 1), dtype=np.uint32)
        pos = np.zeros(max(n.value, 1), dtype=np.uint32)
        self._check(
            self._lib.qs_qubit_map(
                self._handle, _ptr(ids, ctypes.c_uint32), _ptr(pos, ctypes.c_uint32), ctypes.c_size_t(n.value), ctypes.byref(n)
            )
        )
        return {int(ids[i]): int(pos[i]) for i in range(n.value)}

    def device_state(self):
        """The state in device memory without a host copy: (dict id -> position, DeviceState).

        The DeviceState aliases the simulator's buffer (complex64 or complex128, index bit p <->
        the qubit at position p) and is valid until the next call that applies gates, allocates,
        deallocates or measures.  It exposes `__cuda_array_interface__`, so array libraries that
        share this process's HIP runtime can wrap it without a copy (import torch before
        projectq_amd: torch's libraries load their own HIP runtime otherwise)."""
        ptr = ctypes.c_void_p()
        n = ctypes.c_uint64()
        prec = ctypes.c_int()
        self._check(self._lib.qs_device_state(self._handle, ctypes.byref(ptr), ctypes.byref(n), ctypes.byref(prec)))
        return self.qubit_map(), DeviceState(int(ptr.value or 0), int(n.value), prec.value == 1, self)

    def save_state(self, path):
        """Checkpoint the qubit map and the state to a binary file (qs_save_state)."""
        self._check(self._lib.qs_save_state(self._handle, os.fsencode(path)))

    def load_state(self, path):
        """Restore a checkpoint written by save_state (qs_load_state)."""
        self._check(self._lib.qs_load_state(self._handle, os.fsencode(path)))

    def set_wavefunction_device(self, src, ordering):
        """set_wavefunction from device memory on this GPU: a DeviceState, or any object with
        `__cuda_array_interface__` (a 1-d contiguous complex64 or complex128 array, e.g. a torch
        tensor), converted to this simulator's precision if needed."""
        iface = src.__cuda_array_interface__
        shape, typestr, strides = iface["shape"], iface["typestr"], iface.get("strides")
        if len(shape) != 1 or strides not in (None, (np.dtype(typestr).itemsize,)):
            raise ValueError("set_wavefunction_device: need a contiguous 1-d array")
        if typestr == "<c8":
            prec = 1
        elif typestr == "<c16":
            prec = 0
        else:
            raise ValueError("set_wavefunction_device: complex64 or complex128 amplitudes")
        stream = iface.get("stream")
        if hasattr(src, "is_cuda") and stream is None:  # torch: finish its pending work first
            import torch

            torch.cuda.current_stream().synchronize()
        order = np.asarray(list(ordering), dtype=np.uint32)
        self._check(
            self._lib.qs_set_wavefunction_device(
                self._handle, ctypes.c_void_p(int(iface["data"][0])), ctypes.c_uint64(int(shape[0])), ctypes.c_int(prec),
                _ptr(order, ctypes.c_uint32), ctypes.c_size_t(len(order)),
            )
        )

    def cheat(self):
        """simulator.hpp:529-532: (dict id -> bit position, list of complex amplitudes)."""
        mapping, state = self.cheat_numpy()
        return mapping, state.tolist()


class DeviceState:
    """A simulator's state in device memory, without a copy (see CAbiSimulator.device_state).
    `ptr` is the device address, `n_amps` the amplitude count, `c64` the precision."""

    def __init__(self, ptr, n_amps, c64, owner):
        self.ptr, self.n_amps, self.c64 = ptr, n_amps, c64
        self._owner = owner  # keeps the simulator (and so the buffer) alive

    @property
    def __cuda_array_interface__(self):
        return {
            "shape": (self.n_amps,),
            "typestr": "<c8" if self.c64 else "<c16",
            "data": (self.ptr, False),
            "version": 2,
            "strides": None,
        }


class CAbiBatch:
    """Batched circuits over the C ABI (qs_batch_*): `circuits` is a list of (n_qubits, gate
    commands) with qubits 0..n-1 starting in |0...0>; `operators` one Pauli operator per circuit
    in the backend's term format [(list[(qubit, 'X'|'Y'|'Z')], real coefficient)]."""

    def __init__(self, lib, circuits, operators, precision="c64", autotune=True):
        if len(circuits) != len(operators):
            raise ValueError("one operator per circuit")
        self._lib = lib
        self._handle = None
        nq, goff, gk, gnc, gq, mats = [], [0], [], [], [], []
        toff, ooff, oq, ok, co = [0], [0], [], [], []
        for (n, cmds), terms in zip(circuits, operators):
            nq.append(int(n))
            for c in cmds:
                if c[0] != "gate":
                    continue
                _, m, targets, controls = c
                gk.append(len(targets))
                gnc.append(len(controls))
                gq.extend(int(q) for q in targets)
                gq.extend(int(q) for q in controls)
                mats.append(np.asarray(m, dtype=np.complex128).reshape(-1))
            goff.append(len(gk))
            for term, coeff in terms:
                for q, op in term:
                    oq.append(int(q))
                    ok.append(ord(str(op)))
                ooff.append(len(oq))
                co.append(float(coeff))
            toff.append(len(co))
        arrs = dict(
            nq=_u32(nq),
            goff=np.asarray(goff, np.uintp),
            gk=_u32(gk) if gk else np.zeros(1, np.uint32),
            gnc=_u32(gnc) if gnc else np.zeros(1, np.uint32),
            gq=_u32(gq) if gq else np.zeros(1, np.uint32),
            mats=np.ascontiguousarray(np.concatenate(mats)).view(np.float64) if mats else np.zeros(2),
            toff=np.asarray(toff, np.uintp),
            ooff=np.asarray(ooff, np.uintp),
            oq=_u32(oq) if oq else np.zeros(1, np.uint32),
            ok=np.asarray(ok if ok else [0], np.uint8),
            co=np.asarray(co if co else [0.0], np.float64),
        )
        self._keep = arrs
        self.n_circuits = len(nq)
        handle = ctypes.c_void_p()
        status = lib.qs_batch_create(
            PRECISIONS[precision],
            len(nq),
            _ptr(arrs["nq"], ctypes.c_uint32),
            _ptr(arrs["goff"], ctypes.c_size_t),
            _ptr(arrs["gk"], ctypes.c_uint32),
            _ptr(arrs["gnc"], ctypes.c_uint32),
            _ptr(arrs["gq"], ctypes.c_uint32),
            _ptr(arrs["mats"], ctypes.c_double),
            _ptr(arrs["toff"], ctypes.c_size_t),
            _ptr(arrs["ooff"], ctypes.c_size_t),
            _ptr(arrs["oq"], ctypes.c_uint32),
            _ptr(arrs["ok"], ctypes.c_uint8),
            _ptr(arrs["co"], ctypes.c_double),
            1 if autotune else 0,
            ctypes.byref(handle),
        )
        self._check(status)
        self._handle = handle

    def _check(self, status):
        if status == QS_OK:
            return
        msg = (self._lib.qs_last_error() or b"").decode(errors="replace")
        if status == QS_ERR_VALUE:
            raise ValueError(msg)
        if status == QS_ERR_MEMORY:
            raise MemoryError(msg)
        raise RuntimeError(msg)

    def run(self, graph=True):
        """One run of every circuit: (energies ndarray, GPU milliseconds)."""
        out = np.zeros(self.n_circuits, dtype=np.float64)
        ms = ctypes.c_double()
        self._check(self._lib.qs_batch_run(self._handle, 1 if graph else 0, _ptr(out, ctypes.c_double), ctypes.byref(ms)))
        return out, ms.value

    def shape(self):
        t, rb, low, launches = (ctypes.c_uint32() for _ in range(4))
        self._check(self._lib.qs_batch_shape(self._handle, ctypes.byref(t), ctypes.byref(rb), ctypes.byref(low), ctypes.byref(launches)))
        relabel = ctypes.c_int()
        self._check(self._lib.qs_batch_relabel(self._handle, ctypes.byref(relabel)))
        return {"tile_bits": t.value, "reg_bits": rb.value, "low_bits": low.value, "launches": launches.value,
                "relabel": bool(relabel.value)}

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle is not None and handle.value:
            self._lib.qs_batch_destroy(handle)
            self._handle = None


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class MLP:
    """The planner's cost model on the GPU (qs_mlp_*): 24 -> 64 -> 64 -> 1, ReLU, fp32."""

    def __init__(self, lib):
        self._lib = lib
        self.n_weights = int(lib.qs_mlp_num_weights())

    def _check(self, status):
        if status != QS_OK:
            msg = (self._lib.qs_last_error() or b"").decode(errors="replace")
            raise (ValueError if status == QS_ERR_VALUE else RuntimeError)(msg)

    def predict(self, weights, features):
        w, x = _f32(weights).reshape(-1), _f32(features).reshape(-1, PASS_FEATURES)
        out = np.zeros(len(x), dtype=np.float32)
        self._check(self._lib.qs_mlp_predict(_ptr(w, ctypes.c_float), _ptr(x, ctypes.c_float), len(x), _ptr(out, ctypes.c_float)))
        return out

    def loss_grad(self, weights, features, targets):
        w, x, t = _f32(weights).reshape(-1), _f32(features).reshape(-1, PASS_FEATURES), _f32(targets).reshape(-1)
        loss = ctypes.c_float()
        grad = np.zeros(self.n_weights, dtype=np.float32)
        self._check(
            self._lib.qs_mlp_loss_grad(
                _ptr(w, ctypes.c_float), _ptr(x, ctypes.c_float), _ptr(t, ctypes.c_float), len(x), ctypes.byref(loss),
                _ptr(grad, ctypes.c_float),
            )
        )
        return loss.value, grad

    def train(self, features, targets, epochs=200, batch=256, lr=1e-3, seed=1, w_init=None):
        x, t = _f32(features).reshape(-1, PASS_FEATURES), _f32(targets).reshape(-1)
        w_out = np.zeros(self.n_weights, dtype=np.float32)
        loss = ctypes.c_float()
        wi = None if w_init is None else _f32(w_init).reshape(-1)
        self._check(
            self._lib.qs_mlp_train(
                _ptr(x, ctypes.c_float), _ptr(t, ctypes.c_float), len(x), int(epochs), int(batch), float(lr), int(seed),
                None if wi is None else _ptr(wi, ctypes.c_float), _ptr(w_out, ctypes.c_float), ctypes.byref(loss),
            )
        )
        return w_out, loss.value


def qubit_count_for(n_amps):
    """log2 of a power-of-two amplitude count."""
    n = int(math.log2(n_amps))
    if 1 << n != n_amps:
        raise ValueError("amplitude count must be a power of two")
    return n
