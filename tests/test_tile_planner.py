"""CPU checks of the tile planner and the tile-pass index arithmetic.

tests/native/tile_emu.cpp runs the product planner (projectq_amd/csrc/planner.cpp) and replays
each pass with the kernel's index arithmetic on the CPU, bounds-checking every index.  The
result must equal the unfused gate sequence run by the oracle (<= 1e-12), every gate must be
scheduled exactly once, and no index may leave its buffer."""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle.cpu_sim import OracleSimulator
from projectq_amd import circuits, commands
from projectq_amd import gates as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
LIB = os.path.join(NATIVE, "_build", "libtile_emu.so")


@pytest.fixture(scope="module")
def emu():
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = [os.path.join(NATIVE, "tile_emu.cpp"), os.path.join(ROOT, "projectq_amd", "csrc", "planner.cpp")]
    subprocess.run(
        ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "projectq_amd", "csrc"), *srcs, "-o", LIB],
        check=True,
    )
    lib = ctypes.CDLL(LIB)
    lib.emu_run.restype = ctypes.c_int
    return lib


def run_emu(lib, n, t, lmin, cmds, state, rb=5, c64=True):
    ks, tpos, cm, mats = [], [], [], []
    for c in cmds:
        if c[0] != "gate":
            continue
        _, m, targets, controls = c
        ks.append(len(targets))
        tpos.extend(targets)
        mask = 0
        for q in controls:
            mask |= 1 << q
        cm.append(mask)
        mats.append(np.asarray(m, dtype=np.complex128).reshape(-1))
    k = np.asarray(ks, dtype=np.uint32)
    tp = np.asarray(tpos, dtype=np.uint32)
    cmask = np.asarray(cm, dtype=np.uint64)
    mflat = np.ascontiguousarray(np.concatenate(mats)) if mats else np.zeros(1, dtype=np.complex128)
    st = np.ascontiguousarray(state.astype(np.complex128))
    npass = ctypes.c_int()
    err = ctypes.create_string_buffer(256)
    rc = lib.emu_run(
        ctypes.c_uint(n),
        ctypes.c_uint(t),
        ctypes.c_uint(lmin),
        ctypes.c_uint(rb),
        ctypes.c_int(int(c64)),
        ctypes.c_int(len(ks)),
        k.ctypes.data_as(ctypes.c_void_p),
        tp.ctypes.data_as(ctypes.c_void_p),
        cmask.ctypes.data_as(ctypes.c_void_p),
        mflat.ctypes.data_as(ctypes.c_void_p),
        st.ctypes.data_as(ctypes.c_void_p),
        ctypes.byref(npass),
        err,
        ctypes.c_int(256),
    )
    assert rc == 0, err.value.decode()
    return st, npass.value


def oracle_state(cmds, n):
    o = OracleSimulator(1)
    commands.execute(o, cmds, gate_fusion=False)
    return o.cheat_numpy()[1]


def start_state(n):
    s = np.zeros(2**n, dtype=np.complex128)
    s[0] = 1.0
    return s


CASES = [
    ("random_layers", lambda n: circuits.random_layers(n, 6, n // 2, 3, measure=False)),
    ("mixed", lambda n: circuits.mixed_circuit(n, 8, 5)),
    ("qft", lambda n: circuits.allocate(n) + [commands.gate(G.H, [q]) for q in range(n)] + circuits.qft(list(range(n)))),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize(
    "n,t,lmin,rb,c64",
    [(8, 8, 5, 5, True), (12, 10, 4, 4, True), (14, 10, 5, 5, False), (15, 12, 4, 4, True), (15, 12, 4, 4, False),
     (16, 13, 5, 5, True), (4, 4, 4, 5, False)],
)
def test_tile_plan_matches_unfused(emu, name, make, n, t, lmin, rb, c64):
    cmds = make(n)
    want = oracle_state(cmds, n)
    got, npass = run_emu(emu, n, t, lmin, cmds, start_state(n), rb, c64)
    assert np.max(np.abs(got - want)) <= 1e-12
    n_gates = commands.count_gates(cmds)
    assert 1 <= npass <= n_gates


def gate_arrays(cmds):
    ks, tpos, cm, mats = [], [], [], []
    for c in cmds:
        if c[0] != "gate":
            continue
        _, m, targets, controls = c
        ks.append(len(targets))
        tpos.extend(targets)
        cm.append(sum(1 << q for q in controls))
        mats.append(np.asarray(m, dtype=np.complex128).reshape(-1))
    return (np.asarray(ks, dtype=np.uint32), np.asarray(tpos, dtype=np.uint32), np.asarray(cm, dtype=np.uint64),
            np.ascontiguousarray(np.concatenate(mats)) if mats else np.zeros(1, dtype=np.complex128))


def run_emu_relabel(lib, n, t, lmin, cmds, state, rb, c64, relabel, rule):
    """Tile passes that relabel qubits at their stores: returns the state re-indexed back to the
    original qubit positions, the pass count and the final position of every original position."""
    k, tp, cmask, mflat = gate_arrays(cmds)
    st = None if state is None else np.ascontiguousarray(state.astype(np.complex128))
    npass = ctypes.c_int()
    final = np.zeros(max(n, 1), dtype=np.uint32)
    err = ctypes.create_string_buffer(256)
    rc = lib.emu_run_relabel(
        ctypes.c_uint(n), ctypes.c_uint(t), ctypes.c_uint(lmin), ctypes.c_uint(rb), ctypes.c_int(int(c64)),
        ctypes.c_int(int(relabel)), ctypes.c_int(rule), ctypes.c_int(len(k)),
        k.ctypes.data_as(ctypes.c_void_p), tp.ctypes.data_as(ctypes.c_void_p), cmask.ctypes.data_as(ctypes.c_void_p),
        mflat.ctypes.data_as(ctypes.c_void_p), None if st is None else st.ctypes.data_as(ctypes.c_void_p),
        ctypes.byref(npass), final.ctypes.data_as(ctypes.c_void_p), err, ctypes.c_int(256),
    )
    assert rc == 0, err.value.decode()
    if st is None:
        return None, npass.value, final
    idx = np.arange(2**n, dtype=np.int64)
    phys = np.zeros(2**n, dtype=np.int64)
    for q in range(n):
        phys |= ((idx >> q) & 1) << int(final[q])
    return st[phys], npass.value, final


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("n,t,lmin,rb,c64", [(12, 10, 4, 4, True), (14, 10, 5, 5, False), (15, 12, 4, 4, True),
                                              (16, 13, 5, 4, True), (16, 12, 6, 4, False)])
@pytest.mark.parametrize("rule", [0, 1])
def test_relabelled_plan_matches_unfused(emu, name, make, n, t, lmin, rb, c64, rule):
    """Passes that move the qubits needed next into the low positions (exchanged at the tile
    store) give the unfused result once the final layout is undone, and every pass sees the gates
    at the positions the earlier stores left them."""
    cmds = make(n)
    want = oracle_state(cmds, n)
    got, npass, final = run_emu_relabel(emu, n, t, lmin, cmds, start_state(n), rb, c64, True, rule)
    assert sorted(final.tolist()) == list(range(n))
    assert np.max(np.abs(got - want)) <= 1e-12


def test_relabelling_saves_passes(emu):
    """On a 24-qubit random-layer circuit (depth 40, 12 CNOTs per layer) relabelling at the stores
    plans fewer passes than the same scheduler without it, for both slot rules and with 128- and
    256-byte segments."""
    cmds = circuits.random_layers(24, 40, 12, 5, measure=False)
    for lmin in (4, 5):
        for rule in (0, 1):
            _, plain, _ = run_emu_relabel(emu, 24, 12, lmin, cmds, None, 4, True, False, rule)
            _, relab, _ = run_emu_relabel(emu, 24, 12, lmin, cmds, None, 4, True, True, rule)
            assert relab < 0.9 * plain, (lmin, rule, relab, plain)


@pytest.fixture
def emu_cz(emu):
    """The emulator planning the gate list the engine's window planners see: CNOTs between
    1-qubit gates on their target rewritten as H, controlled Z, H (cnot_to_cz)."""
    emu.emu_set_cz(1)
    yield emu
    emu.emu_set_cz(0)


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("n,t,lmin,rb,c64", [(12, 10, 4, 4, True), (14, 10, 5, 5, False), (15, 12, 4, 4, True),
                                              (16, 13, 5, 4, True)])
@pytest.mark.parametrize("rule", [0, 1])
def test_cnots_as_controlled_z_match_unfused(emu_cz, name, make, n, t, lmin, rb, c64, rule):
    """With CNOTs rewritten as H, CZ, H (the H's merged into the neighbouring 1-qubit gates) the
    relabelling tile plan still gives the unfused result."""
    cmds = make(n)
    want = oracle_state(cmds, n)
    got, _, _ = run_emu_relabel(emu_cz, n, t, lmin, cmds, start_state(n), rb, c64, True, rule)
    assert np.max(np.abs(got - want)) <= 1e-12


def test_cnots_as_controlled_z_replace_register_exchanges(emu, emu_cz):
    """On a 24-qubit random-layer circuit most CNOTs sit between 1-qubit gates on their target: the
    rewrite turns a share of the X ops (register exchanges) into Z ops (sign flips in place), and
    CNOTs with a neighbour that is no uncontrolled 1-qubit gate on the target stay X ops."""
    cmds = circuits.random_layers(24, 20, 12, 7, measure=False)
    emu.emu_set_cz(0)
    plain0, lane0 = op_kind_counts(emu, 24, 12, 4, 4, cmds, True)
    emu.emu_set_cz(1)
    plain1, lane1 = op_kind_counts(emu, 24, 12, 4, 4, cmds, True)
    x0, x1 = plain0[3] + lane0[3], plain1[3] + lane1[3]
    z0, z1 = plain0[7] + lane0[7], plain1[7] + lane1[7]
    assert z0 == 0 and z1 > 0.25 * x0 and x1 + z1 == x0, (x0, x1, z1)
    # CX(0 -> 1) stays an X op (qubit 1 is next a control); CX(1 -> 2) sits between H and Rz on
    # qubit 2 and becomes a Z op
    n = 6
    chain = circuits.allocate(n) + [commands.gate(G.H, [q]) for q in range(n)] + [
        commands.gate(G.X, [1], [0]), commands.gate(G.X, [2], [1]), commands.gate(G.Rz(0.3), [2])]
    p, l = op_kind_counts(emu, n, 6, 4, 4, chain, True)
    assert p[3] + l[3] == 1 and p[7] + l[7] == 1
    # two CNOTs onto qubit 2 with nothing else on it in between, between H and Rz: both become Z
    # ops under one pair of H's
    run = circuits.allocate(n) + [commands.gate(G.H, [q]) for q in range(n)] + [
        commands.gate(G.X, [2], [0]), commands.gate(G.Ry(0.2), [4]), commands.gate(G.X, [2], [1]),
        commands.gate(G.Rz(0.3), [2])]
    p, l = op_kind_counts(emu, n, 6, 4, 4, run, True)
    assert p[3] + l[3] == 0 and p[7] + l[7] == 2
    want = oracle_state(run, n)
    got, _ = run_emu(emu, n, 6, 4, run, start_state(n), 4, True)
    assert np.max(np.abs(got - want)) <= 1e-12


def test_random_controlled_blocks(emu):
    rng = np.random.default_rng(4)
    n = 13
    cmds = circuits.allocate(n)
    for _ in range(60):
        k = int(rng.integers(1, 6))
        picks = rng.choice(n, size=k + int(rng.integers(0, 4)), replace=False)
        a = rng.normal(size=(2**k, 2**k)) + 1j * rng.normal(size=(2**k, 2**k))
        u, _ = np.linalg.qr(a)
        cmds.append(commands.gate(u, [int(q) for q in picks[:k]], [int(q) for q in picks[k:]]))
    want = oracle_state(cmds, n)
    for rb in (4, 5):
        for c64 in (True, False):
            got, _ = run_emu(emu, n, 10, 4, cmds, start_state(n), rb, c64)
            assert np.max(np.abs(got - want)) <= 1e-12


def test_pass_count_random32_layer(emu):
    """The planner packs a 32-qubit random-circuit layer (80 gates) into a handful of passes;
    counted with a 20-qubit stand-in of the same shape (t=13, L=5) to stay small on the CPU."""
    n = 20
    cmds = circuits.random_layers(n, 4, 10, 9, measure=False)
    got, npass = run_emu(emu, n, 13, 5, cmds, start_state(n))
    want = oracle_state(cmds, n)
    assert np.max(np.abs(got - want)) <= 1e-12
    assert npass <= 4 * 3


def plan_stage_bits(lib, n, t, lmin, rb, cmds, c64):
    ks, tpos, cm, mats = [], [], [], []
    for c in cmds:
        if c[0] != "gate":
            continue
        _, m, targets, controls = c
        ks.append(len(targets))
        tpos.extend(targets)
        mask = 0
        for q in controls:
            mask |= 1 << q
        cm.append(mask)
        mats.append(np.asarray(m, dtype=np.complex128).reshape(-1))
    k = np.asarray(ks, dtype=np.uint32)
    tp = np.asarray(tpos, dtype=np.uint32)
    cmask = np.asarray(cm, dtype=np.uint64)
    mflat = np.ascontiguousarray(np.concatenate(mats))
    stats = np.zeros(5, dtype=np.int32)
    bits = np.zeros(200000, dtype=np.uint32)
    err = ctypes.create_string_buffer(256)
    rc = lib.emu_plan_stats(
        ctypes.c_uint(n), ctypes.c_uint(t), ctypes.c_uint(lmin), ctypes.c_uint(rb), ctypes.c_int(int(c64)),
        ctypes.c_int(len(ks)), k.ctypes.data_as(ctypes.c_void_p), tp.ctypes.data_as(ctypes.c_void_p),
        cmask.ctypes.data_as(ctypes.c_void_p), mflat.ctypes.data_as(ctypes.c_void_p),
        stats.ctypes.data_as(ctypes.c_void_p), bits.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(len(bits)), err,
        ctypes.c_int(256),
    )
    assert rc == 0, err.value.decode()
    return stats, [tuple(int(b) for b in bits[i * rb:(i + 1) * rb]) for i in range(stats[1])]


@pytest.mark.parametrize("c64", [True, False])
def test_lds_swizzle_is_a_linear_bijection(emu, c64):
    emu.emu_slot.restype = ctypes.c_uint32
    slot = lambda x: emu.emu_slot(ctypes.c_uint32(x), ctypes.c_int(int(c64)))  # noqa: E731
    image = [slot(x) for x in range(1 << 13)]
    assert sorted(image) == list(range(1 << 13))
    rng = np.random.default_rng(7)
    for _ in range(300):
        a = int(rng.integers(1 << 13))
        b = int(rng.integers(1 << 13)) & ~a
        assert slot(a | b) == slot(a) ^ slot(b)
    if c64:  # pairs (2j, 2j + 1) stay adjacent: tiles stream 16-byte pairs as two 8-byte halves
        assert all(slot(2 * j) ^ slot(2 * j + 1) == 1 for j in range(1 << 12))


def bank_conflict_degree(slot, stage, t, lanes_per_cycle, n_banks):
    """Brute force: serve the 256 threads of a register stage lanes_per_cycle at a time and count
    the worst number of lanes that hit one bank, averaged over accesses."""
    nonstage = [b for b in range(t) if b not in stage]
    total = count = 0
    for i in range(1 << len(stage)):
        roff = sum(1 << stage[b] for b in range(len(stage)) if (i >> b) & 1)
        for g0 in range(0, 256, lanes_per_cycle):
            hits = {}
            for g in range(g0, g0 + lanes_per_cycle):
                lb = sum(1 << nonstage[k] for k in range(len(nonstage)) if (g >> k) & 1)
                bank = slot(lb | roff) % n_banks
                hits[bank] = hits.get(bank, 0) + 1
            total += max(hits.values())
            count += 1
    return total / count


def test_register_stages_spread_over_lds_banks(emu):
    """The planner's register stages for 20 layers of the 32-qubit random circuit (t=12, RB=4,
    complex64) read and write LDS with few bank conflicts.  gfx950 LDS serves 32 lanes of 8-byte
    accesses per cycle over 32 bank pairs (64 banks of 4 bytes); brute force over every access
    gives the mean worst-case lanes per bank pair.  The swizzle this one replaced: 1.76."""
    emu.emu_slot.restype = ctypes.c_uint32
    slot = lambda x: emu.emu_slot(ctypes.c_uint32(x), ctypes.c_int(1))  # noqa: E731
    cmds = circuits.random_layers(32, 20, 16, 2024, measure=False)
    stats, stages = plan_stage_bits(emu, 32, 12, 4, 4, cmds, True)
    assert stats[1] == len(stages) > 0
    degrees = {s: bank_conflict_degree(slot, s, 12, 32, 32) for s in set(stages)}
    mean = sum(degrees[s] for s in stages) / len(stages)
    assert mean <= 1.15, mean


def lane_controlled_mix(n, n_gates, seed):
    """Controlled gates of every register-op kind (general, real, diagonal, anti-diagonal, X, Z,
    2-qubit general, 2-qubit diagonal) with 1-2 controls on random qubits, so that many controls
    fall on tile bits outside a stage's register bits."""
    rng = np.random.default_rng(seed)
    makers = [
        lambda a: (G.Rx(a), 1), lambda a: (G.H, 1), lambda a: (G.Ry(a), 1), lambda a: (G.Rz(a), 1), lambda a: (G.R(a), 1),
        lambda a: (G.Y, 1), lambda a: (G.X, 1), lambda a: (G.Rxx(a), 2), lambda a: (G.Rzz(a), 2),
        lambda a: (G.Z, 1),
    ]
    cmds = circuits.allocate(n) + [commands.gate(G.H, [q]) for q in range(n)]
    for _ in range(n_gates):
        m, k = makers[int(rng.integers(len(makers)))](float(rng.uniform(0, 4 * np.pi)))
        picks = [int(q) for q in rng.choice(n, size=k + int(rng.integers(1, 3)), replace=False)]
        cmds.append(commands.gate(m, picks[:k], picks[k:]))
    return cmds


def op_kind_counts(lib, n, t, lmin, rb, cmds, c64):
    ks, tpos, cm, mats = [], [], [], []
    for c in cmds:
        if c[0] != "gate":
            continue
        ks.append(len(c[2]))
        tpos.extend(c[2])
        cm.append(sum(1 << q for q in c[3]))
        mats.append(np.asarray(c[1], dtype=np.complex128).reshape(-1))
    k = np.asarray(ks, dtype=np.uint32)
    tp = np.asarray(tpos, dtype=np.uint32)
    cmask = np.asarray(cm, dtype=np.uint64)
    mflat = np.ascontiguousarray(np.concatenate(mats))
    counts = np.zeros(16, dtype=np.int32)
    err = ctypes.create_string_buffer(256)
    rc = lib.emu_op_kinds(
        ctypes.c_uint(n), ctypes.c_uint(t), ctypes.c_uint(lmin), ctypes.c_uint(rb), ctypes.c_int(int(c64)),
        ctypes.c_int(len(ks)), k.ctypes.data_as(ctypes.c_void_p), tp.ctypes.data_as(ctypes.c_void_p),
        cmask.ctypes.data_as(ctypes.c_void_p), mflat.ctypes.data_as(ctypes.c_void_p),
        counts.ctypes.data_as(ctypes.c_void_p), err, ctypes.c_int(256),
    )
    assert rc == 0, err.value.decode()
    return counts[:8], counts[8:]


LANE_MIX = dict(n=15, n_gates=400, seed=21)


@pytest.mark.parametrize("c64", [True, False])
def test_lane_controlled_mix_covers_every_predicated_op_kind(emu, c64):
    """The circuit of the GPU test below plans (t=12, RB=4, L=4) into register ops of all eight
    kinds both with and without tile-bit controls, so that test runs every plain and every
    predicated op form of the kernel; the CPU replay of the plan matches the unfused oracle."""
    cmds = lane_controlled_mix(**LANE_MIX)
    plain, lane = op_kind_counts(emu, LANE_MIX["n"], 12, 4, 4, cmds, c64)
    assert (lane > 0).all(), lane
    assert (plain > 0).all(), plain
    want = oracle_state(cmds, LANE_MIX["n"])
    got, _ = run_emu(emu, LANE_MIX["n"], 12, 4, cmds, start_state(LANE_MIX["n"]), 4, c64)
    assert np.max(np.abs(got - want)) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_lane_controlled_mix_gpu(precision):
    """Every register-op kind with and without tile-bit (lane) controls on the GPU tile pass."""
    from projectq_amd import Simulator

    n = LANE_MIX["n"]
    cmds = lane_controlled_mix(**LANE_MIX)
    want = oracle_state(cmds, n)
    sim = Simulator(5, precision=precision, fusion="window", planner="tile")
    sim.set_tile_params(12, 4, 4)
    commands.execute(sim, cmds, gate_fusion=True)
    sim.run()
    got = sim.cheat_numpy()[1]
    tol = 1e-12 if precision == "c128" else 2e-5
    assert np.max(np.abs(got - want)) <= tol


def diagonal_mix(n, n_gates, seed):
    """Diagonal gates (Rz, phase, controlled phase, Rzz) interleaved with gates that do not commute
    with them (H, Rx, X, CNOT with the diagonal gates' qubits as targets): the scheduler may move
    diagonal gates and controls past each other but never past a write to their qubits."""
    rng = np.random.default_rng(seed)
    cmds = circuits.allocate(n)
    for _ in range(n_gates):
        r = rng.random()
        a = float(rng.uniform(0, 4 * np.pi))
        q = [int(x) for x in rng.choice(n, size=3, replace=False)]
        if r < 0.2:
            cmds.append(commands.gate(G.Rz(a), [q[0]]))
        elif r < 0.35:
            cmds.append(commands.gate(G.R(a), [q[0]], [q[1]]))
        elif r < 0.45:
            cmds.append(commands.gate(G.Rzz(a), [q[0], q[1]], [q[2]] if rng.random() < 0.3 else []))
        elif r < 0.6:
            cmds.append(commands.gate(G.X, [q[0]], [q[1]]))
        elif r < 0.75:
            cmds.append(commands.gate(G.H, [q[0]]))
        elif r < 0.9:
            cmds.append(commands.gate(G.Rx(a), [q[0]], [q[1]] if rng.random() < 0.5 else []))
        else:
            cmds.append(commands.gate(G.Z, [q[0]], [q[1], q[2]]))
    return cmds


@pytest.mark.parametrize("c64", [True, False])
@pytest.mark.parametrize("n,t,lmin,rb", [(14, 10, 4, 4), (15, 12, 4, 4), (13, 12, 5, 5)])
def test_diagonal_gates_reorder_correctly(emu, c64, n, t, lmin, rb):
    cmds = diagonal_mix(n, 500, 100 + n)
    want = oracle_state(cmds, n)
    start = np.zeros(2**n, dtype=np.complex128)
    start[:] = 1 / np.sqrt(2**n)  # a uniform superposition: every phase is visible
    o = OracleSimulator(1)
    for q in range(n):
        o.allocate_qubit(q)
    o.set_wavefunction(start, list(range(n)))
    commands.execute(o, [c for c in cmds if c[0] == "gate"], gate_fusion=False)
    want = o.cheat_numpy()[1]
    got, npass = run_emu(emu, n, t, lmin, cmds, start, rb, c64)
    assert np.max(np.abs(got - want)) <= 1e-12
    assert npass < 200


def _gate_arrays(cmds):
    ks, tpos, cm, mats = [], [], [], []
    for c in cmds:
        if c[0] != "gate":
            continue
        _, m, targets, controls = c
        ks.append(len(targets))
        tpos.extend(targets)
        mask = 0
        for q in controls:
            mask |= 1 << q
        cm.append(mask)
        mats.append(np.asarray(m, dtype=np.complex128).reshape(-1))
    return (np.asarray(ks, dtype=np.uint32), np.asarray(tpos, dtype=np.uint32), np.asarray(cm, dtype=np.uint64),
            np.ascontiguousarray(np.concatenate(mats)))


def dense_gates(n, count, k, low, seed):
    """`count` random dense k-qubit unitaries on qubits below `low` (all inside one tile)."""
    rng = np.random.default_rng(seed)
    cmds = circuits.allocate(n)
    for _ in range(count):
        picks = rng.choice(low, size=k, replace=False)
        a = rng.normal(size=(2**k, 2**k)) + 1j * rng.normal(size=(2**k, 2**k))
        u, _ = np.linalg.qr(a)
        cmds.append(commands.gate(u, [int(q) for q in picks]))
    return cmds


ARENA_BYTES = 8 << 20  # the engine's upload arena (engine.hpp kArenaBytes)


@pytest.mark.parametrize("k,count", [(4, 3000), (5, 3000)])
def test_pass_upload_fits_the_arena(emu, k, count):
    """Thousands of dense 4- and 5-qubit gates that all fit one 12-bit tile: the scheduler closes
    each pass at its upload budget, so every pass program fits the engine's upload arena (a single
    pass would need 11.7 MiB / 46.9 MiB of matrices)."""
    cmds = dense_gates(20, count, k, 12, 30 + k)
    ks, tp, cm, mf = _gate_arrays(cmds)
    most, npass = ctypes.c_uint64(), ctypes.c_int()
    err = ctypes.create_string_buffer(256)
    emu.emu_max_pass_upload.restype = ctypes.c_int
    rc = emu.emu_max_pass_upload(
        ctypes.c_uint(20), ctypes.c_uint(12), ctypes.c_uint(5), ctypes.c_uint(4), ctypes.c_int(0), ctypes.c_int(len(ks)),
        ks.ctypes.data_as(ctypes.c_void_p), tp.ctypes.data_as(ctypes.c_void_p), cm.ctypes.data_as(ctypes.c_void_p),
        mf.ctypes.data_as(ctypes.c_void_p), ctypes.byref(most), ctypes.byref(npass), err, ctypes.c_int(256))
    assert rc == 0, err.value.decode()
    total = count * 16 * 4**k
    assert most.value <= ARENA_BYTES
    assert npass.value >= total // ARENA_BYTES + 1


def test_split_passes_match_unfused(emu):
    """Passes closed by the upload budget (1100 dense 5-qubit gates in one tile: at least five
    passes) still give the unfused result."""
    n = 12
    cmds = dense_gates(n, 1100, 5, n, 77)
    want = oracle_state(cmds, n)
    got, npass = run_emu(emu, n, 12, 5, cmds, start_state(n), 4, False)
    assert np.max(np.abs(got - want)) <= 1e-12
    assert npass >= 5
