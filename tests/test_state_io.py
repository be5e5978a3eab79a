"""State checkpoints (save_state / load_state), the device view of the state and
set_wavefunction from device memory.

The file format is restated in the CPU oracle (oracle/cpu_sim.cpp), so files cross between the
oracle and the HIP library in both directions: a state written by one and read by the other must
come back bit for bit (complex128) or rounded to nearest (complex64)."""

import os
import struct

import numpy as np
import pytest

from oracle.cpu_sim import OracleSimulator
from projectq_amd import circuits, commands


def prepared_oracle(n=9, seed=4):
    o = OracleSimulator(seed)
    commands.execute(o, circuits.random_layers(n, 4, 3, seed, measure=False), gate_fusion=False)
    # a qubit map that is not the identity: one more qubit with a large id
    o.allocate_qubit(100)
    return o


def test_oracle_state_file_round_trip(tmp_path):
    o = prepared_oracle()
    path = str(tmp_path / "state.qs")
    o.save_state(path)
    with open(path, "rb") as f:
        head = f.read(32)
    magic, version, precision, nq, _, na = struct.unpack("<8sIIIIQ", head)
    assert (magic, version, precision, nq, na) == (b"QSAMDST1", 1, 0, 10, 1024)
    assert os.path.getsize(path) == 32 + 8 * nq + 16 * na
    fresh = OracleSimulator(1)
    fresh.load_state(path)
    assert fresh.qubit_map() == o.qubit_map()
    assert np.array_equal(fresh.cheat_numpy()[1], o.cheat_numpy()[1])
    # a simulator holding other qubit ids refuses the file
    other = OracleSimulator(1)
    other.allocate_qubit(5)
    with pytest.raises(RuntimeError):
        other.load_state(path)


def test_oracle_rejects_malformed_state_files(tmp_path):
    o = prepared_oracle()
    path = str(tmp_path / "state.qs")
    o.save_state(path)
    data = open(path, "rb").read()
    for bad in (b"XXXXXXXX" + data[8:], data[:-16], data[:40], data + b"\0" * 16):
        p = str(tmp_path / "bad.qs")
        with open(p, "wb") as f:
            f.write(bad)
        with pytest.raises(ValueError):
            OracleSimulator(1).load_state(p)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_state_files_cross_between_oracle_and_gpu(tmp_path, precision):
    from projectq_amd import Simulator

    o = prepared_oracle()
    want_map, want = o.qubit_map(), o.cheat_numpy()[1]
    path = str(tmp_path / "oracle.qs")
    o.save_state(path)
    h = Simulator(1, precision=precision)
    h.load_state(path)
    assert h.qubit_map() == want_map
    got = h.cheat_numpy()[1]
    if precision == "c128":
        assert np.array_equal(got, want)
    else:
        assert np.array_equal(got, want.astype(np.complex64).astype(np.complex128))
    # and back: the GPU's file read by the oracle
    back = str(tmp_path / "gpu.qs")
    h.save_state(back)
    o2 = OracleSimulator(1)
    o2.load_state(back)
    assert o2.qubit_map() == want_map
    assert np.array_equal(o2.cheat_numpy()[1], got)


@pytest.mark.gpu
def test_state_file_round_trip_large_gpu(tmp_path):
    """A 26-qubit complex64 state (512 MiB) crosses several 64 MiB staging buffers."""
    from projectq_amd import Simulator

    n = 26
    h = Simulator(3, precision="c64", fusion="window", planner="tile")
    commands.execute(h, circuits.random_layers(n, 2, n // 2, 6, measure=False), gate_fusion=True)
    path = str(tmp_path / "big.qs")
    h.save_state(path)
    assert os.path.getsize(path) == 32 + 8 * n + 8 * 2**n
    g = Simulator(9, precision="c64")
    g.load_state(path)
    assert abs(g.overlap(h) - 1.0) <= 1e-6
    rng = np.random.default_rng(0)
    for idx in rng.integers(0, 2**n, size=32):
        bits = [(int(idx) >> q) & 1 for q in range(n)]
        assert g.get_amplitude(bits, list(range(n))) == h.get_amplitude(bits, list(range(n)))


@pytest.mark.gpu
def test_state_handed_out_after_relabelling_windows(tmp_path):
    """Tile passes that relabel qubits leave the state in another layout; the checkpoint, the qubit
    map and the device state are in the reference's layout (checked against the oracle), and a
    checkpoint written after relabelling loads back into a simulator whose next window relabels
    again."""
    from projectq_amd import Simulator

    n = 18
    cmds = circuits.random_layers(n, 10, 9, 12, measure=False)
    o = OracleSimulator(4)
    commands.execute(o, cmds, gate_fusion=False)
    h = Simulator(4, precision="c128", fusion="window", planner="tile", capacity_qubits=n)
    h.set_tile_params(11, 4, 4)
    h.set_relabel("on")
    commands.execute(h, cmds, gate_fusion=True)
    h.run()
    assert h.last_plan()["relabels"] > 0, h.last_plan()
    path = str(tmp_path / "relabelled.qsst")
    h.save_state(path)
    mapping, view = h.device_state()
    assert view.n_amps == 2**n and not view.c64
    assert mapping == {q: q for q in range(n)} == h.qubit_map()  # the reference's layout
    want = o.cheat_numpy()[1]
    assert np.max(np.abs(h.cheat_numpy()[1] - want)) <= 1e-12
    g = Simulator(4, precision="c128", fusion="window", planner="tile", capacity_qubits=n)
    g.set_tile_params(11, 4, 4)
    g.set_relabel("on")
    g.load_state(path)
    assert np.max(np.abs(g.cheat_numpy()[1] - want)) <= 1e-12
    more = circuits.random_layers(n, 3, 9, 13, measure=False)[n:]
    commands.execute(o, more, gate_fusion=False)
    commands.execute(g, more, gate_fusion=True)
    g.run()
    assert np.max(np.abs(g.cheat_numpy()[1] - o.cheat_numpy()[1])) <= 1e-12


@pytest.mark.gpu
def test_malformed_state_file_leaves_the_gpu_simulator_unchanged(tmp_path):
    """A truncated or padded checkpoint is rejected (ValueError) before the qubit map or the
    state change."""
    from projectq_amd import Simulator

    o = prepared_oracle()
    path = str(tmp_path / "state.qs")
    o.save_state(path)
    data = open(path, "rb").read()
    h = Simulator(1)
    commands.execute(h, circuits.random_layers(9, 3, 4, 9, measure=False), gate_fusion=False)
    h.allocate_qubit(100)  # the file's qubit ids, at other positions
    h_map, h_state = h.qubit_map(), h.cheat_numpy()[1]
    assert set(h_map) == set(o.qubit_map())
    for bad in (data[:-16], data[:-1], data + b"\0" * 16):
        p = str(tmp_path / "bad.qs")
        with open(p, "wb") as f:
            f.write(bad)
        with pytest.raises(ValueError):
            h.load_state(p)
        assert h.qubit_map() == h_map
        assert np.array_equal(h.cheat_numpy()[1], h_state)


@pytest.mark.gpu
@pytest.mark.parametrize("src_precision", ["c128", "c64"])
@pytest.mark.parametrize("dst_precision", ["c128", "c64"])
def test_device_state_and_set_wavefunction_device(src_precision, dst_precision):
    """One simulator's state handed to another through device memory (no host copy)."""
    from projectq_amd import Simulator

    n = 12
    a = Simulator(2, precision=src_precision)
    commands.execute(a, circuits.random_layers(n, 3, 5, 8, measure=False), gate_fusion=False)
    mapping, view = a.device_state()
    assert mapping == a.qubit_map() and view.n_amps == 2**n and view.c64 == (src_precision == "c64")
    assert view.ptr != 0
    host = a.cheat_numpy()[1]
    b = Simulator(5, precision=dst_precision)
    for q in range(n):
        b.allocate_qubit(q)
    order = [q for q, _ in sorted(mapping.items(), key=lambda kv: kv[1])]
    b.set_wavefunction_device(view, order)
    got = b.cheat_numpy()[1]
    want = host.astype(np.complex64).astype(np.complex128) if dst_precision == "c64" else host
    assert np.array_equal(got, want)
    assert b.qubit_map() == mapping


TORCH_CHILD = r"""
import torch  # first: projectq_amd then shares torch's HIP runtime
import numpy as np
from projectq_amd import Simulator, circuits, commands
sim = Simulator(2, precision="c64")
commands.execute(sim, circuits.random_layers(10, 2, 4, 3, measure=False), gate_fusion=False)
mapping, view = sim.device_state()
t = torch.as_tensor(view, device="cuda")
assert t.dtype == torch.complex64 and t.numel() == 1024
assert np.array_equal(t.cpu().numpy(), sim.cheat_numpy()[1].astype(np.complex64))
g = Simulator(4, precision="c128")
for q in range(10):
    g.allocate_qubit(q)
g.set_wavefunction_device((2 * t).to(torch.complex128), sorted(mapping, key=mapping.get))
assert np.array_equal(g.cheat_numpy()[1], 2 * sim.cheat_numpy()[1])
print("torch interop ok")
"""


@pytest.mark.gpu
def test_torch_interop_in_a_process_that_imports_torch_first():
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", TORCH_CHILD], cwd=root, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    assert "torch interop ok" in out.stdout
