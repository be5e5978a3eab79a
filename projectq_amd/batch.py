"""Batched variational circuits on one GPU (extension of the simulator; no reference counterpart).

    from projectq_amd.batch import BatchExpectation
    b = BatchExpectation([(n, gate_commands), ...], [operator, ...], precision="c64")
    energies, gpu_ms = b.run()          # one hipGraph launch for the whole batch

Each circuit starts from |0...0> on its own n-qubit register; energies[i] = <psi_i|H_i|psi_i>.
The tile shape is autotuned on the batch when `autotune` is set.
"""

from ._binding import CAbiBatch
from ._cppsim import library


class BatchExpectation(CAbiBatch):
    def __init__(self, circuits, operators, precision="c64", autotune=True):
        super().__init__(library(), circuits, operators, precision, autotune)
