"""MI355X-native state-vector simulator with ProjectQ's `_cppsim.Simulator` surface.

Drop-in use with ProjectQ::

    from projectq.backends import Simulator
    import projectq_amd
    sim = Simulator(gate_fusion=True)
    sim._simulator = projectq_amd.Simulator(seed)
"""

from ._cppsim import Simulator, library  # noqa: F401

__all__ = ["Simulator", "library"]
