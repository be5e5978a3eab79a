"""CPU checkers for the simulator -- test infrastructure only.

`oracle.cpu_sim` is an in-repo C++ restatement of ProjectQ's complex128 engine exporting the
same C ABI as the HIP library.  `oracle.ref` loads ProjectQ's own simulator compiled from the
reference sources by `make -C oracle ref` (present only where that build ran).  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this package.
"""
