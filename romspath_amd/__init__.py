"""MI355X implementation of ROMSPath's per-particle time step.

librp.so (built from csrc/ by build.py) holds the HIP kernels and the C-ABI
of include/rp_abi.h; this package binds it (abi), drives it through the
reference's time schedule (schedule, runner) and generates the synthetic
grids and fields the tests and benchmark use (synth).
"""
__all__ = ["abi", "build", "nesting", "runner", "schedule", "synth"]
