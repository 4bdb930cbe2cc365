"""MI355X-native MLS-MPM backend ("hip"): ctypes host layer over libmpm_hip.so."""

from mpm_core.backends.hip.solver import HipSolver

__all__ = ["HipSolver"]
