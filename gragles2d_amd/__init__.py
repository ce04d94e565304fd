"""gragles2d_amd -- MI355X-native per-grain level-set timestep of GraGLeS2D.

The product is libggx.so (gragles2d_amd/csrc, C-ABI in include/ggx.h); this package is the thin
host-side mirror of the reference driver over that C-ABI.
"""
from .configs import WORKLOADS, Workload  # noqa: F401
