"""avbd_amd -- MI355X-native (gfx950) Augmented Vertex Block Descent solver.

``avbd_amd.sim`` mirrors the reference solver API (Simulation / SimulationContext /
VBDSolver) over the C ABI declared in include/avbd.h; ``avbd_amd.scenes`` builds
the synthetic workloads.  Importing ``avbd_amd.sim`` loads the native library
and raises if it is missing -- there is no CPU fallback.
"""
from . import scenes  # noqa: F401

__all__ = ["scenes"]
