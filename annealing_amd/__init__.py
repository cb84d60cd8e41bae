"""MI355X-native simulated-bifurcation solver (drop-in for SB/CUDA_GPU of
strongshih/Annealing).  Compute lives in libsbamd.so (HIP, gfx950); this
package is the Python binding of its C ABI."""
from . import sb  # noqa: F401
from .sb import Context, Problem, SBError, load  # noqa: F401

__all__ = ["sb", "Context", "Problem", "SBError", "load"]
