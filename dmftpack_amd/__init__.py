"""MI355X-native lattice Green's function engine (DMFTpack's GreenFtn_w /
FourierTransform path) -- HIP kernels for gfx950 behind the C ABI of
include/dmftgpu.h.  Python here is plumbing for tests and benchmarks."""
from . import capi, lattice  # noqa: F401

__all__ = ["capi", "lattice"]
