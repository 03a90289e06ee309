// Schur-complement lattice path (placeholder until the kernel lands).
#include "lattice_kernels.hpp"

namespace dg {
bool schur_supported(int, int) { return false; }
int schur_slots(int, int) { return 64; }
hipError_t launch_lattice_schur(const LatArgs&, int, bool, hipStream_t) { return hipErrorNotSupported; }
hipError_t launch_schur_prep(const cd*, const cd*, int, int, int, cd*, hipStream_t) { return hipErrorNotSupported; }
}  // namespace dg
