// One-matrix-per-lane Schur kernels for n_c = 2 (lattice_schur_lane.hpp).
#include "lattice_schur_lane.hpp"

namespace dg {

hipError_t launch_lane_2(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s, int wg,
                         const SchurMulti* m)
{
    return launch_lane_nc<2>(a, mode, fp32, sig_all, s, wg, m);
}

}  // namespace dg
