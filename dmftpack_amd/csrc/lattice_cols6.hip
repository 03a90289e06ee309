// Column-per-lane Schur kernels, bucket N = 6 (n_c up to 6); see lattice_schur_cols.hpp.
#include "lattice_schur_cols.hpp"

namespace dg {

hipError_t launch_cols_6(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s)
{
    return launch_cols_bucket<6>(a, mode, fp32, sig_all, s);
}

}  // namespace dg
