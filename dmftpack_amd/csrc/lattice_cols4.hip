// Column-per-lane Schur kernels, bucket N = 4 (n_c up to 4); see lattice_schur_cols.hpp.
#include "lattice_schur_cols.hpp"

namespace dg {

hipError_t launch_cols_4(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s)
{
    return launch_cols_bucket<4>(a, mode, fp32, sig_all, s);
}

}  // namespace dg
