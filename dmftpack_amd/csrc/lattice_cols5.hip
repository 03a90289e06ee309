// Column-per-lane Schur kernels, bucket N = 5 (n_c up to 5); see lattice_schur_cols.hpp.
#include "lattice_schur_cols.hpp"

namespace dg {

hipError_t launch_cols_5(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s)
{
    return launch_cols_bucket<5>(a, mode, fp32, sig_all, s);
}

}  // namespace dg
