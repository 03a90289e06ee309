// Column-per-lane Schur kernels, bucket N = 12 (n_c up to 12); see lattice_schur_cols.hpp.
#include "lattice_schur_cols.hpp"

namespace dg {

hipError_t launch_cols_12(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s)
{
    return launch_cols_bucket<12>(a, mode, fp32, sig_all, s);
}

hipError_t launch_cols_multi_12(const LatArgs& proto, int mode, bool fp32, hipStream_t s, const SchurMulti& m)
{
    return launch_cols_bucket<12>(proto, mode, fp32, nullptr, s, &m);
}

}  // namespace dg
