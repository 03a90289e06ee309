// Column-per-lane Schur kernels, bucket N = 8 (n_c up to 8); see lattice_schur_cols.hpp.
#include "lattice_schur_cols.hpp"

namespace dg {

hipError_t launch_cols_8(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s)
{
    return launch_cols_bucket<8>(a, mode, fp32, sig_all, s);
}

hipError_t launch_cols_multi_8(const LatArgs& proto, int mode, bool fp32, hipStream_t s, const SchurMulti& m)
{
    return launch_cols_bucket<8>(proto, mode, fp32, nullptr, s, &m);
}

}  // namespace dg
