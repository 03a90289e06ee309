// Column-per-lane Gauss-Jordan inverse without pivoting, at wavefront level.
//
// A "slot" of lanes owns one N x N matrix (bucketed: only the leading n x n
// block is eliminated, the padding columns/rows hold the identity).  Lane c
// holds column c of the matrix in col[].  At step j lane j publishes its
// column through the slot's LDS buffer bc (2*N entries, double buffered so no
// second wave barrier is needed), and every lane applies the rank-1 update to
// its own column.  The pivot column is not rescaled in place: lane j records
// the factor -1/A_jj and applies it once at the end, which keeps every step
// free of divergent branches.  On return col[] holds column c of A^{-1}.
//
// Elimination without pivoting is backward stable for the matrices inverted
// here: for causal Sigma the Hermitian part of -iA is >= w_n I (positive
// definite), the case covered by Golub & Van Loan sec. 4.4 / Higham sec. 10.4.
#pragma once
#include "dev_common.hpp"

namespace dg {

template <int N>
__device__ __forceinline__ void gj_columns(cd (&col)[N], cd* bc, int c, int n)
{
    cd scl = cd{1.0, 0.0};
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j < n) {
            cd* buf = bc + (j & 1) * N;
            if (c == j) {
#pragma unroll
                for (int i = 0; i < N; ++i) buf[i] = col[i];
            }
            wave_sync();
            cd b[N];
#pragma unroll
            for (int i = 0; i < N; ++i) b[i] = buf[i];
            const cd p = crcp(b[j]);
            const bool isj = (c == j);
            const cd t = isj ? cd{0.0, 0.0} : cmul(col[j], p);
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (i != j) col[i] = cfms(b[i], t, col[i]);
            col[j] = isj ? cd{-1.0, 0.0} : t;
            scl = isj ? cneg(p) : scl;
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) col[i] = cmul(col[i], scl);
}

// FP32 form of the same elimination (the first solve of FP32 + FP64-residual
// refinement).
template <int N>
__device__ __forceinline__ void gj_columns_f32(cf (&col)[N], cf* bc, int c, int n)
{
    cf scl = cf{1.f, 0.f};
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j < n) {
            cf* buf = bc + (j & 1) * N;
            if (c == j) {
#pragma unroll
                for (int i = 0; i < N; ++i) buf[i] = col[i];
            }
            wave_sync();
            cf b[N];
#pragma unroll
            for (int i = 0; i < N; ++i) b[i] = buf[i];
            const cf p = frcp(b[j]);
            const bool isj = (c == j);
            const cf t = isj ? cf{0.f, 0.f} : fmul(col[j], p);
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (i != j) col[i] = ffms(b[i], t, col[i]);
            col[j] = isj ? cf{-1.f, 0.f} : t;
            scl = isj ? cf{-p.x, -p.y} : scl;
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) col[i] = fmul(col[i], scl);
}

}  // namespace dg
