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

// Lane J of each 16-lane row of the wavefront, to every lane of that row
// (64-bit DPP row_newbcast, gfx90a and later): one v_mov_b64_dpp, no LDS.
template <int J>
__device__ __forceinline__ double row_bcast(double x)
{
    return __builtin_amdgcn_update_dpp(0.0, x, 0x150 + J, 0xF, 0xF, false);
}

__device__ __forceinline__ double row_bcast_j(double x, int j)
{
    switch (j) {  // j is a constant after unrolling: the switch folds away
    case 0: return row_bcast<0>(x);
    case 1: return row_bcast<1>(x);
    case 2: return row_bcast<2>(x);
    case 3: return row_bcast<3>(x);
    case 4: return row_bcast<4>(x);
    case 5: return row_bcast<5>(x);
    case 6: return row_bcast<6>(x);
    case 7: return row_bcast<7>(x);
    case 8: return row_bcast<8>(x);
    case 9: return row_bcast<9>(x);
    case 10: return row_bcast<10>(x);
    case 11: return row_bcast<11>(x);
    case 12: return row_bcast<12>(x);
    case 13: return row_bcast<13>(x);
    case 14: return row_bcast<14>(x);
    default: return row_bcast<15>(x);
    }
}

// The same elimination for a slot that is one 16-lane row of the wavefront
// (N <= 16, slot lanes 16 s .. 16 s + 15): the pivot column is broadcast from
// lane j's registers by DPP instead of through LDS, so a step has no LDS
// round trip and no wait.
template <int N>
__device__ __forceinline__ void gj_columns_row16(cd (&col)[N], int c, int n)
{
    static_assert(N <= 16, "a slot of one 16-lane row");
    cd scl = cd{1.0, 0.0};
#pragma unroll
    for (int j = 0; j < N; ++j) {
        if (j < n) {
            cd b[N];
#pragma unroll
            for (int i = 0; i < N; ++i) b[i] = cd{row_bcast_j(col[i].x, j), row_bcast_j(col[i].y, j)};
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
