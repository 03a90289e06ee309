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

// Rank-1 update c_i -= b_i t of one Gauss-Jordan step where b = lane J's
// column, taken by the FMA itself: v_fmac_f64 with a DPP64 row_newbcast:J
// source operand, so no separate broadcast move.  Per element the four FMAs
// follow cfms's rounding order,
//   A  x += b.y * t.y     B  y += b.y * (-t.x)
//   C  x += b.x * (-t.x)  D  y += b.x * (-t.y)
// issued phase by phase over a group of elements.  A DPP operand must not
// have been written by either of the two preceding VALU instructions; in this
// order every operand read by DPP was written at least two instructions
// earlier, and "s_nop 1" at both ends of a block covers the compiler's own
// instructions on either side, whose hazard tracking does not look inside the
// asm.  The pivot lane has t = 0, so its registers -- the DPP sources of the
// whole slot -- keep their values through the four phases.
template <int J>
__device__ __forceinline__ void fms_bcast4(cd& c0, cd& c1, cd& c2, cd& c3, double ty, double mtx, double mty)
{
    asm("s_nop 1\n\t"
        "v_fmac_f64_dpp %0, %1, %8 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %2, %3, %8 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %4, %5, %8 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %6, %7, %8 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %1, %1, %9 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %3, %3, %9 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %5, %5, %9 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %7, %7, %9 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %0, %0, %9 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %2, %2, %9 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %4, %4, %9 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %6, %6, %9 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %1, %0, %10 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %3, %2, %10 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %5, %4, %10 row_newbcast:%c11\n\t"
        "v_fmac_f64_dpp %7, %6, %10 row_newbcast:%c11\n\t"
        "s_nop 1"
        : "+v"(c0.x), "+v"(c0.y), "+v"(c1.x), "+v"(c1.y), "+v"(c2.x), "+v"(c2.y), "+v"(c3.x), "+v"(c3.y)
        : "v"(ty), "v"(mtx), "v"(mty), "n"(J));
}

template <int J>
__device__ __forceinline__ void fms_bcast3(cd& c0, cd& c1, cd& c2, double ty, double mtx, double mty)
{
    asm("s_nop 1\n\t"
        "v_fmac_f64_dpp %0, %1, %6 row_newbcast:%c9\n\t"
        "v_fmac_f64_dpp %2, %3, %6 row_newbcast:%c9\n\t"
        "v_fmac_f64_dpp %4, %5, %6 row_newbcast:%c9\n\t"
        "v_fmac_f64_dpp %1, %1, %7 row_newbcast:%c9\n\t"
        "v_fmac_f64_dpp %3, %3, %7 row_newbcast:%c9\n\t"
        "v_fmac_f64_dpp %5, %5, %7 row_newbcast:%c9\n\t"
        "v_fmac_f64_dpp %0, %0, %7 row_newbcast:%c9\n\t"
        "v_fmac_f64_dpp %2, %2, %7 row_newbcast:%c9\n\t"
        "v_fmac_f64_dpp %4, %4, %7 row_newbcast:%c9\n\t"
        "v_fmac_f64_dpp %1, %0, %8 row_newbcast:%c9\n\t"
        "v_fmac_f64_dpp %3, %2, %8 row_newbcast:%c9\n\t"
        "v_fmac_f64_dpp %5, %4, %8 row_newbcast:%c9\n\t"
        "s_nop 1"
        : "+v"(c0.x), "+v"(c0.y), "+v"(c1.x), "+v"(c1.y), "+v"(c2.x), "+v"(c2.y)
        : "v"(ty), "v"(mtx), "v"(mty), "n"(J));
}

// two elements: one wait state between C1 and D0 is padded to two
template <int J>
__device__ __forceinline__ void fms_bcast2(cd& c0, cd& c1, double ty, double mtx, double mty)
{
    asm("s_nop 1\n\t"
        "v_fmac_f64_dpp %0, %1, %4 row_newbcast:%c7\n\t"
        "v_fmac_f64_dpp %2, %3, %4 row_newbcast:%c7\n\t"
        "v_fmac_f64_dpp %1, %1, %5 row_newbcast:%c7\n\t"
        "v_fmac_f64_dpp %3, %3, %5 row_newbcast:%c7\n\t"
        "v_fmac_f64_dpp %0, %0, %5 row_newbcast:%c7\n\t"
        "v_fmac_f64_dpp %2, %2, %5 row_newbcast:%c7\n\t"
        "s_nop 0\n\t"
        "v_fmac_f64_dpp %1, %0, %6 row_newbcast:%c7\n\t"
        "v_fmac_f64_dpp %3, %2, %6 row_newbcast:%c7\n\t"
        "s_nop 1"
        : "+v"(c0.x), "+v"(c0.y), "+v"(c1.x), "+v"(c1.y)
        : "v"(ty), "v"(mtx), "v"(mty), "n"(J));
}

template <int J>
__device__ __forceinline__ void fms_bcast1(cd& c0, double ty, double mtx, double mty)
{
    asm("s_nop 1\n\t"
        "v_fmac_f64_dpp %0, %1, %2 row_newbcast:%c5\n\t"
        "v_fmac_f64_dpp %1, %1, %3 row_newbcast:%c5\n\t"
        "s_nop 0\n\t"
        "v_fmac_f64_dpp %0, %0, %3 row_newbcast:%c5\n\t"
        "s_nop 1\n\t"
        "v_fmac_f64_dpp %1, %0, %4 row_newbcast:%c5\n\t"
        "s_nop 1"
        : "+v"(c0.x), "+v"(c0.y)
        : "v"(ty), "v"(mtx), "v"(mty), "n"(J));
}

// acc += s_M w, where s_M is lane M's value of s (DPP64 row_newbcast:M) and w
// is the lane's own; cfma's rounding order.  Nothing the block writes is read
// through DPP, so only the leading s_nop is needed.
template <int M>
__device__ __forceinline__ void cfma_bcast(cd& acc, const cd& s, const cd& w, double mwy)
{
    asm("s_nop 1\n\t"
        "v_fmac_f64_dpp %0, %3, %6 row_newbcast:%c7\n\t"
        "v_fmac_f64_dpp %1, %3, %4 row_newbcast:%c7\n\t"
        "v_fmac_f64_dpp %0, %2, %4 row_newbcast:%c7\n\t"
        "v_fmac_f64_dpp %1, %2, %5 row_newbcast:%c7"
        : "+v"(acc.x), "+v"(acc.y)
        : "v"(s.x), "v"(s.y), "v"(w.x), "v"(w.y), "v"(mwy), "n"(M));
}

// c_i -= b_i t for all i < N (b = lane J's column): groups of four, then the rest
template <int J, int N>
__device__ __forceinline__ void fms_bcast_all(cd (&col)[N], double ty, double mtx, double mty)
{
    constexpr int Q = N / 4, R = N % 4;
#pragma unroll
    for (int g = 0; g < Q; ++g) fms_bcast4<J>(col[4 * g], col[4 * g + 1], col[4 * g + 2], col[4 * g + 3], ty, mtx, mty);
    if constexpr (R == 3) fms_bcast3<J>(col[4 * Q], col[4 * Q + 1], col[4 * Q + 2], ty, mtx, mty);
    if constexpr (R == 2) fms_bcast2<J>(col[4 * Q], col[4 * Q + 1], ty, mtx, mty);
    if constexpr (R == 1) fms_bcast1<J>(col[4 * Q], ty, mtx, mty);
}

// Steps J .. N-1 of the row-slot elimination; the step index is a template
// argument (it is the DPP control immediate), so the recursion replaces an
// unrolled loop.
template <int N, int J>
__device__ __forceinline__ void gj_row16_steps(cd (&col)[N], cd& scl, int c, int n)
{
    if constexpr (J < N) {
        if (J < n) {
            const cd bj = cd{row_bcast<J>(col[J].x), row_bcast<J>(col[J].y)};
            const cd p = crcp(bj);
            const bool isj = (c == J);
            const cd t = isj ? cd{0.0, 0.0} : cmul(col[J], p);
            fms_bcast_all<J, N>(col, t.y, -t.x, -t.y);
            col[J] = isj ? cd{-1.0, 0.0} : t;
            scl = isj ? cneg(p) : scl;
            gj_row16_steps<N, J + 1>(col, scl, c, n);
        }
    }
}

// The same elimination for a slot that is one 16-lane row of the wavefront
// (N <= 16, slot lanes 16 s .. 16 s + 15): the pivot column is read from
// lane j's registers by the update FMAs themselves (DPP64 row_newbcast), so a
// step has no LDS round trip, no wait and no broadcast moves.  The update of
// element j is computed like the others and then overwritten.
template <int N>
__device__ __forceinline__ void gj_columns_row16(cd (&col)[N], int c, int n)
{
    static_assert(N <= 16, "a slot of one 16-lane row");
    cd scl = cd{1.0, 0.0};
    gj_row16_steps<N, 0>(col, scl, c, n);
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
