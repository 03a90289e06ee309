// Schur-complement lattice path: G_cc without the full n x n inverse.
//
// Same outputs as the plain path (the hot loop of GreenFtn_w,
// GreenFtn.cpp:218-308, and the densities of GiwToNele :513-635), computed
// from the block structure of A = (z+mu) I - H(k) - P Sigma P^H with the
// correlated orbitals first (downfolding.cpp:69-79):
//   H_dd(k) = V_d diag(e_d) V_d^H           (once per k, in dmftgpu_prepare)
//   W(k)    = H_cd V_d,  P_d = w_d w_d^H    (w_d = column d of W)
//   g_d     = 1/(z + mu - e_d)
//   S       = (z+mu) I - H_cc - Sigma - sum_d g_d P_d        (Schur complement)
//   G_cc    = S^{-1}
//   tr G    = sum_d g_d + tr G_cc + sum_d g_d^2 M_d,   M_d = w_d^H G_cc w_d
//   tr G^2  = sum_d g_d^2 + tr((G_cc S')^2) + 2 sum_d g_d^3 M_d,
//             S' = I + X, X = sum_d g_d^2 P_d
//   (G^2)_cc = G_cc S' G_cc
// (tr G = d/dz log det A = d/dz[log det A_dd + log det S]; tr G^2 = -d/dz tr G.)
// Per (k, iw) this costs ~ n_c^2 n_d + n_c^3 complex operations instead of n^3.
//
// This header holds the kernel templates; lattice_lane{1..6}.hip instantiate
// one n_c each (launch_lane_<n_c>), lattice_schur.hip dispatches.
//
// Mapping: one (k, iw) matrix per lane, everything in registers (n_c <= 6):
// a wavefront owns 64 consecutive frequencies and walks a chunk of k points;
// the per-k data (H_cc, e_d, P_d for pairs a <= b; built once by k_schur_kdata)
// is read with wave-uniform scalar loads.  The n_c x n_c inversion is
// Gauss-Jordan without pivoting (the Hermitian part of -iS is >= w_n I for
// causal Sigma, so elimination without pivoting is stable).
//
// FP32R: the Schur complement S is formed in FP64, inverted in FP32, and the
// FP32 inverse X0 preconditions iterative refinement with FP64 residuals
//   X = X0 + D,  D <- D + X0 (I - S X),
// `refine` times; the correction D is kept in FP32.
#pragma once
#include "lattice_kernels.hpp"

namespace dg {

template <int NC>
__device__ __forceinline__ void gj_inverse(cd (&S)[NC][NC])
{
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const cd p = crcp(S[j][j]);
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (c != j) S[j][c] = cmul(S[j][c], p);
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            if (i == j) continue;
            const cd f = S[i][j];
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (c != j) S[i][c] = cfms(f, S[j][c], S[i][c]);
            S[i][j] = cneg(cmul(f, p));
        }
        S[j][j] = p;
    }
}

template <int NC>
__device__ __forceinline__ void gj_inverse_f32(cf (&S)[NC][NC])
{
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const cf p = frcp(S[j][j]);
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (c != j) S[j][c] = fmul(S[j][c], p);
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            if (i == j) continue;
            const cf f = S[i][j];
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (c != j) S[i][c] = ffms(f, S[j][c], S[i][c]);
            const cf fp = fmul(f, p);
            S[i][j] = cf{-fp.x, -fp.y};
        }
        S[j][j] = p;
    }
}

// Scalar-cache fill of the next k's block before its d loops need it: one
// dword from each of the first 48 64-byte lines (3 KB; the per-k blocks are
// followed by kKdataPad bytes of padding, so the last block's fill stays in
// the allocation), all into one throw-away SGPR, waited for inside the asm
// block so that no load is in flight when it ends.  Without it every d
// block's scalar loads wait on L2 in turn; with it they hit the scalar cache.
__device__ __forceinline__ void kcache_fill(const void* base)
{
    unsigned t;
    asm volatile(
        "s_load_dword %0, %1, 0x0\n\t"
        "s_load_dword %0, %1, 0x40\n\t"
        "s_load_dword %0, %1, 0x80\n\t"
        "s_load_dword %0, %1, 0xc0\n\t"
        "s_load_dword %0, %1, 0x100\n\t"
        "s_load_dword %0, %1, 0x140\n\t"
        "s_load_dword %0, %1, 0x180\n\t"
        "s_load_dword %0, %1, 0x1c0\n\t"
        "s_load_dword %0, %1, 0x200\n\t"
        "s_load_dword %0, %1, 0x240\n\t"
        "s_load_dword %0, %1, 0x280\n\t"
        "s_load_dword %0, %1, 0x2c0\n\t"
        "s_load_dword %0, %1, 0x300\n\t"
        "s_load_dword %0, %1, 0x340\n\t"
        "s_load_dword %0, %1, 0x380\n\t"
        "s_load_dword %0, %1, 0x3c0\n\t"
        "s_load_dword %0, %1, 0x400\n\t"
        "s_load_dword %0, %1, 0x440\n\t"
        "s_load_dword %0, %1, 0x480\n\t"
        "s_load_dword %0, %1, 0x4c0\n\t"
        "s_load_dword %0, %1, 0x500\n\t"
        "s_load_dword %0, %1, 0x540\n\t"
        "s_load_dword %0, %1, 0x580\n\t"
        "s_load_dword %0, %1, 0x5c0\n\t"
        "s_load_dword %0, %1, 0x600\n\t"
        "s_load_dword %0, %1, 0x640\n\t"
        "s_load_dword %0, %1, 0x680\n\t"
        "s_load_dword %0, %1, 0x6c0\n\t"
        "s_load_dword %0, %1, 0x700\n\t"
        "s_load_dword %0, %1, 0x740\n\t"
        "s_load_dword %0, %1, 0x780\n\t"
        "s_load_dword %0, %1, 0x7c0\n\t"
        "s_load_dword %0, %1, 0x800\n\t"
        "s_load_dword %0, %1, 0x840\n\t"
        "s_load_dword %0, %1, 0x880\n\t"
        "s_load_dword %0, %1, 0x8c0\n\t"
        "s_load_dword %0, %1, 0x900\n\t"
        "s_load_dword %0, %1, 0x940\n\t"
        "s_load_dword %0, %1, 0x980\n\t"
        "s_load_dword %0, %1, 0x9c0\n\t"
        "s_load_dword %0, %1, 0xa00\n\t"
        "s_load_dword %0, %1, 0xa40\n\t"
        "s_load_dword %0, %1, 0xa80\n\t"
        "s_load_dword %0, %1, 0xac0\n\t"
        "s_load_dword %0, %1, 0xb00\n\t"
        "s_load_dword %0, %1, 0xb40\n\t"
        "s_load_dword %0, %1, 0xb80\n\t"
        "s_load_dword %0, %1, 0xbc0\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&s"(t)
        : "s"(base));
    asm volatile("" ::"s"(t));
}

// per-k block (k_schur_kdata): [Hcc: NC*NC cd][P: nd*NP cd][e_d: nd doubles]
template <int NC>
__host__ __device__ __forceinline__ int kblock_size(int nd)
{
    constexpr int NP = NC * (NC + 1) / 2;
    return NC * NC + nd * NP + (nd + 1) / 2;
}

// One (frequency tile, k chunk) work item.  The per-k block is read from
// a.kdata through the constant address space: its addresses are wave-uniform,
// so the loads become scalar loads and the products against it take SGPR
// operands, with no LDS traffic and no barriers in the k loop.
// KS = 2 (the FP32R forward only): the WG lanes are KS wavefronts over the same
// WG / KS frequencies, each walking every KS-th k of the chunk and adding into
// one set of LDS accumulators, which then costs a wavefront half the LDS.
template <int NC, int MODE, bool F32R, int WG, bool TRACE, int REFINE, int KS = 1>
__device__ __forceinline__ void schur_item(const LatArgs& a, const cd* __restrict__ sig_all, int wtile, int chunk,
                                           cd* lds)
{
    using BT = kconst_cd;
    using BD = kconst_double;
    constexpr int NP = NC * (NC + 1) / 2;
    constexpr int M2 = NC * NC;
    constexpr bool HERM = MODE == MODE_FWD_HERM || MODE == MODE_DENS_HERM;
    constexpr bool FWDLIKE = MODE == MODE_FWD || MODE == MODE_FWD_HERM || MODE == MODE_FWD_REAL;
    // accumulators that do not fit next to the working matrices live in LDS
    constexpr bool ACC_LDS = (MODE == MODE_VJP) || (MODE == MODE_DENS_HERM) || (FWDLIKE && F32R);
    constexpr int FW = WG / KS;       // frequencies per workgroup
    static_assert(KS == 1 || (KS == 2 && MODE == MODE_FWD && ACC_LDS && FW % 64 == 0),
                  "k pairs: LDS-accumulating forward, whole wavefronts per part");
    const int nd = a.n - NC;
    const int bs = kblock_size<NC>(nd);
    cd* sAcc = lds;                   // [M2][FW] complex, frequency-contiguous
    double* sAccH = (double*)lds;     // HERM: [M2][FW] reals
    cd* sTr = lds + M2 * FW;          // KS > 1: [FW] sums of the tr G shares
    const int lane = KS == 1 ? (int)threadIdx.x : (int)threadIdx.x % FW;  // this lane's frequency in the tile
    // which k of every KS: wave-uniform (FW is a multiple of 64), so that the k
    // walk and the per-k block addresses stay scalar
    const int part = KS == 1 ? 0 : __builtin_amdgcn_readfirstlane((int)threadIdx.x / FW);
    const int nw_hi = a.nw_end ? a.nw_end : a.nw_tot;
    const int nw = a.nw_begin + wtile * FW + lane;
    const bool wvalid = nw < nw_hi;
    const int nwc = wvalid ? nw : nw_hi - 1;  // clamp for loads
    const int k0 = chunk * a.k_per_chunk;
    const int k1 = min(a.nk, k0 + a.k_per_chunk);
    // z + mu = (mu + shift) + i w: the Matsubara point, or on the real axis E + mu + i eta
    double mu, w;
    if constexpr (MODE == MODE_FWD_REAL) {
        freq_point(a, nwc, mu, w);
    } else {
        w = matsubara(nwc, a.beta);
        mu = a.mu;
    }
    const bool vjp_here = (MODE == MODE_VJP) && wvalid && nw < a.n_freq;
    const cd* sig = sig_all + (size_t)nwc * M2;
    const cd* gb = (MODE == MODE_VJP) ? a.gbar + (size_t)(nw < a.n_freq ? nw : 0) * M2 : nullptr;
    // nbar: host value, or (batched adjoint) a device scalar written by an earlier kernel
    const double cn = 2.0 * (a.nbar_dev ? *a.nbar_dev : a.nbar) / a.beta;

    cd acc[ACC_LDS || HERM ? 1 : NC][ACC_LDS || HERM ? 1 : NC];
    // HERM: sums of Re G_ii and, for i < j, of (G_ij + G_ji).x, (G_ij - G_ji).y,
    // in the order (j ascending, i <= j ascending)
    double accH[HERM && !ACC_LDS ? M2 : 1];
    if (HERM) {
        if (ACC_LDS) {
            for (int e = 0; e < M2; ++e) sAccH[e * FW + lane] = 0.0;
        } else {
#pragma unroll
            for (int e = 0; e < (HERM && !ACC_LDS ? M2 : 1); ++e) accH[e] = 0.0;
        }
    } else if (ACC_LDS) {
        // (KS > 1: every part clears the same entries, before the barrier below)
        for (int e = 0; e < M2; ++e) sAcc[e * FW + lane] = cd{0.0, 0.0};
        if constexpr (KS > 1) {
            sTr[lane] = cd{0.0, 0.0};
            __syncthreads();
        }
    } else {
#pragma unroll
        for (int i = 0; i < (ACC_LDS ? 1 : NC); ++i)
#pragma unroll
            for (int j = 0; j < (ACC_LDS ? 1 : NC); ++j) acc[i][j] = cd{0.0, 0.0};
    }
    cd trA = cd{0.0, 0.0}, tr2A = cd{0.0, 0.0}, muA = cd{0.0, 0.0};

    // KS > 1: both parts take the same number of steps (the barriers below); a
    // part past the chunk's end repeats its last k and adds nothing
    for (int kb = k0; kb < k1; kb += KS) {
        const int k = KS == 1 ? kb : min(kb + part, k1 - 1);
        const bool kin = KS == 1 || kb + part < k1;
        // Sigma(iw) is re-read from the cache every k rather than hoisted out
        // of the loop, where it would hold 4*NC*NC registers for its duration.
        // The k block is read through the constant address space, which the
        // compiler treats as invariant: those loads stay scalar despite the
        // clobber.
        __asm__ __volatile__("" ::: "memory");
        kcache_fill(a.kdata + (size_t)k * bs);
        // Sigma's loads are issued here, all together, so that they complete under
        // the d loop below instead of waiting in turn where S is combined; not in
        // the FP64 forward kernels, whose G_cc sums already fill the registers
        constexpr bool SIG_EARLY = F32R || !FWDLIKE;
        cd sg[SIG_EARLY ? NC : 1][SIG_EARLY ? NC : 1];
        if constexpr (SIG_EARLY) {
#pragma unroll
            for (int ib = 0; ib < NC; ++ib)
#pragma unroll
                for (int ia = 0; ia < NC; ++ia) sg[ia][ib] = sig[ia + ib * NC];
        }
        auto sig_at = [&](int ia, int ib) -> cd {
            if constexpr (SIG_EARLY) return sg[ia][ib];
            else return sig[ia + ib * NC];
        };
        const BT* B = (const BT*)(a.kdata + (size_t)k * bs);
        const BT* Hcc = B;
        const BT* P = B + M2;
        const BD* ed = (const BD*)(P + nd * NP);
        // ---- Schur complement S = (z+mu) I - H_cc - Sigma - sum_d g_d P_d.
        // sum_d g_d P_d = R + i J with R = sum_d Re(g_d) P_d and
        // J = sum_d Im(g_d) P_d both Hermitian: their upper triangles (real
        // diagonals) are accumulated in place of S, G[a][b] = R_ab,
        // G[b][a] = J_ab (a < b), G[a][a] = (R_aa, J_aa), a real-by-complex
        // update per stored entry (4 FMA per pair instead of 8).
        cd G[NC][NC];
#pragma unroll
        for (int j = 0; j < NC; ++j)
#pragma unroll
            for (int i = 0; i < NC; ++i) G[i][j] = cd{0.0, 0.0};
        // g_d = 1/(z + mu - e_d) one d ahead: its reciprocal chain is independent
        // of the d step's FMAs, so the two interleave instead of the chain
        // stalling the top of every step; e_{d+1} is read one step earlier still,
        // with that step's P_d, so that the chain does not wait for it either
        auto g_of = [&](double e) -> cd {
            const double x = mu - e;
            const double r = rcp_pos(fma(x, x, w * w));
            return cd{x * r, -w * r};
        };
        auto e_at = [&](int d) -> double { return (double)ed[d < nd ? d : nd - 1]; };
        cd g_next = nd > 0 ? g_of(e_at(0)) : cd{0.0, 0.0};
        double e_next = nd > 0 ? e_at(1) : 0.0;
        for (int d = 0; d < nd; ++d) {
            const double gr = g_next.x, gi = g_next.y;
            g_next = g_of(e_next);
            e_next = e_at(d + 2);
            const BT* Pd = P + d * NP;
#pragma unroll
            for (int ib = 0; ib < NC; ++ib)
#pragma unroll
                for (int ia = 0; ia <= ib; ++ia) {
                    const cd p = ldc(Pd + ib * (ib + 1) / 2 + ia);
                    if (ia == ib) {  // P_aa is real
                        G[ia][ia].x = fma(gr, p.x, G[ia][ia].x);
                        G[ia][ia].y = fma(gi, p.x, G[ia][ia].y);
                    } else {
                        G[ia][ib] = cd{fma(gr, p.x, G[ia][ib].x), fma(gr, p.y, G[ia][ib].y)};
                        G[ib][ia] = cd{fma(gi, p.x, G[ib][ia].x), fma(gi, p.y, G[ib][ia].y)};
                    }
                }
        }
#pragma unroll
        for (int ib = 0; ib < NC; ++ib)
#pragma unroll
            for (int ia = 0; ia <= ib; ++ia) {
                cd bab = csub(cneg(ldc(Hcc + ia + ib * NC)), sig_at(ia, ib));
                if (ia == ib) {
                    bab = cadd(bab, cd{mu, w});
                    G[ia][ia] = csub(bab, G[ia][ia]);  // - (R_aa + i J_aa)
                } else {
                    const cd R = G[ia][ib], J = G[ib][ia];
                    const cd bba = csub(cneg(ldc(Hcc + ib + ia * NC)), sig_at(ib, ia));
                    G[ia][ib] = csub(bab, cd{R.x - J.y, R.y + J.x});   // - (R + i J)_ab
                    G[ib][ia] = csub(bba, cd{R.x + J.y, J.x - R.y});   // - (R + i J)_ba
                }
            }
        // ---- G_cc = S^{-1}
        if (!F32R) {
            gj_inverse<NC>(G);
        } else {
            // G holds S (FP64); X0 = S^{-1} in FP32
            cf X0[NC][NC];
#pragma unroll
            for (int i = 0; i < NC; ++i)
#pragma unroll
                for (int j = 0; j < NC; ++j) X0[i][j] = cf{(float)G[i][j].x, (float)G[i][j].y};
            gj_inverse_f32<NC>(X0);
            if constexpr (REFINE == 1) {
                // One step: the whole residual R = I - S X0 (FP64 products,
                // FP32 storage) is formed first, after which S is dead and
                // G = X0 + X0 R is written over it.  Peak: S + X0 + R.
                // R is formed row by row so that each row of S dies as soon as
                // its row of R is done.
                cf Rm[NC][NC];
#pragma unroll
                for (int i = 0; i < NC; ++i)
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        cd s = cd{i == c ? 1.0 : 0.0, 0.0};
#pragma unroll
                        for (int l = 0; l < NC; ++l) s = cfms(G[i][l], cd{(double)X0[l][c].x, (double)X0[l][c].y}, s);
                        Rm[i][c] = cf{(float)s.x, (float)s.y};
                    }
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int i = 0; i < NC; ++i) {
                        cf t = cf{0.f, 0.f};
#pragma unroll
                        for (int l = 0; l < NC; ++l) t = ffma(X0[i][l], Rm[l][c], t);
                        G[i][c] = cd{(double)X0[i][c].x + (double)t.x, (double)X0[i][c].y + (double)t.y};
                    }
            } else {
            // X = X0 + D with the correction D held in FP32: D is ~kappa*eps32 of
            // |X|, so its own FP32 rounding (~kappa*eps32^2 of |X|) is the
            // accuracy floor of the refinement, the same as adding an FP32
            // correction into an FP64 X, at half the registers.
            cf D[NC][NC];
#pragma unroll
            for (int i = 0; i < NC; ++i)
#pragma unroll
                for (int j = 0; j < NC; ++j) D[i][j] = cf{0.f, 0.f};
            for (int it = 0; it < REFINE; ++it) {
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    cd xc[NC];  // column c of X in FP64
#pragma unroll
                    for (int l = 0; l < NC; ++l)
                        xc[l] = cd{(double)X0[l][c].x + (double)D[l][c].x, (double)X0[l][c].y + (double)D[l][c].y};
                    cf rc[NC];  // r = e_c - S x_c, formed in FP64, stored in FP32
#pragma unroll
                    for (int i = 0; i < NC; ++i) {
                        cd s = cd{i == c ? 1.0 : 0.0, 0.0};
#pragma unroll
                        for (int l = 0; l < NC; ++l) s = cfms(G[i][l], xc[l], s);
                        rc[i] = cf{(float)s.x, (float)s.y};
                    }
#pragma unroll
                    for (int i = 0; i < NC; ++i) {
                        cf t = D[i][c];
#pragma unroll
                        for (int l = 0; l < NC; ++l) t = ffma(X0[i][l], rc[l], t);
                        D[i][c] = t;
                    }
                }
            }
#pragma unroll
            for (int i = 0; i < NC; ++i)
#pragma unroll
                for (int j = 0; j < NC; ++j)
                    G[i][j] = cd{(double)X0[i][j].x + (double)D[i][j].x, (double)X0[i][j].y + (double)D[i][j].y};
            }
        }

        if constexpr (HERM) {  // Hermitian-part accumulation (before G is folded below)
            int q = 0;
#pragma unroll
            for (int j = 0; j < NC; ++j)
#pragma unroll
                for (int i = 0; i <= j; ++i) {
                    const double h0 = i == j ? G[i][i].x : G[i][j].x + G[j][i].x;
                    const double h1 = G[i][j].y - G[j][i].y;
                    if (ACC_LDS) {
                        lds_add(&sAccH[q * FW + lane], h0);
                        if (i != j) lds_add(&sAccH[(q + 1) * FW + lane], h1);
                    } else {
                        accH[q] += h0;
                        if (i != j) accH[q + 1] += h1;
                    }
                    q += i == j ? 1 : 2;
                }
        }
        if (FWDLIKE) {
            if constexpr (HERM) {
            } else if (ACC_LDS) {
                auto add_g = [&]() {
#pragma unroll
                    for (int j = 0; j < NC; ++j)
#pragma unroll
                        for (int i = 0; i < NC; ++i) lds_add(&sAcc[(i + j * NC) * FW + lane], G[i][j]);
                };
                if constexpr (KS > 1) {
                    // the parts add in turn, part 0 first, so that every sum has one
                    // order and the result is the same on every run
                    __syncthreads();
                    if (part == 0) add_g();
                    __syncthreads();
                    if (part == 1 && kin) add_g();
                } else {
                    add_g();
                }
            } else {
#pragma unroll
                for (int i = 0; i < (ACC_LDS || HERM ? 1 : NC); ++i)
#pragma unroll
                    for (int j = 0; j < (ACC_LDS || HERM ? 1 : NC); ++j) acc[i][j] = cadd(acc[i][j], G[i][j]);
            }
            if (TRACE) {  // tr G = sum_d g_d + tr G_cc + sum_d g_d^2 M_d
                cd t = cd{0.0, 0.0};
#pragma unroll
                for (int i = 0; i < NC; ++i) t = cadd(t, G[i][i]);
                // M_d = sum_ab G_ab P_d(b, a): for a < b the pair contributes
                // Re = (G_ab + G_ba).x p.x + (G_ab - G_ba).y p.y and
                // Im = (G_ab + G_ba).y p.x + (G_ba - G_ab).x p.y (p = P_d(a, b)),
                // so G is folded once into those sums (in place; G is not needed
                // after this) and each pair costs 4 FMA per d.
#pragma unroll
                for (int ib = 0; ib < NC; ++ib)
#pragma unroll
                    for (int ia = 0; ia < ib; ++ia) {
                        const cd x1 = G[ia][ib], x2 = G[ib][ia];
                        G[ia][ib] = cd{x1.x + x2.x, x1.y - x2.y};
                        G[ib][ia] = cd{x1.y + x2.y, x2.x - x1.x};
                    }
                g_next = nd > 0 ? g_of(e_at(0)) : cd{0.0, 0.0};
                e_next = nd > 0 ? e_at(1) : 0.0;
                for (int d = 0; d < nd; ++d) {
                    const cd g = g_next;
                    g_next = g_of(e_next);
                    e_next = e_at(d + 2);
                    const BT* Pd = P + d * NP;
                    cd Md = cd{0.0, 0.0};
#pragma unroll
                    for (int ib = 0; ib < NC; ++ib)
#pragma unroll
                        for (int ia = 0; ia <= ib; ++ia) {
                            const cd p = ldc(Pd + ib * (ib + 1) / 2 + ia);
                            if (ia == ib) {
                                Md.x = fma(G[ia][ia].x, p.x, Md.x);
                                Md.y = fma(G[ia][ia].y, p.x, Md.y);
                            } else {
                                const cd sd = G[ia][ib], sm = G[ib][ia];
                                Md.x = fma(sd.x, p.x, fma(sd.y, p.y, Md.x));
                                Md.y = fma(sm.x, p.x, fma(sm.y, p.y, Md.y));
                            }
                        }
                    t = cadd(t, cadd(g, cmul(cmul(g, g), Md)));
                }
                if (wvalid && kin) trA = cadd(trA, t);
            }
        } else {
            // the adjoint's cotangent, loaded all together here so that it arrives
            // under the d loop below (it is k-invariant; keeping it across k would
            // hold 4 NC^2 registers for the whole walk)
            cd gbv[MODE == MODE_VJP ? NC : 1][MODE == MODE_VJP ? NC : 1];
            if constexpr (MODE == MODE_VJP) {
#pragma unroll
                for (int j = 0; j < NC; ++j)
#pragma unroll
                    for (int i = 0; i < NC; ++i) gbv[i][j] = gb[i + j * NC];
            }
            // ---- Y = sum_d g_d^2 P_d (X of the header), trace terms via M_d = w_d^H G w_d
            cd Y[NC][NC];
#pragma unroll
            for (int i = 0; i < NC; ++i)
#pragma unroll
                for (int j = 0; j < NC; ++j) Y[i][j] = cd{0.0, 0.0};
            cd sg = cd{0.0, 0.0}, sg2 = cd{0.0, 0.0}, sgM2 = cd{0.0, 0.0}, sgM3 = cd{0.0, 0.0};
            // Y is accumulated as R + i J (R = sum Re(g^2) P_d, J = sum Im(g^2) P_d,
            // upper triangles in place: Y[a][b] = R_ab, Y[b][a] = J_ab), as S is
            // M_d = sum_ab conj(w_a) G_ab w_b = sum_ab G_ab P_{ba} with G folded
            // into pair sums for the loop (as in the forward trace) and
            // unfolded after it: for a < b, G[a][b] <- (G_ab + G_ba).x, (G_ab - G_ba).y
            // and G[b][a] <- (G_ab + G_ba).y, (G_ba - G_ab).x
#pragma unroll
            for (int ib = 0; ib < NC; ++ib)
#pragma unroll
                for (int ia = 0; ia < ib; ++ia) {
                    const cd x1 = G[ia][ib], x2 = G[ib][ia];
                    G[ia][ib] = cd{x1.x + x2.x, x1.y - x2.y};
                    G[ib][ia] = cd{x1.y + x2.y, x2.x - x1.x};
                }
            g_next = nd > 0 ? g_of(e_at(0)) : cd{0.0, 0.0};
            e_next = nd > 0 ? e_at(1) : 0.0;
            for (int d = 0; d < nd; ++d) {
                const cd g = g_next;
                g_next = g_of(e_next);
                e_next = e_at(d + 2);
                const cd g2 = cmul(g, g);
                const BT* Pd = P + d * NP;
                cd Md = cd{0.0, 0.0};
#pragma unroll
                for (int ib = 0; ib < NC; ++ib)
#pragma unroll
                    for (int ia = 0; ia <= ib; ++ia) {
                        const cd p = ldc(Pd + ib * (ib + 1) / 2 + ia);  // P_{ia ib} = w_a conj(w_b)
                        if (ia == ib) {
                            Y[ia][ia].x = fma(g2.x, p.x, Y[ia][ia].x);
                            Y[ia][ia].y = fma(g2.y, p.x, Y[ia][ia].y);
                            Md.x = fma(G[ia][ia].x, p.x, Md.x);
                            Md.y = fma(G[ia][ia].y, p.x, Md.y);
                        } else {
                            Y[ia][ib] = cd{fma(g2.x, p.x, Y[ia][ib].x), fma(g2.x, p.y, Y[ia][ib].y)};
                            Y[ib][ia] = cd{fma(g2.y, p.x, Y[ib][ia].x), fma(g2.y, p.y, Y[ib][ia].y)};
                            const cd sd = G[ia][ib], sm = G[ib][ia];
                            Md.x = fma(sd.x, p.x, fma(sd.y, p.y, Md.x));
                            Md.y = fma(sm.x, p.x, fma(sm.y, p.y, Md.y));
                        }
                    }
                sg = cadd(sg, g);
                sg2 = cadd(sg2, g2);
                sgM2 = cfma(g2, Md, sgM2);
                sgM3 = cfma(cmul(g2, g), Md, sgM3);
            }
#pragma unroll
            for (int ib = 0; ib < NC; ++ib)
#pragma unroll
                for (int ia = 0; ia < ib; ++ia) {
                    const cd a1 = G[ia][ib], a2 = G[ib][ia];
                    G[ia][ib] = cd{0.5 * (a1.x - a2.y), 0.5 * (a2.x + a1.y)};
                    G[ib][ia] = cd{0.5 * (a1.x + a2.y), 0.5 * (a2.x - a1.y)};
                }
#pragma unroll
            for (int ib = 0; ib < NC; ++ib)
#pragma unroll
                for (int ia = 0; ia < ib; ++ia) {
                    const cd R = Y[ia][ib], J = Y[ib][ia];
                    Y[ia][ib] = cd{R.x - J.y, R.y + J.x};
                    Y[ib][ia] = cd{R.x + J.y, J.x - R.y};
                }
            cd trG = cadd(sg, sgM2);
#pragma unroll
            for (int i = 0; i < NC; ++i) trG = cadd(trG, G[i][i]);
            // Z = G S' = G + G Y.  DENS: in place of Y.  VJP: apart, as S' = I + Y
            // is used again below.
            cd Z[MODE == MODE_VJP ? NC : 1][MODE == MODE_VJP ? NC : 1];
            auto Zr = [&](int i, int j) -> cd& { return MODE == MODE_VJP ? Z[i][j] : Y[i][j]; };
#pragma unroll
            for (int j = 0; j < NC; ++j) {
                cd t[NC];
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    cd s = G[i][j];
#pragma unroll
                    for (int l = 0; l < NC; ++l) s = cfma(G[i][l], Y[l][j], s);
                    t[i] = s;
                }
#pragma unroll
                for (int i = 0; i < NC; ++i) Zr(i, j) = t[i];
            }
            // tr(Z^2) = sum_i Z_ii^2 + 2 sum_{i<j} Z_ij Z_ji
            cd tzz = cd{0.0, 0.0};
            {
                cd off = cd{0.0, 0.0};
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    tzz = cfma(Zr(i, i), Zr(i, i), tzz);
#pragma unroll
                    for (int j = i + 1; j < NC; ++j) off = cfma(Zr(i, j), Zr(j, i), off);
                }
                tzz = cadd(tzz, cscale(off, 2.0));
            }
            const cd tz = cadd(cadd(sg2, cscale(sgM3, 2.0)), tzz);
            if (wvalid) {
                trA = cadd(trA, trG);
                tr2A = cadd(tr2A, tz);
            }
            if (MODE == MODE_VJP && vjp_here) {
                // sigma_bar += G^H gbar G^H + cn ((G^2)_cc)^H with (G^2)_cc = G S' G:
                // both terms are C = G^H V G^H, V = gbar + cn S'^H.
                // mu term tr(gbar^H (G^2)_cc) = tr(S' G gbar^H G)
                //   = sum_ab S'_ab conj(C_ab) - cn tr(Z^2)
                // (G gbar^H G is (G^H gbar G^H)^H = (C - cn G^H S'^H G^H)^H, and
                // tr(S' G S' G) = tr(Z^2)).
                cd V[NC][NC];
#pragma unroll
                for (int j = 0; j < NC; ++j)
#pragma unroll
                    for (int i = 0; i < NC; ++i) {
                        const cd gv = gbv[i][j];
                        const double dg = i == j ? cn : 0.0;
                        V[i][j] = cd{fma(cn, Y[j][i].x, gv.x + dg), fma(-cn, Y[j][i].y, gv.y)};
                    }
                cd ms = cscale(tzz, -cn);
#pragma unroll
                for (int b = 0; b < NC; ++b) {
                    cd t[NC];  // column b of L = V G^H: t_i = sum_l V_il conj(G_bl)
#pragma unroll
                    for (int i = 0; i < NC; ++i) {
                        cd s = cd{0.0, 0.0};
#pragma unroll
                        for (int l = 0; l < NC; ++l) s = cfma(V[i][l], cconj(G[b][l]), s);
                        t[i] = s;
                    }
#pragma unroll
                    for (int aa = 0; aa < NC; ++aa) {  // C_ab = sum_i conj(G_ia) t_i
                        cd s = cd{0.0, 0.0};
#pragma unroll
                        for (int i = 0; i < NC; ++i) s = cfmac(G[i][aa], t[i], s);
                        lds_add(&sAcc[(aa + b * NC) * FW + lane], s);
                        const cd sp = aa == b ? cadd(Y[aa][b], cd{1.0, 0.0}) : Y[aa][b];
                        ms = cfmac(s, sp, ms);  // S'_ab conj(C_ab)
                    }
                }
                muA = cadd(muA, ms);
            }
        }
    }

    if constexpr (KS > 1) {  // the parts' tr G shares and G_cc sums meet in LDS; part 0 writes them
        lds_add(&sTr[lane], trA);
        __syncthreads();
        if (part != 0) return;
        trA = sTr[lane];
    }
    if (!wvalid) return;
    const size_t ch = chunk;
    if (HERM) {
        // (G + G^H)/2 in full layout: NumMatrix sums G_ij + conj(G_ji) over frequencies
        a.pTr[ch * a.nw_tot + nw] = trA;
        if (MODE == MODE_DENS_HERM) a.pTr2[ch * a.nw_tot + nw] = tr2A;
        cd* out = a.pG + (ch * a.nw_tot + nw) * M2;
        int q = 0;
#pragma unroll
        for (int j = 0; j < NC; ++j)
#pragma unroll
            for (int i = 0; i <= j; ++i) {
                const double h0 = ACC_LDS ? sAccH[q * FW + lane] : accH[q];
                if (i == j) {
                    out[i + i * NC] = cd{h0, 0.0};
                    q += 1;
                } else {
                    const double h1 = ACC_LDS ? sAccH[(q + 1) * FW + lane] : accH[q + 1];
                    out[i + j * NC] = cd{0.5 * h0, 0.5 * h1};
                    out[j + i * NC] = cd{0.5 * h0, -0.5 * h1};
                    q += 2;
                }
            }
    } else if (MODE == MODE_FWD || MODE == MODE_FWD_REAL) {
        a.pTr[ch * a.nw_tot + nw] = trA;
        if (ACC_LDS) {
            for (int e = 0; e < M2; ++e) a.pG[(ch * a.nw_tot + nw) * M2 + e] = sAcc[e * FW + lane];
        } else {
#pragma unroll
            for (int i = 0; i < (ACC_LDS || HERM ? 1 : NC); ++i)
#pragma unroll
                for (int j = 0; j < (ACC_LDS || HERM ? 1 : NC); ++j)
                    a.pG[(ch * a.nw_tot + nw) * M2 + i + j * NC] = acc[i][j];
        }
    } else {
        a.pTr[ch * a.nw_tot + nw] = trA;
        a.pTr2[ch * a.nw_tot + nw] = tr2A;
        if (MODE == MODE_VJP && vjp_here) {
            a.pMu[ch * a.n_freq + nw] = muA;
            for (int e = 0; e < M2; ++e) a.pSb[(ch * a.n_freq + nw) * M2 + e] = sAcc[e * FW + lane];
        }
    }
}

// Waves per SIMD the register allocator is held to.  At n_c = 5 the forward and
// density kernels (FP64, or FP32 with one refinement step) need a few registers
// more than the 256 of two waves; two waves with a few bytes of scratch beat one.
template <int NC, int MODE, int REFINE>
constexpr int schur_waves() { return (NC == 5 && MODE != MODE_VJP && REFINE <= 1) ? 2 : 1; }

template <int NC, int MODE, bool F32R, int WG, bool TRACE, int REFINE, int KS = 1>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(schur_waves<NC, MODE, REFINE>())))
void k_lattice_schur(LatArgs a, const cd* __restrict__ sig_all)
{
    extern __shared__ cd lds[];
    schur_item<NC, MODE, F32R, WG, TRACE, REFINE, KS>(a, sig_all, blockIdx.x, blockIdx.y, lds);
}

// Persistent form over a work list spanning several lattices of the same n_c:
// workgroup b takes items b, b + gridDim.x, ... (static striding, no atomics,
// so every replay does the same work in the same order).  Each system's per-k
// blocks come from its LatArgs (kdata), read through the constant address space.
template <int NC, int MODE, bool F32R, int WG, bool TRACE, int REFINE, int KS = 1>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(schur_waves<NC, MODE, REFINE>())))
void k_lattice_schur_multi(const LatArgs* __restrict__ args,
                                                            const int4* __restrict__ items, int n_items)
{
    extern __shared__ cd lds[];
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const int4 item = items[it];  // {system, wtile, chunk, -}
        const LatArgs a = args[item.x];
        schur_item<NC, MODE, F32R, WG, TRACE, REFINE, KS>(a, a.pSigAll, item.y, item.z, lds);
        __syncthreads();
    }
}

template <int NC, int MODE, bool F32R, int WG, int KS = 1>
constexpr size_t schur_shm()
{
    if constexpr (MODE == MODE_FWD_HERM) return F32R ? sizeof(double) * (size_t)NC * NC * WG : 0;
    if constexpr (MODE == MODE_DENS_HERM) return sizeof(double) * (size_t)NC * NC * WG;
    constexpr bool ACC_LDS = (MODE == MODE_VJP) || ((MODE == MODE_FWD || MODE == MODE_FWD_REAL) && F32R);
    // KS > 1: one accumulator set per frequency, plus the tr G sums
    return ACC_LDS ? sizeof(cd) * (size_t)(NC * NC + (KS > 1 ? 1 : 0)) * (WG / KS) : 0;
}

template <int NC, int MODE, bool F32R, bool TRACE, int REFINE, int WG, int KS = 1>
hipError_t launch_nc_w(const LatArgs& a, const cd* sig_all, hipStream_t s, const SchurMulti* multi)
{
    constexpr size_t shm = schur_shm<NC, MODE, F32R, WG, KS>();
    if (multi) {
        hipError_t e = hipFuncSetAttribute((const void*)k_lattice_schur_multi<NC, MODE, F32R, WG, TRACE, REFINE, KS>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_lattice_schur_multi<NC, MODE, F32R, WG, TRACE, REFINE, KS>), dim3(multi->grid), dim3(WG),
                           shm, s, multi->args, multi->items, multi->n_items);
        return hipGetLastError();
    }
    constexpr int FW = WG / KS;  // frequencies per workgroup
    const int span = (a.nw_end ? a.nw_end : a.nw_tot) - a.nw_begin;
    dim3 grid((span + FW - 1) / FW, a.n_chunks);
    hipError_t e = hipFuncSetAttribute((const void*)k_lattice_schur<NC, MODE, F32R, WG, TRACE, REFINE, KS>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_lattice_schur<NC, MODE, F32R, WG, TRACE, REFINE, KS>), grid, dim3(WG), shm, s, a, sig_all);
    return hipGetLastError();
}

template <int NC, int MODE, bool F32R, bool TRACE, int REFINE>
hipError_t launch_nc_r(const LatArgs& a, const cd* sig_all, hipStream_t s, int wg, const SchurMulti* m)
{
    if constexpr (MODE == MODE_FWD_REAL) {  // one width on the real axis
        (void)wg;
        return launch_nc_w<NC, MODE, F32R, TRACE, REFINE, 128>(a, sig_all, s, m);
    } else {
        if constexpr (MODE == MODE_FWD && F32R && NC >= 5) {  // two wavefronts over 64 frequencies
            if (wg == SCHUR_WG_KPAIR) return launch_nc_w<NC, MODE, F32R, TRACE, REFINE, 128, 2>(a, sig_all, s, m);
        }
        if (wg == 128) return launch_nc_w<NC, MODE, F32R, TRACE, REFINE, 128>(a, sig_all, s, m);
        return launch_nc_w<NC, MODE, F32R, TRACE, REFINE, 64>(a, sig_all, s, m);
    }
}

template <int NC, int MODE, bool F32R, bool TRACE>
hipError_t launch_nc(const LatArgs& a, const cd* sig_all, hipStream_t s, int wg, const SchurMulti* m)
{
    if constexpr (!F32R) {
        return launch_nc_r<NC, MODE, false, TRACE, 0>(a, sig_all, s, wg, m);
    } else {
        if (a.refine <= 1) return launch_nc_r<NC, MODE, true, TRACE, 1>(a, sig_all, s, wg, m);
        return launch_nc_r<NC, MODE, true, TRACE, 2>(a, sig_all, s, wg, m);
    }
}

// All variants of one n_c: mode, precision, trace.
template <int NC>
hipError_t launch_lane_nc(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s, int wg,
                          const SchurMulti* m)
{
    if (mode == MODE_FWD) {
        if (a.want_trace)
            return fp32 ? launch_nc<NC, MODE_FWD, true, true>(a, sig_all, s, wg, m)
                        : launch_nc<NC, MODE_FWD, false, true>(a, sig_all, s, wg, m);
        return fp32 ? launch_nc<NC, MODE_FWD, true, false>(a, sig_all, s, wg, m)
                    : launch_nc<NC, MODE_FWD, false, false>(a, sig_all, s, wg, m);
    }
    if (mode == MODE_FWD_HERM) {
        if (a.want_trace)
            return fp32 ? launch_nc<NC, MODE_FWD_HERM, true, true>(a, sig_all, s, wg, m)
                        : launch_nc<NC, MODE_FWD_HERM, false, true>(a, sig_all, s, wg, m);
        return fp32 ? launch_nc<NC, MODE_FWD_HERM, true, false>(a, sig_all, s, wg, m)
                    : launch_nc<NC, MODE_FWD_HERM, false, false>(a, sig_all, s, wg, m);
    }
    if (mode == MODE_FWD_REAL) {
        if (a.want_trace)
            return fp32 ? launch_nc<NC, MODE_FWD_REAL, true, true>(a, sig_all, s, wg, m)
                        : launch_nc<NC, MODE_FWD_REAL, false, true>(a, sig_all, s, wg, m);
        return fp32 ? launch_nc<NC, MODE_FWD_REAL, true, false>(a, sig_all, s, wg, m)
                    : launch_nc<NC, MODE_FWD_REAL, false, false>(a, sig_all, s, wg, m);
    }
    if (mode == MODE_DENS)
        return fp32 ? launch_nc<NC, MODE_DENS, true, false>(a, sig_all, s, wg, m)
                    : launch_nc<NC, MODE_DENS, false, false>(a, sig_all, s, wg, m);
    if (mode == MODE_DENS_HERM)
        return fp32 ? launch_nc<NC, MODE_DENS_HERM, true, false>(a, sig_all, s, wg, m)
                    : launch_nc<NC, MODE_DENS_HERM, false, false>(a, sig_all, s, wg, m);
    return fp32 ? launch_nc<NC, MODE_VJP, true, false>(a, sig_all, s, wg, m)
                : launch_nc<NC, MODE_VJP, false, false>(a, sig_all, s, wg, m);
}

}  // namespace dg
