// densityMatDFT (GreenFtn.cpp:291-319) in block form, for n <= 16, n_c <= 6.
//
// With the correlated orbitals first (downfolding.cpp:69-79) and
// H_dd(k) = V_d diag(e_d) V_d^H, the lattice Green's function of one (k, iw)
// in the basis (c, V_d) is
//   G' = E G_cc F + diag(0, g),   E = [I; g W^H] (n x n_c),  F = [I, W g] (n_c x n),
//   g = diag(1/(z + mu - e_d)),   W = H_cd V_d,   G_cc = S^{-1},
//   S = (z + mu) I - H_cc - Sigma - W g W^H          (Schur complement),
// so the frequency sum of G' is one complex product per k,
//   sum_w G' = sum_w E_w (G_cc F)_w + diag(0, sum_w g),
// of inner dimension n_c N_w, and S(k) = sum_w G = T (sum_w G') T^H with
// T = diag(I, V_d).  The density matrix then follows as in the plain kernel,
//   D(k) = (S + S^H)/beta + U diag(w) U^H.
// Per (k, iw) this is n_c^2 n_d + n_c^3 (G_cc) + n_c^2 n_d (G_cc W) +
// n^2 n_c (the product) instead of the n^3 of a full inverse.
//
// Mapping: one wavefront per k, walking the frequencies PW = 48 at a time.
//   phase 1  lane = frequency: G_cc from the Schur complement (per-k data by
//            scalar loads, as in the one-matrix-per-lane kernels; FP64
//            Gauss-Jordan), stored to LDS;
//   phase 2  16-lane slots, slot s = one frequency, lane j = row / column j:
//            lane j forms E[j][a] and R[a][j] = (G_cc F)[a][j], which are the
//            A and B operands of v_mfma_f64_16x16x4 (A[i][q] on lane i + 16 q,
//            B[q][j] on lane j + 16 q) with the four slots as the inner index q,
//            so the 16 x 16 sum accumulates in the MFMA result registers with
//            no transposes.  A complex product takes three real MFMAs:
//            P1 = Er Rr, P2 = Ei Ri, P3 = (Er + Ei)(Rr + Ri), Re = P1 - P2,
//            Im = P3 - P1 - P2.
// The wave then applies T (.) T^H and the static weights for its k.
#include "lattice_schur_lane.hpp"

namespace dg {

typedef double dv4 __attribute__((ext_vector_type(4)));

// two waves per SIMD up to n_c = 5 (a few registers spill at n_c = 5)
template <int NC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NC <= 5 ? 2 : 1))) void k_density_lr(LatArgs a, const cd* __restrict__ Wg, const cd* __restrict__ Vd,
                                                    const double* __restrict__ edg, const cd* __restrict__ U,
                                                    const double* __restrict__ wt, cd* __restrict__ dens)
{
    using BT = kconst_cd;
    using BD = kconst_double;
    constexpr int NP = NC * (NC + 1) / 2;
    constexpr int M2 = NC * NC;
    // A pass takes PW = 48 frequencies (lanes 48..63 idle in phase 1): the
    // pass's G_cc then needs 19.2 KB of LDS at n_c = 5, which with the
    // registers allows two waves per SIMD (64 frequencies, 25.6 KB, allow 1.5).
    constexpr int PW = 48;
    // G_cc of the pass's frequencies, [w][i + b NC]; after the frequency loop
    // the same memory holds two 16 x 16 column-major work matrices
    __shared__ cd lds[PW * M2 > 512 ? PW * M2 : 512];
    __shared__ double sW[PW];  // the pass's Matsubara frequencies
    cd* sG = lds;
    cd* sA = lds;
    cd* sB = lds + 256;
    const int lane = threadIdx.x;
    const int k = blockIdx.x;
    const int n = a.n, nd = n - NC;
    const int s = lane >> 4, j = lane & 15;
    // phase 2: lane j = NC + d holds column d of W and e_d
    const bool jd = j >= NC && j < n;
    const int dj = jd ? j - NC : 0;
    cd wcol[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) wcol[q] = jd ? Wg[(size_t)k * NC * nd + q + (size_t)dj * NC] : cd{0.0, 0.0};
    const double e_j = jd ? edg[(size_t)k * nd + dj] : 0.0;
    // phase 1: the per-k block [H_cc][P_d pairs][e_d] through the constant address space
    const BT* B = (const BT*)(a.kdata + (size_t)k * kblock_size<NC>(nd));
    const BT* Hcc = B;
    const BT* P = B + M2;
    const BD* ed = (const BD*)(P + nd * NP);
    const double mu = a.mu;

    dv4 P1 = {0.0, 0.0, 0.0, 0.0}, P2 = P1, P3 = P1;
    cd gsum = cd{0.0, 0.0};
    for (int w0 = 0; w0 < a.nw_tot; w0 += PW) {
        // ---- phase 1: G_cc of frequency w0 + lane (lane < PW)
        {
            const int nw = w0 + lane;
            const bool wv = nw < a.nw_tot && lane < PW;
            const int nwc = wv ? nw : a.nw_tot - 1;
            const double w = matsubara(nwc, a.beta);
            const cd* sig = a.pSigAll + (size_t)nwc * M2;  // Sigma, moment tail beyond n_freq
            // sum_d g_d P_d = R + i J accumulated as in the Schur kernels:
            // G[a][b] = R_ab, G[b][a] = J_ab (a < b), G[a][a] = (R_aa, J_aa)
            cd G[NC][NC];
#pragma unroll
            for (int jj = 0; jj < NC; ++jj)
#pragma unroll
                for (int i = 0; i < NC; ++i) G[i][jj] = cd{0.0, 0.0};
            for (int d = 0; d < nd; ++d) {
                const double x = mu - ed[d];
                const double r = rcp_pos(fma(x, x, w * w));
                const double gr = x * r, gi = -w * r;
                const BT* Pd = P + d * NP;
#pragma unroll
                for (int ib = 0; ib < NC; ++ib)
#pragma unroll
                    for (int ia = 0; ia <= ib; ++ia) {
                        const cd p = ldc(Pd + ib * (ib + 1) / 2 + ia);
                        if (ia == ib) {
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
                    cd bab = csub(cneg(ldc(Hcc + ia + ib * NC)), sig[ia + ib * NC]);
                    if (ia == ib) {
                        bab = cadd(bab, cd{mu, w});
                        G[ia][ia] = csub(bab, G[ia][ia]);
                    } else {
                        const cd R = G[ia][ib], J = G[ib][ia];
                        const cd bba = csub(cneg(ldc(Hcc + ib + ia * NC)), sig[ib + ia * NC]);
                        G[ia][ib] = csub(bab, cd{R.x - J.y, R.y + J.x});
                        G[ib][ia] = csub(bba, cd{R.x + J.y, J.x - R.y});
                    }
                }
            gj_inverse<NC>(G);
            if (lane < PW) {
                sW[lane] = w;
#pragma unroll
                for (int b = 0; b < NC; ++b)
#pragma unroll
                    for (int i = 0; i < NC; ++i) sG[lane * M2 + i + b * NC] = wv ? G[i][b] : cd{0.0, 0.0};
            }
        }
        wave_sync();
        // ---- phase 2: four frequencies per step, one per 16-lane slot.  Row q
        // of R and column q of E feed three MFMAs; row q + 1 is formed while
        // they run (in-order issue: the MFMAs of a row are followed by the
        // next row's VALU work, not by more MFMAs).  No divergent branches:
        // lanes j < NC have wcol = 0 and g = 0, lanes j >= NC a zero mask m1.
#pragma unroll 1
        for (int t = 0; t < PW / 4; ++t) {
            const int wl = 4 * t + s;
            const bool v2 = w0 + wl < a.nw_tot;
            const double w2 = sW[wl];
            const cd* Gs = sG + wl * M2;
            const double x = mu - e_j;
            const double rr = rcp_pos(fma(x, x, w2 * w2));
            const double md = (jd && v2) ? 1.0 : 0.0;  // lane of a d column, valid frequency
            const double m1 = (j < NC && v2) ? 1.0 : 0.0;  // lane of a c column, valid frequency
            const cd g = cd{md * x * rr, -md * w2 * rr};
            gsum = cadd(gsum, g);
            const int jc = j < NC ? j : 0;
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                // R[q][j] = G_cc[q][j] (j < NC) or (G_cc W)_{q d} g_d (j = NC + d)
                cd tq = cd{0.0, 0.0};
#pragma unroll
                for (int b = 0; b < NC; ++b) tq = cfma(Gs[q + b * NC], wcol[b], tq);
                const cd gc = Gs[q + jc * NC];
                const cd tg = cmul(tq, g);
                const cd r = cd{fma(m1, gc.x, tg.x), fma(m1, gc.y, tg.y)};
                // E[j][q] = delta_jq (j < NC) or g_d conj(W_qd) (j = NC + d)
                const cd eg = cmulc(g, wcol[q]);
                const cd e = cd{(j == q ? m1 : 0.0) + eg.x, eg.y};
                P1 = __builtin_amdgcn_mfma_f64_16x16x4f64(e.x, r.x, P1, 0, 0, 0);
                P2 = __builtin_amdgcn_mfma_f64_16x16x4f64(e.y, r.y, P2, 0, 0, 0);
                P3 = __builtin_amdgcn_mfma_f64_16x16x4f64(e.x + e.y, r.x + r.y, P3, 0, 0, 0);
            }
        }
        wave_sync();  // sG is written again by the next pass
    }

    __syncthreads();  // the last pass's G_cc is dead; sA / sB reuse its memory
    // sum_w g_d over the four slots (lanes j, j + 16, j + 32, j + 48)
    gsum.x += __shfl_xor(gsum.x, 16);
    gsum.y += __shfl_xor(gsum.y, 16);
    gsum.x += __shfl_xor(gsum.x, 32);
    gsum.y += __shfl_xor(gsum.y, 32);
    // Out = sum_w G' from the MFMA result layout: lane jcol + 16 q, register r holds row 4 r + q
    {
        const int q = lane >> 4, jcol = lane & 15;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = 4 * r + q;
            cd v = cd{P1[r] - P2[r], P3[r] - P1[r] - P2[r]};
            if (i == jcol && jd) v = cadd(v, gsum);
            sA[i + 16 * jcol] = v;
        }
    }
    __syncthreads();
    const cd* V = Vd + (size_t)k * nd * nd;  // column-major, column l = eigenvector l of H_dd
    const int nn = n * n;
    // Y = Out T^H: columns NC + e are sum_l Out[:, NC + l] conj(V[e][l])
    for (int el = lane; el < nn; el += 64) {
        const int i = el % n, jj = el / n;
        cd y;
        if (jj < NC) {
            y = sA[i + 16 * jj];
        } else {
            const int e = jj - NC;
            y = cd{0.0, 0.0};
            for (int l = 0; l < nd; ++l) y = cfma(sA[i + 16 * (NC + l)], cconj(V[e + l * nd]), y);
        }
        sB[i + 16 * jj] = y;
    }
    __syncthreads();
    // S = T Y: rows NC + e are sum_l V[e][l] Y[NC + l, :]
    for (int el = lane; el < nn; el += 64) {
        const int i = el % n, jj = el / n;
        cd v;
        if (i < NC) {
            v = sB[i + 16 * jj];
        } else {
            const int e = i - NC;
            v = cd{0.0, 0.0};
            for (int l = 0; l < nd; ++l) v = cfma(V[e + l * nd], sB[NC + l + 16 * jj], v);
        }
        sA[i + 16 * jj] = v;
    }
    __syncthreads();
    // D = (S + S^H)/beta + U diag(w) U^H
    const cd* Uk = U + (size_t)k * nn;
    const double* wk = wt + (size_t)k * n;
    for (int el = lane; el < nn; el += 64) {
        const int i = el % n, jj = el / n;
        cd v = cscale(cadd(sA[i + 16 * jj], cconj(sA[jj + 16 * i])), 1.0 / a.beta);
        for (int l = 0; l < n; ++l) v = cadd(v, cscale(cmulc(Uk[i + l * n], Uk[jj + l * n]), wk[l]));
        dens[(size_t)k * nn + el] = v;
    }
}

template <int NC>
static hipError_t launch_dlr(const LatArgs& a, const cd* W, const cd* Vd, const double* ed, const cd* U,
                             const double* wt, cd* dens, hipStream_t s)
{
    hipLaunchKernelGGL((k_density_lr<NC>), dim3(a.nk), dim3(64), 0, s, a, W, Vd, ed, U, wt, dens);
    return hipGetLastError();
}

bool density_lr_supported(int n, int nc) { return nc >= 1 && nc <= 6 && n >= nc && n <= 16; }

hipError_t launch_density_matrix_lr(const LatArgs& a, const cd* W, const cd* Vd, const double* ed, const cd* U,
                                    const double* wt, cd* dens, hipStream_t s)
{
    if (!density_lr_supported(a.n, a.nc) || !a.kdata) return hipErrorInvalidValue;
    switch (a.nc) {
        case 1: return launch_dlr<1>(a, W, Vd, ed, U, wt, dens, s);
        case 2: return launch_dlr<2>(a, W, Vd, ed, U, wt, dens, s);
        case 3: return launch_dlr<3>(a, W, Vd, ed, U, wt, dens, s);
        case 4: return launch_dlr<4>(a, W, Vd, ed, U, wt, dens, s);
        case 5: return launch_dlr<5>(a, W, Vd, ed, U, wt, dens, s);
        default: return launch_dlr<6>(a, W, Vd, ed, U, wt, dens, s);
    }
}

}  // namespace dg
