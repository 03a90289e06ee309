// Plain FP64 lattice Green's function: one full n x n inverse per (k, iw_n).
//
// Replaces the hot loop of GreenFtn_w (GreenFtn.cpp:218-308):
//     A = (iw_n + mu) I - H(k) - P Sigma_n P^H ,  G = A^{-1}   (:245-261)
// and the k sums that follow it (:264-290, 298-305).
//
// Mapping (CDNA4, wave64):
//   * a "slot" of L lanes (L = 4/8/16/32, L >= n) owns one matrix; lane c holds
//     column c of A in registers (complex128, fully unrolled over N);
//   * a workgroup owns SLOTS consecutive Matsubara frequencies (one per slot)
//     and a contiguous chunk of k points; it walks the chunk, staging H(k) once
//     into LDS (double buffered) and reusing it for all SLOTS frequencies;
//   * Gauss-Jordan elimination without pivoting, at wavefront level: at step j
//     lane j publishes its column through LDS, every lane applies the rank-1
//     update to its own column.  The pivot column is not rescaled in place;
//     lane j instead records the factor -1/A_jj and applies it once at the end,
//     which keeps the step free of divergent branches.  Elimination without
//     pivoting is backward stable here: for causal Sigma the Hermitian part of
//     -iA is >= w_n I (positive definite), the case covered by Golub & Van Loan
//     sec. 4.4 / Higham sec. 10.4; its result agrees with partial-pivot LU to
//     O(kappa * eps) ~ 1e-14;
//   * per-(k,iw) results are summed over the k chunk in LDS/registers and
//     written as per-chunk partials; a second pass adds the chunks in a fixed
//     order, so results are deterministic (no floating-point atomics).
//
// Modes: FWD   G_cc and tr G               (G_loc, NumMatrix, N)
//        DENS  tr G and tr G^2             (N and dN/dmu)
//        VJP   sigma_bar, mu_bar terms and tr G^2
#include "gj.hpp"
#include "lattice_kernels.hpp"

namespace dg {

// Adjoint sums of a 16-lane slot, column c on lane c (lane M's column read by
// DPP in the FMAs themselves):
//   S2_:c += sum_{M < n} G_{:nc, M} G_{M c}     ((G^2)_cc, rows i < nc)
template <int N, int NCR, int M>
__device__ __forceinline__ void s2_from_lanes(const cd (&col)[N], cd (&S2)[NCR], int nc, int n)
{
    if constexpr (M < N) {
        if (M < n) {
            const double mwy = -col[M].y;
#pragma unroll
            for (int i = 0; i < NCR; ++i)
                if (i < nc) cfma_bcast<M>(S2[i], col[i], col[M], mwy);
            s2_from_lanes<N, NCR, M + 1>(col, S2, nc, n);
        }
    }
}

//   XH_:c += sum_{I < nc} G_{:nc, I} z_I      (z = gbar^H G_:c, so XH = G_cc gbar^H G_cc)
template <int N, int NCR, int I>
__device__ __forceinline__ void xh_from_lanes(const cd (&col)[N], const cd (&z)[NCR], cd (&XH)[NCR], int nc)
{
    if constexpr (I < NCR) {
        if (I < nc) {
            const double mzy = -z[I].y;
#pragma unroll
            for (int r = 0; r < NCR; ++r)
                if (r < nc) cfma_bcast<I>(XH[r], col[r], z[I], mzy);
            xh_from_lanes<N, NCR, I + 1>(col, z, XH, nc);
        }
    }
}

template <int N, int L, int MODE, int WG>
__global__ __launch_bounds__(WG) void k_lattice_plain(LatArgs a)
{
    constexpr int SLOTS = WG / L;
    // 16-lane slots with n_c <= NCR keep the adjoint sums in registers: lane
    // c < n_c holds column c of S2 = sum_k (G^2)_cc and of XH = sum_k G_cc
    // gbar^H G_cc; the terms linear in (G^2)_cc (the N constraint, mu_bar)
    // are applied once after the k loop
    constexpr int NCR = (MODE == MODE_VJP && L == 16) ? (N < 6 ? N : 6) : 1;
    constexpr int BS = (MODE == MODE_FWD) ? 2 * N : (N * N > 2 * N ? N * N : 2 * N);
    constexpr int NN = N * N;
    constexpr int PF = (NN + WG - 1) / WG;
    extern __shared__ cd lds[];
    const int n = a.n, nc = a.nc, m2 = nc * nc;
    cd* sH = lds;                      // [2][N*N] column-major, leading dim N
    cd* sB = sH + 2 * NN;              // [SLOTS][BS] broadcast / transpose
    cd* sSig = sB + SLOTS * BS;        // [SLOTS][m2]
    cd* sAcc = sSig + SLOTS * m2;      // [SLOTS][m2]
    cd* sGb = sAcc + SLOTS * m2;       // [SLOTS][m2] (VJP)
    cd* sV = sGb + SLOTS * m2;         // [SLOTS][m2] (VJP)
    const int tid = threadIdx.x, slot = tid / L, c = tid % L;
    const int nw = blockIdx.x * SLOTS + slot;
    const bool wvalid = nw < a.nw_tot;
    const int k0 = blockIdx.y * a.k_per_chunk;
    const int k1 = min(a.nk, k0 + a.k_per_chunk);
    double mun, wn;
    freq_point(a, nw, mun, wn);
    const cd zmu = cd{mun, wn};
    const bool vjp_here = (MODE == MODE_VJP) && wvalid && nw < a.n_freq;

    for (int e = tid; e < 2 * NN; e += WG) sH[e] = cd{0.0, 0.0};
    for (int e = c; e < m2; e += L) {
        const int i = e % nc, j = e / nc;
        sSig[slot * m2 + e] = wvalid ? sigma_elem(a.sigma, a.mom, nc, a.n_freq, a.beta, nw, i, j) : cd{0.0, 0.0};
        sAcc[slot * m2 + e] = cd{0.0, 0.0};
        if (MODE == MODE_VJP) sGb[slot * m2 + e] = vjp_here ? a.gbar[(size_t)nw * m2 + e] : cd{0.0, 0.0};
    }
    __syncthreads();
    if (k0 < k1)
        for (int e = tid; e < n * n; e += WG) sH[(e % n) + (e / n) * N] = a.Hk[(size_t)k0 * n * n + e];

    cd trA = cd{0.0, 0.0}, tr2A = cd{0.0, 0.0}, muA = cd{0.0, 0.0};
    cd* bc = sB + slot * BS;
    const bool vjp_reg = (MODE == MODE_VJP) && L == 16 && nc <= NCR;
    cd S2[NCR], XH[NCR];
#pragma unroll
    for (int i = 0; i < NCR; ++i) S2[i] = XH[i] = cd{0.0, 0.0};
    __syncthreads();

    for (int k = k0; k < k1; ++k) {
        const int cur = (k - k0) & 1;
        // prefetch H(k+1) into registers
        cd pre[PF];
        const bool more = (k + 1) < k1;
#pragma unroll
        for (int r = 0; r < PF; ++r) {
            const int e = tid + r * WG;
            pre[r] = (more && e < n * n) ? a.Hk[(size_t)(k + 1) * n * n + e] : cd{0.0, 0.0};
        }
        // column c of A = (iw+mu) I - H - P Sigma P^H
        cd col[N];
        const cd* Hc = sH + cur * NN + (c < N ? c : 0) * N;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            cd h = Hc[i];
            col[i] = (c < n) ? cneg(h) : cd{0.0, 0.0};
            if (i == c) col[i] = (c < n) ? cadd(col[i], zmu) : cd{1.0, 0.0};
        }
        if (c < nc) {
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (i < nc) col[i] = csub(col[i], sSig[slot * m2 + i + c * nc]);
        }
        if constexpr (L == 16) gj_columns_row16<N>(col, c, n);
        else gj_columns<N>(col, bc, c, n);
        // lane c < n now holds column c of G
        cd gcc = cd{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (i == c) gcc = col[i];
        if (c < n && wvalid) trA = cadd(trA, gcc);

        if (MODE == MODE_FWD) {
            if (c < nc) {
#pragma unroll
                for (int i = 0; i < N; ++i)
                    if (i < nc) lds_add(&sAcc[slot * m2 + i + c * nc], col[i]);
            }
        } else {
            wave_sync();
            if (c < n) {
#pragma unroll
                for (int i = 0; i < N; ++i) bc[i + c * N] = col[i];
            }
            wave_sync();
            if (c < n) {  // (G^2)_cc = sum_m G_cm G_mc
                cd s = cd{0.0, 0.0};
#pragma unroll
                for (int m = 0; m < N; ++m)
                    if (m < n) s = cfma(bc[c + m * N], col[m], s);
                if (wvalid) tr2A = cadd(tr2A, s);
            }
            if constexpr (MODE == MODE_VJP && L == 16) {
                if (vjp_reg && vjp_here) {  // every lane of the slot takes part
                    s2_from_lanes<N, NCR, 0>(col, S2, nc, n);
                    cd z[NCR];
#pragma unroll
                    for (int i = 0; i < NCR; ++i) {
                        z[i] = cd{0.0, 0.0};
#pragma unroll
                        for (int j = 0; j < NCR; ++j)
                            if (i < nc && j < nc) z[i] = cfmac(sGb[slot * m2 + j + i * nc], col[j], z[i]);
                    }
                    xh_from_lanes<N, NCR, 0>(col, z, XH, nc);
                }
            }
            if (MODE == MODE_VJP && !vjp_reg) {
                if (vjp_here && c < nc) {
                    // v_i = sum_j gbar_ij conj(G_cj)  (row c of G_cc)
                    for (int i = 0; i < nc; ++i) {
                        cd v = cd{0.0, 0.0};
                        for (int j = 0; j < nc; ++j) v = cfma(sGb[slot * m2 + i + j * nc], cconj(bc[c + j * N]), v);
                        sV[slot * m2 + i + c * nc] = v;
                    }
                    const double cn = 2.0 * a.nbar / a.beta;
                    cd mu_s = cd{0.0, 0.0};
                    for (int aa = 0; aa < nc; ++aa) {
                        // X_{a c} = sum_i conj(G_ia) v_i ;  r_a = (G^2)_{c a} = sum_m G_cm G_ma
                        cd x = cd{0.0, 0.0};
                        for (int i = 0; i < nc; ++i) x = cfmac(bc[i + aa * N], sV[slot * m2 + i + c * nc], x);
                        cd r = cd{0.0, 0.0};
                        for (int m = 0; m < n; ++m) r = cfma(bc[c + m * N], bc[m + aa * N], r);
                        // sigma_bar(a, c) += X_ac + cn * conj((G^2)_{c a})
                        cd* dst = &sAcc[slot * m2 + aa + c * nc];
                        *dst = cadd(*dst, cadd(x, cscale(cconj(r), cn)));
                        // tr(gbar^H G2) row c: conj(gbar_{c a}) (G^2)_{c a}
                        mu_s = cfmac(sGb[slot * m2 + c + aa * nc], r, mu_s);
                    }
                    muA = cadd(muA, mu_s);
                }
            }
            wave_sync();
        }
        // publish prefetched H(k+1)
        if (more) {
#pragma unroll
            for (int r = 0; r < PF; ++r) {
                const int e = tid + r * WG;
                if (e < n * n) sH[(cur ^ 1) * NN + (e % n) + (e / n) * N] = pre[r];
            }
        }
        __syncthreads();
    }

    if (MODE == MODE_VJP && vjp_reg) {
        // sigma_bar(c, r) = conj(XH_rc + cn S2_rc); mu_bar's tr(gbar^H S2) by columns
        const double cn = 2.0 * (a.nbar_dev ? *a.nbar_dev : a.nbar) / a.beta;
        cd mu_s = cd{0.0, 0.0};
        if (c < nc) {
#pragma unroll
            for (int r = 0; r < NCR; ++r)
                if (r < nc) {
                    sAcc[slot * m2 + c + r * nc] = cconj(cadd(XH[r], cscale(S2[r], cn)));
                    mu_s = cfmac(sGb[slot * m2 + r + c * nc], S2[r], mu_s);
                }
        }
        muA = mu_s;
        wave_sync();  // sAcc is stored below by other lanes of the slot
    }
    // reduce lane partials of the slot (lanes of a slot are contiguous)
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) {
        trA.x += __shfl_xor(trA.x, o);
        trA.y += __shfl_xor(trA.y, o);
        if (MODE != MODE_FWD) {
            tr2A.x += __shfl_xor(tr2A.x, o);
            tr2A.y += __shfl_xor(tr2A.y, o);
        }
        if (MODE == MODE_VJP) {
            muA.x += __shfl_xor(muA.x, o);
            muA.y += __shfl_xor(muA.y, o);
        }
    }
    if (!wvalid) return;
    const size_t ch = blockIdx.y;
    if (c == 0) {
        a.pTr[ch * a.nw_tot + nw] = trA;
        if (MODE != MODE_FWD) a.pTr2[ch * a.nw_tot + nw] = tr2A;
        if (MODE == MODE_VJP && nw < a.n_freq) a.pMu[ch * a.n_freq + nw] = muA;
    }
    if (MODE == MODE_FWD) {
        for (int e = c; e < m2; e += L) a.pG[(ch * a.nw_tot + nw) * m2 + e] = sAcc[slot * m2 + e];
    } else if (MODE == MODE_VJP && nw < a.n_freq) {
        for (int e = c; e < m2; e += L) a.pSb[(ch * a.n_freq + nw) * m2 + e] = sAcc[slot * m2 + e];
    }
}


// densityMatDFT (GreenFtn.cpp:291-319): one workgroup per k walks all
// high*N_freq frequencies, SLOTS at a time, summing G(k, iw) column-wise in
// registers; then
//   D(k) = (S + S^H)/beta + U diag(w) U^H,   S = sum_n G(k, iw_n),
// where w_l = f(eps_l - mu) - (1/beta) sum_n 2 Re 1/(iw_n + mu - eps_l) folds
// the G0 subtraction and the Fermi add-back (the static weights).
template <int N, int L>
__global__ __launch_bounds__(256) void k_density_plain(LatArgs a, const cd* __restrict__ U,
                                                       const double* __restrict__ wt, cd* __restrict__ dens)
{
    constexpr int WG = 256;
    constexpr int SLOTS = WG / L;
    constexpr int NN = N * N;
    extern __shared__ cd lds[];
    const int n = a.n, nc = a.nc, m2 = nc * nc;
    cd* sH = lds;                   // [N*N]
    cd* sB = sH + NN;               // [SLOTS][2N]
    cd* sS = sB + SLOTS * 2 * N;    // [SLOTS][N*N] slot sums
    const int tid = threadIdx.x, slot = tid / L, c = tid % L;
    const int k = blockIdx.x;
    for (int e = tid; e < NN; e += WG) sH[e] = cd{0.0, 0.0};
    __syncthreads();
    for (int e = tid; e < n * n; e += WG) sH[(e % n) + (e / n) * N] = a.Hk[(size_t)k * n * n + e];
    __syncthreads();
    cd acc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = cd{0.0, 0.0};
    cd* bc = sB + slot * 2 * N;
    const cd* Hc = sH + (c < N ? c : 0) * N;
    for (int nw0 = 0; nw0 < a.nw_tot; nw0 += SLOTS) {
        const int nw = nw0 + slot;
        const bool wvalid = nw < a.nw_tot;
        const int nwc = wvalid ? nw : a.nw_tot - 1;
        const cd zmu = cd{a.mu, matsubara(nwc, a.beta)};
        cd col[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            col[i] = (c < n) ? cneg(Hc[i]) : cd{0.0, 0.0};
            if (i == c) col[i] = (c < n) ? cadd(col[i], zmu) : cd{1.0, 0.0};
        }
        if (c < nc) {
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (i < nc) col[i] = csub(col[i], sigma_elem(a.sigma, a.mom, nc, a.n_freq, a.beta, nwc, i, c));
        }
        if constexpr (L == 16) gj_columns_row16<N>(col, c, n);
        else gj_columns<N>(col, bc, c, n);
        if (wvalid && c < n) {
#pragma unroll
            for (int i = 0; i < N; ++i) acc[i] = cadd(acc[i], col[i]);
        }
        wave_sync();
    }
    if (c < N) {
#pragma unroll
        for (int i = 0; i < N; ++i) sS[slot * NN + i + c * N] = acc[i];
    }
    __syncthreads();
    const cd* Uk = U + (size_t)k * n * n;
    const double* w = wt + (size_t)k * n;
    for (int e = tid; e < n * n; e += WG) {
        const int i = e % n, j = e / n;
        cd sij = cd{0.0, 0.0}, sji = cd{0.0, 0.0};
        for (int q = 0; q < SLOTS; ++q) {
            sij = cadd(sij, sS[q * NN + i + j * N]);
            sji = cadd(sji, sS[q * NN + j + i * N]);
        }
        cd v = cscale(cadd(sij, cconj(sji)), 1.0 / a.beta);
        for (int l = 0; l < n; ++l) v = cadd(v, cscale(cmulc(Uk[i + l * n], Uk[j + l * n]), w[l]));
        dens[(size_t)k * n * n + e] = v;
    }
    (void)m2;
}

template <int N>
static hipError_t launch_dens_n(const LatArgs& a, const cd* U, const double* wt, cd* dens, hipStream_t s)
{
    constexpr int L = N <= 4 ? 4 : N <= 8 ? 8 : N <= 16 ? 16 : 32;
    constexpr int SLOTS = 256 / L;
    size_t shm = sizeof(cd) * ((size_t)N * N + (size_t)SLOTS * 2 * N + (size_t)SLOTS * N * N);
    hipError_t e = hipFuncSetAttribute((const void*)k_density_plain<N, L>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)shm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_density_plain<N, L>), dim3(a.nk), dim3(256), shm, s, a, U, wt, dens);
    return hipGetLastError();
}

hipError_t launch_density_matrix(const LatArgs& a, const cd* U, const double* wt, cd* dens, hipStream_t s)
{
    switch (a.n) {
#define CASE(NV) \
    case NV: return launch_dens_n<NV>(a, U, wt, dens, s);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13)
        CASE(14) CASE(15) CASE(16)
#undef CASE
        default: break;
    }
    if (a.n <= 20) return launch_dens_n<20>(a, U, wt, dens, s);
    if (a.n <= 24) return launch_dens_n<24>(a, U, wt, dens, s);
    if (a.n <= 28) return launch_dens_n<28>(a, U, wt, dens, s);
    if (a.n <= 32) return launch_dens_n<32>(a, U, wt, dens, s);
    return hipErrorInvalidValue;
}

static int lanes_for(int N) { return N <= 4 ? 4 : N <= 8 ? 8 : N <= 16 ? 16 : 32; }
static int wg_for(int N, int mode) { return (mode == MODE_FWD || N <= 8) ? 256 : 128; }

int plain_slots(int n, int mode)
{
    return wg_for(n <= 16 ? n : 32, mode) / lanes_for(n);
}

template <int N, int MODE>
static hipError_t launch_n(const LatArgs& a, hipStream_t s, int* wtile)
{
    constexpr int L = N <= 4 ? 4 : N <= 8 ? 8 : N <= 16 ? 16 : 32;
    constexpr int WG = (MODE == MODE_FWD || N <= 8) ? 256 : 128;
    constexpr int SLOTS = WG / L;
    // 16-lane slots with n_c <= NCR keep the adjoint sums in registers: lane
    // c < n_c holds column c of S2 = sum_k (G^2)_cc and of XH = sum_k G_cc
    // gbar^H G_cc; the terms linear in (G^2)_cc (the N constraint, mu_bar)
    // are applied once after the k loop
    constexpr int NCR = (MODE == MODE_VJP && L == 16) ? (N < 6 ? N : 6) : 1;
    constexpr int BS = (MODE == MODE_FWD) ? 2 * N : (N * N > 2 * N ? N * N : 2 * N);
    const int m2 = a.nc * a.nc;
    size_t shm = sizeof(cd) * (2 * (size_t)N * N + (size_t)SLOTS * BS + (size_t)SLOTS * m2 * (MODE == MODE_VJP ? 4 : 2));
    dim3 grid((a.nw_tot + SLOTS - 1) / SLOTS, a.n_chunks);
    if (wtile) *wtile = SLOTS;
    hipError_t e = hipFuncSetAttribute((const void*)k_lattice_plain<N, L, MODE, WG>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_lattice_plain<N, L, MODE, WG>), grid, dim3(WG), shm, s, a);
    return hipGetLastError();
}

template <int MODE>
static hipError_t dispatch(const LatArgs& a, hipStream_t s, int* wtile)
{
    switch (a.n) {
#define CASE(NV) \
    case NV: return launch_n<NV, MODE>(a, s, wtile);
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12) CASE(13)
        CASE(14) CASE(15) CASE(16)
#undef CASE
        default: break;
    }
    // sizes above 16 run in padded buckets (steps j >= n are skipped)
    if (a.n <= 20) return launch_n<20, MODE>(a, s, wtile);
    if (a.n <= 24) return launch_n<24, MODE>(a, s, wtile);
    if (a.n <= 28) return launch_n<28, MODE>(a, s, wtile);
    if (a.n <= 32) return launch_n<32, MODE>(a, s, wtile);
    return hipErrorInvalidValue;
}

hipError_t launch_lattice_plain(const LatArgs& a, int mode, hipStream_t s, int* wtile_out)
{
    if (mode == MODE_FWD) return dispatch<MODE_FWD>(a, s, wtile_out);
    if (mode == MODE_DENS) return dispatch<MODE_DENS>(a, s, wtile_out);
    return dispatch<MODE_VJP>(a, s, wtile_out);
}

}  // namespace dg
