// Schur-complement lattice path for larger correlated blocks (7 <= n_c <= 16),
// column per lane.
//
// Same mathematics as lattice_schur.hip (the hot loop of GreenFtn_w,
// GreenFtn.cpp:218-308, and the densities of GiwToNele :513-635, from the
// block structure of A = (z+mu) I - H(k) - P Sigma P^H with the correlated
// orbitals first, downfolding.cpp:69-79):
//   S     = (z+mu) I - H_cc - Sigma - sum_d g_d w_d w_d^H,  g_d = 1/(z+mu-e_d)
//   G_cc  = S^{-1}
//   tr G  = sum_d g_d + tr G_cc + sum_d g_d^2 M_d,        M_d = w_d^H G_cc w_d
//   tr G^2 = sum_d g_d^2 + tr(Z^2) + 2 sum_d g_d^3 M_d,   Z = G_cc (I + Y),
//            Y = sum_d g_d^2 w_d w_d^H
//   (G^2)_cc = Z G_cc
// An n_c x n_c complex matrix no longer fits one lane's registers, so a slot
// of N lanes (N = n_c rounded up to 8/10/12/16) owns one (k, iw) matrix and
// lane c holds column c of S, then of G_cc:
//   * S column c:  S_ic = E_ic - H_cc,ic - sum_d W_id (g_d conj(W_cd)),
//     E = (z+mu) I - Sigma(iw) kept in registers across the k loop;
//     the g_d of the slot's frequency are computed once per slot and shared
//     through LDS;
//   * inversion: column-per-lane Gauss-Jordan (gj.hpp); FP32R inverts in FP32
//     and refines each column against the FP64 S held in LDS,
//     x_c <- x_c + X0 (e_c - S x_c);
//   * FWD accumulates column c of G_cc over the k chunk in registers; the
//     trace terms are per-lane partials summed over the slot at the end;
//   * DENS/VJP exchange G_cc, Z and V = gbar + cn (I + Y)^H through per-slot
//     LDS matrices for the products Z = G + G Y, tr(Z^2) and
//     sigma_bar += G^H V G^H; lane c accumulates column c of sigma_bar.
// The per-k data (H_cc, W = H_cd V_d, e_d) is read from global memory through
// the constant address space (wave-uniform reads become scalar loads).
#pragma once
#include "gj.hpp"
#include "lattice_kernels.hpp"

namespace dg {

// Smallest n_c the bucket of width N serves (lattice_schur.hip, cols_bucket):
// rows i below it are inside every served block, so in the d loops the tests
// i < n_c on them fold away at compile time and the W loads issue
// unconditionally (a load under a uniform branch waits on its own).  The
// adjoint's LDS products keep their tests: unconditional, their reads would
// all be hoisted and spill.
__host__ __device__ constexpr int cols_nc_min(int N)
{
    return N <= 4 ? 1 : N <= 6 ? N : N == 8 ? 7 : N == 10 ? 9 : N == 12 ? 11 : 13;
}

// one wavefront per workgroup: the per-k data comes from global memory, so
// wavefronts share nothing but the cache
__host__ __device__ constexpr int cols_wg(int, int) { return 64; }

// per-slot LDS (cd units): bc 2N (GJ broadcast), gsh 32 (g_d), MA N*N, MB N*N, GB N*N.
// The FP32R forward keeps only X0 (FP32) in MB: N*N/2 cd, so that at N = 10 a
// wavefront needs 22.6 KB instead of 28.2 KB and 7 fit in a CU's LDS instead of 5.
__host__ __device__ constexpr int cols_slot_size(int N, int mode, int prec)
{
    const int mb = mode == MODE_FWD ? (prec > 0 ? (N * N + 1) / 2 : 0) : N * N;
    return 2 * N + 32 + ((prec > 0 || mode != MODE_FWD) ? N * N + mb : 0) + (mode == MODE_VJP ? N * N : 0);
}

__host__ __device__ constexpr int cols_slots(int N, int mode, int prec) { return cols_wg(mode, prec) / 64 * (64 / N); }

inline size_t cols_lds_bytes(int N, int mode, int prec, int, int)
{
    const int nwave = cols_wg(mode, prec) / 64;
    const int spw = 64 / N;
    return sizeof(cd) * ((size_t)nwave * (spw + 1) * cols_slot_size(N, mode, prec));
}

// One (frequency tile, k chunk) work item of one lattice: the body of the
// per-lattice kernel and of the persistent multi-lattice kernel below.
template <int N, int MODE, int PREC, bool TRACE>
__device__ __forceinline__ void cols_item(const LatArgs& a, const cd* __restrict__ sig_all, int wtile, int chunk,
                                          cd* lds)
{
    constexpr int SPW = 64 / N;
    constexpr int SLOTS = cols_slots(N, MODE, PREC);
    constexpr int SS = cols_slot_size(N, MODE, PREC);
    constexpr bool MATS = (PREC > 0 || MODE != MODE_FWD);
    const int n = a.n, nc = a.nc, nd = a.n - nc, m2 = nc * nc;
    constexpr int NMIN = cols_nc_min(N);
    auto inb = [nc](int i) { return i < NMIN || i < nc; };  // row / column i inside the n_c block
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int sl = lane / N;  // == SPW for the spare lanes of a wavefront
    const int c = lane - sl * N;
    const bool active = sl < SPW;
    const bool live = active && c < nc;  // lanes c >= nc hold identity padding
    const int cc = live ? c : 0;         // clamped column index for loads
    cd* my = lds + (size_t)(wave * (SPW + 1) + sl) * SS;  // spare lanes get a scratch slot
    cd* bc = my;
    cd* gsh = my + 2 * N;
    cd* MA = gsh + 32;
    cd* MB = MA + N * N;
    cd* GB = MB + N * N;
    const int nw = wtile * SLOTS + wave * SPW + (active ? sl : 0);
    const bool wvalid = active && nw < a.nw_tot;
    const int nwc = nw < a.nw_tot ? nw : a.nw_tot - 1;
    const int k0 = chunk * a.k_per_chunk;
    const int k1 = min(a.nk, k0 + a.k_per_chunk);
    double mu, w;
    freq_point(a, nwc, mu, w);
    const bool vjp_here = (MODE == MODE_VJP) && wvalid && nw < a.n_freq;
    const double cn = (MODE == MODE_VJP) ? 2.0 * (a.nbar_dev ? *a.nbar_dev : a.nbar) / a.beta : 0.0;

    // E = (z+mu) I - Sigma(iw), column c
    cd E[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        cd v = cd{0.0, 0.0};
        if (live && i < nc) v = cneg(sig_all[(size_t)nwc * m2 + i + cc * nc]);
        if (live && i == c) v = cadd(v, cd{mu, w});
        E[i] = v;
    }
    if (MODE == MODE_VJP) {
#pragma unroll
        for (int i = 0; i < N; ++i)
            GB[i + c * N] = (vjp_here && live && i < nc) ? a.gbar[(size_t)nw * m2 + i + cc * nc] : cd{0.0, 0.0};
    }
    cd acc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = cd{0.0, 0.0};
    cd trA = cd{0.0, 0.0}, tr2A = cd{0.0, 0.0}, muA = cd{0.0, 0.0};
    __syncthreads();

    for (int k = k0; k < k1; ++k) {
        // per-k data straight from global memory through the constant address
        // space: the wave-uniform reads (W_id, e_d) become scalar loads
        const kconst_cd* Hkk = (const kconst_cd*)(a.Hk + (size_t)k * n * n);  // H_cc(i, c) at i + c*n
        const kconst_cd* Wk = (const kconst_cd*)(a.W + (size_t)k * nc * nd);
        const kconst_double* ed = (const kconst_double*)(a.ed + (size_t)k * nd);
        // This k's W and e_d into the caches at once: vector loads into L1 for
        // the lane-dependent reads (W_cd, e_d), the scalar-cache fill for the
        // wave-uniform ones (columns of W), one round trip per k instead of
        // one per d
        if (nd > 0) {
            const cd* Wg = a.W + (size_t)k * nc * nd;
            const int wn = nc * nd;  // complex entries of W(k)
            // indices clamped, not branched on: the loads issue together
            double pf = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int e = tid + r * 64;
                pf = fma(Wg[e < wn ? e : wn - 1].x, 0.0, pf);
            }
            pf = fma(a.ed[(size_t)k * nd + (tid < nd ? tid : nd - 1)], 0.0, pf);
            kcache_fill64(Wg);
            asm volatile("" ::"v"(pf));
        }
        // g_d = 1/(iw + mu - e_d) of this slot's frequency, shared by its lanes
        for (int d = c; d < nd; d += N) {
            const double x = mu - ed[d];
            const double r = rcp_pos(fma(x, x, w * w));
            gsh[d] = cd{x * r, -w * r};
        }
        wave_sync();

        // ---- column c of S (identity column for padding lanes)
        cd col[N];
        cd y[MODE == MODE_FWD ? 1 : N];  // column c of Y = sum_d g_d^2 w_d w_d^H
        // (the loads are unconditional, rows clamped in range, and pinned by an
        // empty asm so that the compiler cannot sink each into a branch of its
        // own: they issue together instead of one round trip per entry)
        cd hcol[N];
#pragma unroll
        for (int i = 0; i < N; ++i) hcol[i] = ldc(Hkk + (i < nc ? i : 0) + (size_t)cc * n);
#pragma unroll
        for (int i = 0; i < N; ++i) asm volatile("" ::"v"(hcol[i].x), "v"(hcol[i].y));
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const cd pad = (i == c) ? cd{1.0, 0.0} : cd{0.0, 0.0};
            col[i] = live ? ((i < nc) ? csub(E[i], hcol[i]) : cd{0.0, 0.0}) : pad;
        }
        if constexpr (MODE != MODE_FWD) {
#pragma unroll
            for (int i = 0; i < N; ++i) y[i] = cd{0.0, 0.0};
        }
        for (int d = 0; d < nd; ++d) {
            const cd g = gsh[d];
            const kconst_cd* Wd = Wk + d * nc;
            const cd u = live ? cmulc(g, ldc(Wd + cc)) : cd{0.0, 0.0};  // g_d conj(W_cd)
#pragma unroll
            for (int i = 0; i < N; ++i)
                if (inb(i)) col[i] = cfms(ldc(Wd + i), u, col[i]);
            if constexpr (MODE != MODE_FWD) {
                const cd u2 = cmul(g, u);
#pragma unroll
                for (int i = 0; i < N; ++i)
                    if (inb(i)) y[i] = cfma(ldc(Wd + i), u2, y[i]);
            }
        }

        // ---- G_cc = S^{-1}, column c
        if constexpr (PREC == 0) {
            gj_columns<N>(col, bc, c, nc);
        } else {
            cf* X0 = (cf*)MB;
#pragma unroll
            for (int i = 0; i < N; ++i) MA[i + c * N] = col[i];
            cf x0[N];
#pragma unroll
            for (int i = 0; i < N; ++i) x0[i] = cf{(float)col[i].x, (float)col[i].y};
            gj_columns_f32<N>(x0, (cf*)bc, c, nc);
#pragma unroll
            for (int i = 0; i < N; ++i) X0[i + c * N] = x0[i];
            wave_sync();
#pragma unroll
            for (int i = 0; i < N; ++i) col[i] = cd{(double)x0[i].x, (double)x0[i].y};
            const int steps = a.refine < 1 ? 1 : (a.refine > 2 ? 2 : a.refine);
#pragma unroll 1
            for (int it = 0; it < steps; ++it) {
                wave_sync();  // keeps the S and X0 reads inside the step (no hoisting into registers)
                cf r[N];  // e_c - S x_c in FP64, stored in FP32
                // each row of S (then of X0) is read from LDS as a whole before
                // its products, pinned by empty asm statements: left to itself the
                // compiler issues one read, waits for it and uses it, one LDS
                // round trip per entry
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    cd srow[N];
#pragma unroll
                    for (int l = 0; l < N; ++l) srow[l] = MA[i + l * N];
#pragma unroll
                    for (int l = 0; l < N; ++l) asm volatile("" ::"v"(srow[l].x), "v"(srow[l].y));
                    cd s = (i == c) ? cd{1.0, 0.0} : cd{0.0, 0.0};
#pragma unroll
                    for (int l = 0; l < N; ++l) s = cfms(srow[l], col[l], s);
                    r[i] = cf{(float)s.x, (float)s.y};
                }
#pragma unroll
                for (int i = 0; i < N; ++i) {
                    cf xrow[N];
#pragma unroll
                    for (int l = 0; l < N; ++l) xrow[l] = X0[i + l * N];
#pragma unroll
                    for (int l = 0; l < N; ++l) asm volatile("" ::"v"(xrow[l].x), "v"(xrow[l].y));
                    cf t = cf{0.f, 0.f};
#pragma unroll
                    for (int l = 0; l < N; ++l) t = ffma(xrow[l], r[l], t);
                    col[i] = cadd(col[i], cd{(double)t.x, (double)t.y});
                }
            }
            wave_sync();
        }
        cd gdiag = cd{0.0, 0.0};  // G_cc(c, c)
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (i == c) gdiag = col[i];

        if constexpr (MODE == MODE_FWD) {
#pragma unroll
            for (int i = 0; i < N; ++i) acc[i] = cadd(acc[i], col[i]);
            if constexpr (TRACE) {
                // lane share of tr G: G_cc + sum_d g_d^2 W_cd (sum_i conj(W_id) G_ic)
                cd t = live ? gdiag : cd{0.0, 0.0};
                for (int d = 0; d < nd; ++d) {
                    const cd g = gsh[d];
                    const kconst_cd* Wd = Wk + d * nc;
                    cd v = cd{0.0, 0.0};
#pragma unroll
                    for (int i = 0; i < N; ++i)
                        if (inb(i)) v = cfmac(ldc(Wd + i), col[i], v);
                    const cd m = live ? cmul(ldc(Wd + cc), v) : cd{0.0, 0.0};
                    t = cfma(cmul(g, g), m, t);
                    if (c == 0) t = cadd(t, g);
                }
                if (wvalid) trA = cadd(trA, t);
            }
        } else {
#pragma unroll
            for (int i = 0; i < N; ++i) MA[i + c * N] = col[i];
            wave_sync();
            // Z = G (I + Y), column c
            cd z[N];
#pragma unroll
            for (int i = 0; i < N; ++i) z[i] = col[i];
#pragma unroll
            for (int l = 0; l < N; ++l) {
                if (l < nc) {
                    const cd yl = y[l];
#pragma unroll
                    for (int i = 0; i < N; ++i) z[i] = cfma(MA[i + l * N], yl, z[i]);
                }
            }
            // trace shares: tr G and tr G^2 = sum g^2 + tr(Z^2) + 2 sum g^3 M_d
            cd tg = live ? gdiag : cd{0.0, 0.0};
            cd t2 = cd{0.0, 0.0};
            for (int d = 0; d < nd; ++d) {
                const cd g = gsh[d];
                const cd g2 = cmul(g, g);
                const kconst_cd* Wd = Wk + d * nc;
                cd v = cd{0.0, 0.0};
#pragma unroll
                for (int i = 0; i < N; ++i)
                    if (inb(i)) v = cfmac(ldc(Wd + i), col[i], v);
                const cd m = live ? cmul(ldc(Wd + cc), v) : cd{0.0, 0.0};
                tg = cfma(g2, m, tg);
                t2 = cfma(cscale(cmul(g2, g), 2.0), m, t2);
                if (c == 0) {
                    tg = cadd(tg, g);
                    t2 = cadd(t2, g2);
                }
            }
#pragma unroll
            for (int i = 0; i < N; ++i) MB[i + c * N] = z[i];
            wave_sync();
            cd zz = cd{0.0, 0.0};  // lane share of tr(Z^2): sum_i Z_ic Z_ci
            if (live) {
#pragma unroll
                for (int i = 0; i < N; ++i)
                    if (i < nc) zz = cfma(z[i], MB[c + i * N], zz);
            }
            t2 = cadd(t2, zz);
            if (wvalid) {
                trA = cadd(trA, tg);
                tr2A = cadd(tr2A, t2);
            }
            if constexpr (MODE == MODE_VJP) {
                // sigma_bar += G^H gbar G^H + cn ((G^2)_cc)^H = C = G^H V G^H with
                // V = gbar + cn S'^H, S' = I + Y ((G^2)_cc = G S' G); lane c forms
                // column c of C.  mu term tr(gbar^H (G^2)_cc) = sum_ab S'_ab conj(C_ab)
                // - cn tr(Z^2) (as in lattice_schur.hip).
                wave_sync();  // every lane has read Z from MB
                if (live) {  // row c of V into MB
#pragma unroll
                    for (int j = 0; j < N; ++j)
                        if (j < nc)
                            MB[c + j * N] = cadd(GB[c + j * N], cscale(cadd(cconj(y[j]), cd{j == c ? 1.0 : 0.0, 0.0}), cn));
                }
                wave_sync();
                if (vjp_here && live) {
                    cd l[N];  // column c of L = V G^H: l_i = sum_j V_ij conj(G_cj)
#pragma unroll
                    for (int i = 0; i < N; ++i) l[i] = cd{0.0, 0.0};
#pragma unroll
                    for (int j = 0; j < N; ++j) {
                        if (j < nc) {
                            const cd gc = cconj(MA[c + j * N]);
#pragma unroll
                            for (int i = 0; i < N; ++i)
                                if (i < nc) l[i] = cfma(MB[i + j * N], gc, l[i]);
                        }
                    }
                    cd ms = cscale(zz, -cn);
#pragma unroll
                    for (int aa = 0; aa < N; ++aa) {
                        if (aa < nc) {  // C_ac = sum_i conj(G_ia) l_i
                            cd s = cd{0.0, 0.0};
#pragma unroll
                            for (int i = 0; i < N; ++i)
                                if (i < nc) s = cfmac(MA[i + aa * N], l[i], s);
                            acc[aa] = cadd(acc[aa], s);
                            const cd sp = aa == c ? cadd(y[aa], cd{1.0, 0.0}) : y[aa];  // S'_ac
                            ms = cfmac(s, sp, ms);
                        }
                    }
                    muA = cadd(muA, ms);
                }
            }
            wave_sync();
        }
        wave_sync();  // gsh of this k is read by every lane of the slot before it is rewritten
    }

    // ---- sum the trace shares over the lanes of each slot
    my[c] = trA;
    my[N + c] = tr2A;
    my[2 * N + c] = muA;
    wave_sync();
    if (!wvalid) return;
    const size_t ch = chunk;
    if (c == 0) {
        cd s1 = cd{0.0, 0.0}, s2 = cd{0.0, 0.0}, s3 = cd{0.0, 0.0};
        for (int l = 0; l < N; ++l) {
            s1 = cadd(s1, my[l]);
            s2 = cadd(s2, my[N + l]);
            s3 = cadd(s3, my[2 * N + l]);
        }
        a.pTr[ch * a.nw_tot + nw] = s1;
        if (MODE != MODE_FWD) a.pTr2[ch * a.nw_tot + nw] = s2;
        if (MODE == MODE_VJP && nw < a.n_freq) a.pMu[ch * a.n_freq + nw] = s3;
    }
    if (!live) return;
    if constexpr (MODE == MODE_FWD) {
#pragma unroll
        for (int i = 0; i < N; ++i)
            if (i < nc) a.pG[(ch * a.nw_tot + nw) * m2 + i + c * nc] = acc[i];
    } else if constexpr (MODE == MODE_VJP) {
        if (nw < a.n_freq) {
#pragma unroll
            for (int b = 0; b < N; ++b)
                if (b < nc) a.pSb[(ch * a.n_freq + nw) * m2 + b + c * nc] = acc[b];  // column c
        }
    }
    (void)MATS;
}

template <int N, int MODE, int PREC, bool TRACE>
__global__ __launch_bounds__(cols_wg(MODE, PREC)) void k_lattice_schur_cols(LatArgs a, const cd* __restrict__ sig_all)
{
    extern __shared__ cd lds[];
    cols_item<N, MODE, PREC, TRACE>(a, sig_all, blockIdx.x, blockIdx.y, lds);
}

// Persistent form over a work list spanning several lattices with n_c in this
// bucket: workgroup b takes items b, b + gridDim.x, ... ({system, frequency
// tile, k chunk}; static striding, no atomics, so every replay does the same
// work in the same order).  Each system's tables come from its LatArgs.
template <int N, int MODE, int PREC, bool TRACE>
__global__ __launch_bounds__(cols_wg(MODE, PREC)) void k_lattice_schur_cols_multi(const LatArgs* __restrict__ args,
                                                                                 const int4* __restrict__ items,
                                                                                 int n_items)
{
    extern __shared__ cd lds[];
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const int4 item = items[it];  // {system, wtile, chunk, -}
        const LatArgs a = args[item.x];
        cols_item<N, MODE, PREC, TRACE>(a, a.pSigAll, item.y, item.z, lds);
        __syncthreads();
    }
}

template <int N, int MODE, int PREC, bool TRACE>
hipError_t launch_cols_k(const LatArgs& a, const cd* sig_all, hipStream_t s, const SchurMulti* m)
{
    if (a.nc < cols_nc_min(N) || a.nc > N) return hipErrorInvalidValue;  // the kernels assume this bucket
    const int nd = a.n - a.nc;
    const size_t shm = cols_lds_bytes(N, MODE, PREC, a.nc, nd);
    constexpr int SLOTS = cols_slots(N, MODE, PREC);
    if (m) {
        hipError_t e = hipFuncSetAttribute((const void*)k_lattice_schur_cols_multi<N, MODE, PREC, TRACE>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_lattice_schur_cols_multi<N, MODE, PREC, TRACE>), dim3(m->grid),
                           dim3(cols_wg(MODE, PREC)), shm, s, m->args, m->items, m->n_items);
        return hipGetLastError();
    }
    hipError_t e = hipFuncSetAttribute((const void*)k_lattice_schur_cols<N, MODE, PREC, TRACE>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
    dim3 grid((a.nw_tot + SLOTS - 1) / SLOTS, a.n_chunks);
    hipLaunchKernelGGL((k_lattice_schur_cols<N, MODE, PREC, TRACE>), grid, dim3(cols_wg(MODE, PREC)), shm, s, a,
                       sig_all);
    return hipGetLastError();
}

template <int N, int PREC>
hipError_t launch_cols_p(const LatArgs& a, int mode, const cd* sig_all, hipStream_t s, const SchurMulti* m)
{
    if (mode == MODE_FWD)
        return a.want_trace ? launch_cols_k<N, MODE_FWD, PREC, true>(a, sig_all, s, m)
                            : launch_cols_k<N, MODE_FWD, PREC, false>(a, sig_all, s, m);
    if (mode == MODE_DENS) return launch_cols_k<N, MODE_DENS, PREC, false>(a, sig_all, s, m);
    return launch_cols_k<N, MODE_VJP, PREC, false>(a, sig_all, s, m);
}

// entry of one bucket (one translation unit each); PREC 0 = FP64, 1 = FP32 +
// a.refine (1 or 2) FP64-residual steps; with m, the persistent multi-lattice
// launch over m's work list (a is then only a prototype: n_c, refine, want_trace)
template <int N>
hipError_t launch_cols_bucket(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s,
                              const SchurMulti* m = nullptr)
{
    if (!fp32) return launch_cols_p<N, 0>(a, mode, sig_all, s, m);
    return launch_cols_p<N, 1>(a, mode, sig_all, s, m);
}

}  // namespace dg
