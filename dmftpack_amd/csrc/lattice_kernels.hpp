// Host-visible launchers of the lattice kernels (lattice_plain.hip,
// lattice_schur.hip, eig.hip).  All pointers are device pointers.
#pragma once
#include <hip/hip_runtime.h>

#include "dev_common.hpp"

namespace dg {

enum LatMode { MODE_FWD = 0, MODE_DENS = 1, MODE_VJP = 2 };
// One-matrix-per-lane kernels only: the forward over frequencies n >= N_freq,
// whose G_cc enters NumMatrix alone (through G + G^H), accumulates the
// Hermitian part (G + G^H)/2 (n_c^2 reals instead of n_c^2 complex).
constexpr int MODE_FWD_HERM = 3;
// The same Hermitian-part forward fused with the density terms (tr G^2): when an
// adjoint follows at the same (Sigma, mu), the frequencies beyond N_freq are done
// once, in the forward, instead of once there and once in the adjoint.
constexpr int MODE_DENS_HERM = 4;
// One-matrix-per-lane kernels only: the forward on the real axis (retarded_GreenFtn2),
// frequency point nw being z = e0 + nw de + i eta (freq_point)
constexpr int MODE_FWD_REAL = 5;

// Data read through the constant address space: the compiler treats it as
// invariant for the kernel's duration, so wave-uniform loads become scalar
// loads (s_load) whose results feed the VALU as SGPR operands.  Used for the
// per-k data of the Schur kernels, which no kernel writes while they run.
typedef __attribute__((address_space(4))) cd kconst_cd;
// bytes of padding after the per-k blocks of the one-matrix-per-lane kernels
// and after W (the scalar-cache fills read up to 4 KB from a block's start)
constexpr int kKdataPadBytes = 4096;

// Scalar-cache fill of the first 4 KB from base: one dword from each 64-byte
// line, all into one throw-away SGPR, waited for inside the asm block so that
// no load is in flight when it ends.  The lines arrive together (one L2
// round trip) and the scalar loads that follow hit the cache.
__device__ __forceinline__ void kcache_fill64(const void* base)
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
        "s_load_dword %0, %1, 0xc00\n\t"
        "s_load_dword %0, %1, 0xc40\n\t"
        "s_load_dword %0, %1, 0xc80\n\t"
        "s_load_dword %0, %1, 0xcc0\n\t"
        "s_load_dword %0, %1, 0xd00\n\t"
        "s_load_dword %0, %1, 0xd40\n\t"
        "s_load_dword %0, %1, 0xd80\n\t"
        "s_load_dword %0, %1, 0xdc0\n\t"
        "s_load_dword %0, %1, 0xe00\n\t"
        "s_load_dword %0, %1, 0xe40\n\t"
        "s_load_dword %0, %1, 0xe80\n\t"
        "s_load_dword %0, %1, 0xec0\n\t"
        "s_load_dword %0, %1, 0xf00\n\t"
        "s_load_dword %0, %1, 0xf40\n\t"
        "s_load_dword %0, %1, 0xf80\n\t"
        "s_load_dword %0, %1, 0xfc0\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&s"(t)
        : "s"(base));
    asm volatile("" ::"s"(t));
}
typedef __attribute__((address_space(4))) double kconst_double;

__device__ __forceinline__ cd ldc(const cd* p) { return *p; }
__device__ __forceinline__ cd ldc(const kconst_cd* p)
{
    const kconst_double* q = (const kconst_double*)p;
    return cd{q[0], q[1]};
}

struct LatArgs {
    const cd* Hk;     // [nk][n][n]
    const cd* sigma;  // [n_freq][nc][nc]
    const cd* mom;    // [4][nc][nc]
    int n, nc, nk, n_freq, nw_tot;
    double beta, mu;
    int k_per_chunk, n_chunks;
    // partial sums, [chunk][nw][...]
    cd* pG;    // FWD: sum_k G_cc       [chunks][nw_tot][nc*nc]
    cd* pTr;   // sum_k tr G            [chunks][nw_tot]
    cd* pTr2;  // sum_k tr G^2          [chunks][nw_tot]   (DENS, VJP)
    // VJP
    const cd* gbar;  // [n_freq][nc][nc]
    double nbar;
    cd* pSb;  // [chunks][n_freq][nc*nc] sum_k G^H gbar G^H + (2 nbar/beta) ((G^2)_cc)^H
    cd* pMu;  // [chunks][n_freq]         sum_k tr(gbar^H (G^2)_cc)
    // Schur-complement data (per k), see lattice_schur.hip
    const double* ed;  // [nk][nd]      eigenvalues of H_dd(k)
    const cd* W;       // [nk][nc][nd]  H_cd(k) V_d(k), column-major nc x nd
    int refine;        // FP32R: number of FP64-residual refinement steps
    cd* pSigAll;       // Schur: Sigma expanded to all nw_tot frequencies [nw_tot][nc*nc]
    int want_trace;    // FWD: also accumulate tr G (total N)
    const double* nbar_dev;  // VJP: device copy of nbar (overrides nbar when set)
    // real axis (retarded_GreenFtn2): frequency point nw is z = e0 + nw*de + i eta
    // instead of i w_n (plain and column-per-lane kernels, FWD mode)
    int real_axis;
    double e0, de, eta;
    // per-k blocks of the one-matrix-per-lane Schur kernels [nk][schur_kdata_size]
    const cd* kdata;
    // one-matrix-per-lane Schur kernels: frequency range [nw_begin, nw_end) of
    // this launch (nw_end = 0: up to nw_tot); partial sums keep the nw_tot stride
    int nw_begin, nw_end;
};

// (real shift added to mu, imaginary part) of frequency point nw
__device__ __forceinline__ void freq_point(const LatArgs& a, int nw, double& mu, double& w)
{
    if (a.real_axis) {
        mu = a.mu + a.e0 + nw * a.de;
        w = a.eta;
    } else {
        mu = a.mu;
        w = matsubara(nw, a.beta);
    }
}

// plain FP64 full inverse; returns 0 or a HIP error
hipError_t launch_lattice_plain(const LatArgs& a, int mode, hipStream_t s, int* wtile_out);
int plain_slots(int n, int mode);
// densityMatDFT: dens [nk][n][n] from U, static weights wt [nk][n]
hipError_t launch_density_matrix(const LatArgs& a, const cd* U, const double* wt, cd* dens, hipStream_t s);
// densityMatDFT in block form (density_lr.hip): n <= 16, n_c <= 6, a.kdata set
bool density_lr_supported(int n, int nc);
hipError_t launch_density_matrix_lr(const LatArgs& a, const cd* W, const cd* Vd, const double* ed, const cd* U,
                                    const double* wt, cd* dens, hipStream_t s);

// Schur complement (FP64 or FP32 + refinement); wg = workgroup width of the
// one-matrix-per-lane kernels: 64 or 128 frequencies, or 0 for the default of
// each launch (the FP32R forward over the first n_freq frequencies in
// SCHUR_WG_KPAIR form, the adjoint at 128, the others at 64); cols selects the
// column-per-lane kernels for n_c <= 6 (always used above 6)
// fused (one-matrix-per-lane kernels, schur_can_fuse): a forward that also
// writes tr G^2 (a.pTr2) beyond n_freq, and the adjoint at the same (Sigma, mu)
// that then covers the first n_freq frequencies only (same k chunking).
// wg value of the FP32R forward kernel (n_c >= 5) as two wavefronts over 64
// frequencies, each walking every other k of the chunk into shared LDS
// accumulators (half the LDS per wavefront of the 64-lane form); its work items
// cover 64 frequencies
constexpr int SCHUR_WG_KPAIR = 1;
hipError_t launch_lattice_schur(const LatArgs& a, int mode, bool fp32, hipStream_t s, int wg = 64, bool cols = false,
                                bool fused = false);
bool schur_can_fuse(int nc, bool cols, bool real_axis);
// Sigma expanded over all nw_tot frequencies into a.pSigAll
hipError_t launch_sigma_all(const LatArgs& a, hipStream_t s);
// Persistent launch over a work list of several lattices sharing n_c (and mode,
// precision, refinement): args [systems] in device memory, items {system,
// wtile, chunk, 0}; proto carries nc / refine / want_trace.
struct SchurMulti {
    const LatArgs* args;
    const int4* items;
    int n_items;
    int grid;
    int max_nd;
};
hipError_t launch_lattice_schur_multi(const LatArgs& proto, int mode, bool fp32, hipStream_t s, int wg,
                                      const SchurMulti& m);
// the same for 6 < n_c <= 16: column-per-lane kernels (MODE_FWD, MODE_VJP over every frequency)
hipError_t launch_lattice_schur_cols_multi(const LatArgs& proto, int mode, bool fp32, hipStream_t s,
                                           const SchurMulti& m);
// frequencies per workgroup of the Schur kernel for n_c (grid planning)
int schur_slots(int nc, int mode, bool fp32, bool cols);
bool schur_supported(int n, int nc);
// the persistent multi-lattice launch covers the one-matrix-per-lane kernels
bool schur_multi_supported(int nc);
// one-matrix-per-lane Schur kernels of one n_c (lattice_lane{1..6}.hip)
hipError_t launch_lane_1(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s, int wg,
                         const SchurMulti* m);
hipError_t launch_lane_2(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s, int wg,
                         const SchurMulti* m);
hipError_t launch_lane_3(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s, int wg,
                         const SchurMulti* m);
hipError_t launch_lane_4(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s, int wg,
                         const SchurMulti* m);
hipError_t launch_lane_5(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s, int wg,
                         const SchurMulti* m);
hipError_t launch_lane_6(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s, int wg,
                         const SchurMulti* m);
// column-per-lane Schur kernels (lattice_cols{4,5,6,8,10,12,16}.hip)
hipError_t launch_cols_4(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s);
hipError_t launch_cols_5(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s);
hipError_t launch_cols_6(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s);
hipError_t launch_cols_8(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s);
// persistent multi-lattice launches of the column-per-lane buckets for n_c > 6
hipError_t launch_cols_multi_8(const LatArgs& proto, int mode, bool fp32, hipStream_t s, const SchurMulti& m);
hipError_t launch_cols_multi_10(const LatArgs& proto, int mode, bool fp32, hipStream_t s, const SchurMulti& m);
hipError_t launch_cols_multi_12(const LatArgs& proto, int mode, bool fp32, hipStream_t s, const SchurMulti& m);
hipError_t launch_cols_multi_16(const LatArgs& proto, int mode, bool fp32, hipStream_t s, const SchurMulti& m);
hipError_t launch_cols_10(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s);
hipError_t launch_cols_12(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s);
hipError_t launch_cols_16(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s);

// W = H_cd V_d per k (Schur preparation)
hipError_t launch_schur_prep(const cd* Hk, const cd* Vd, int n, int nc, int nk, cd* W, hipStream_t s);
// per-k blocks [H_cc][P_d pairs][e_d] of the one-matrix-per-lane kernels (n_c <= 6)
int schur_kdata_size(int n, int nc);
hipError_t launch_schur_kdata(const cd* Hk, const cd* W, const double* ed, int n, int nc, int nk, cd* out,
                              hipStream_t s);

// Batched Hermitian eigen-decomposition of X_b[off:off+m, off:off+m] (+ add on
// the leading nadd x nadd block).  w [batch][m] ascending, V [batch][m][m] (may be null).
hipError_t launch_herm_eig(const cd* X, int ldx, int off, int m, const cd* add, int nadd, double* w, cd* V,
                           int batch, hipStream_t s);

// Static part of the densities: per (k,l) weights f - s and dN terms, then
// per-chunk partials of sum_k (U diag(f-s) U^H)_cc, sum (f-s), sum dN.
hipError_t launch_static_terms(const double* ev, int n, int nk, int nw_tot, double beta, double mu, double* wt,
                               double* dt, hipStream_t s);
hipError_t launch_static_reduce(const double* wt, const double* dt, const cd* U, int n, int nc, int nk,
                                int k_per_chunk, int n_chunks, cd* pNum, double* pN, double* pdN, hipStream_t s);

// Core-band term of GiwToNele: per-chunk sums of f(eps - mu) and beta f (1 - f)
// over the Kohn-Sham bands flagged out[k][l] = 1 (outside the model window)
hipError_t launch_core_bands(const double* eps, const unsigned char* out, int n_orb, int nk, int k_per_chunk,
                             int n_chunks, double beta, double mu, double* pN, double* pdN, hipStream_t s);

// loss cotangents gbar = 2 (G_loc - G_tgt) inside cluster blocks (npc orbitals per
// cluster), 0 across them; *loss = sum |G_loc - G_tgt|^2 (device scalar)
hipError_t launch_loss_cotangent(const cd* gloc, const cd* gtgt, int n_freq, int nc, int npc, cd* gbar, double* loss,
                                 hipStream_t s);
// GwHF (optional output): gloc_hf[nw][a][b] = (1/nkg) sum_k sum_l U_al conj(U_bl)/(iw+mu-e_kl)
hipError_t launch_gloc_hf(const double* ev, const cd* U, int n, int nc, int nk, int n_freq, double beta, double mu,
                          double inv_nkg, int n_cluster, cd* out, hipStream_t s);

// Deterministic reductions over chunks: out[i] = sum_c part[c*len + i]
hipError_t launch_reduce_chunks(const cd* part, int n_chunks, size_t len, cd* out, hipStream_t s);
hipError_t launch_reduce_chunks_d(const double* part, int n_chunks, size_t len, double* out, hipStream_t s);

// Final reductions over frequency: for FWD/DENS/VJP totals.
//   sumG[e]  = sum_nw (G[nw][e] + conj(G[nw][e^T]))        (e < m2)   if G != null
//   sumTr    = sum_nw 2 Re tr[nw]
//   sumTr2   = sum_nw 2 Re tr2[nw]                          if tr2 != null
//   sumMu    = sum_nw Re mu[nw] over n_mu entries           if mu != null
// out layout: [m2 complex] [3 doubles as complex .x]
hipError_t launch_freq_totals(const cd* G, int m2, int nc, const cd* tr, const cd* tr2, int nw_tot, const cd* mu,
                              int n_mu, cd* out, hipStream_t s);

// gloc[nw][e] = Gsum[nw][e] * scale for nw < n_freq, zero outside cluster blocks
hipError_t launch_gloc_finalize(const cd* Gsum, int n_freq, int nc, int n_cluster, double scale, cd* gloc,
                                hipStream_t s);
// out[m] = scale * in[m]^H over a stack of count d x d matrices (in != out)
hipError_t launch_adjoint_stack(const cd* in, size_t count, int d, double scale, cd* out, hipStream_t s);
// out[i] = in[i] * scale
hipError_t launch_scale(const cd* in, size_t len, double scale, cd* out, hipStream_t s);

}  // namespace dg
