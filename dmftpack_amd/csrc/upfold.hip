// upfolding_density (upfolding.cpp:38-139) with upfolding_k (:10-35): the
// per-k density matrix of the lattice path, taken from the model basis to the
// full DFT basis, and its k -> R sums (the values DM_DFT_DMFT.HWR lists).
//
// Per k (downfolding == 1 branch, :45-51):
//   E  = (T^H)^{-1} K_o                        T = transformMatrix_k, K_o = KS_eigenVectors_orthoBasis
//   M  = C^H X C                               C = DF_CorrBase, X = model-basis densityMatDFT (:13)
//   K  = diag f(eps - mu) with K(v_i, v_j) = M(i, j)   (:15-32)
//   D  = E K E^H                               (:34)
//   DDir = S^{-1} D,  DD = S^{-1} D S^{-1}     (:49-50)
// then dmr[x][r] = sum_k Y_x(k)(i0, m0) exp(+i kphase_k . R) / n_k_global, (:67-86).
//
// All of it is O(n_orb^3) per k next to the lattice path's O(n^3) per (k, iw),
// so the kernels are plain: a batched square complex GEMM with 16x16 LDS
// tiles, an in-place Gauss-Jordan inverse with partial pivoting (one
// workgroup per matrix, the matrix in global memory), a gather that builds K,
// and a k -> R sum that stages a tile of k values in LDS and gives every
// thread one R of one matrix element.  k is processed in chunks that bound the
// scratch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <vector>

#include "../../include/dmftgpu.h"
#include "api_internal.hpp"

using dg::cd;
using namespace dgi;

#define HIPCHK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(h, DMFTGPU_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace {

using namespace dg;

constexpr int UPF_MAX_ORB = 512;  // LDS of the Gauss-Jordan kernel: a column and the pivots
constexpr size_t UPF_CHUNK_BYTES = size_t(512) << 20;

// C_b = op(A_b) op(B_b), N x N column-major, op = identity or ^H (HA, HB).
// Each operand tile is read with consecutive threads on consecutive
// addresses; the ^H tiles are transposed on their way into LDS.
template <bool HA, bool HB>
__global__ __launch_bounds__(256) void k_zgemm_sq(int N, const cd* __restrict__ A, size_t sA, const cd* __restrict__ B,
                                                  size_t sB, cd* __restrict__ C, size_t sC)
{
    __shared__ cd ta[16][17];  // ta[i][l] = op(A)(i0 + i, l0 + l)
    __shared__ cd tb[16][17];  // tb[l][j] = op(B)(l0 + l, j0 + j)
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int i0 = blockIdx.x * 16, j0 = blockIdx.y * 16;
    const cd* Ab = A + (size_t)blockIdx.z * sA;
    const cd* Bb = B + (size_t)blockIdx.z * sB;
    cd acc = cd{0.0, 0.0};
    for (int l0 = 0; l0 < N; l0 += 16) {
        if (!HA) {  // A(i0 + tx, l0 + ty)
            const int r = i0 + tx, c = l0 + ty;
            ta[tx][ty] = (r < N && c < N) ? Ab[r + (size_t)c * N] : cd{0.0, 0.0};
        } else {  // conj A(l0 + tx, i0 + ty) -> op(A)(i0 + ty, l0 + tx)
            const int r = l0 + tx, c = i0 + ty;
            ta[ty][tx] = (r < N && c < N) ? cconj(Ab[r + (size_t)c * N]) : cd{0.0, 0.0};
        }
        if (!HB) {  // B(l0 + tx, j0 + ty)
            const int r = l0 + tx, c = j0 + ty;
            tb[tx][ty] = (r < N && c < N) ? Bb[r + (size_t)c * N] : cd{0.0, 0.0};
        } else {  // conj B(j0 + tx, l0 + ty) -> op(B)(l0 + ty, j0 + tx)
            const int r = j0 + tx, c = l0 + ty;
            tb[ty][tx] = (r < N && c < N) ? cconj(Bb[r + (size_t)c * N]) : cd{0.0, 0.0};
        }
        __syncthreads();
#pragma unroll
        for (int l = 0; l < 16; ++l) acc = cfma(ta[tx][l], tb[l][ty], acc);
        __syncthreads();
    }
    const int r = i0 + tx, c = j0 + ty;
    if (r < N && c < N) C[(size_t)blockIdx.z * sC + r + (size_t)c * N] = acc;
}

template <bool HA, bool HB>
hipError_t zgemm_sq(int N, int batch, const cd* A, const cd* B, cd* C, hipStream_t s)
{
    const size_t nn = (size_t)N * N;
    const unsigned t = (unsigned)((N + 15) / 16);
    hipLaunchKernelGGL((k_zgemm_sq<HA, HB>), dim3(t, t, (unsigned)batch), dim3(256), 0, s, N, A, nn, B, nn, C, nn);
    return hipGetLastError();
}

// In-place inverse of each N x N matrix (column-major, stride N*N) by
// Gauss-Jordan elimination with partial (row) pivoting, the pivot being the
// largest |A(i, j)|, i >= j, the first of equals; the row exchanges are undone
// on the columns at the end.  One workgroup per matrix; the matrix stays in
// global memory (the workgroup's barriers order its own loads and stores).
// A zero pivot sets *singular.  IN_LDS: the matrix is staged in LDS for the
// elimination (N <= 64, one wavefront of BS = 64 per matrix), the same steps in
// the same order.
template <bool IN_LDS, int BS>
__global__ __launch_bounds__(BS) void k_gj_inverse(int N, cd* A, int* singular)
{
    __shared__ cd col[IN_LDS ? 64 : UPF_MAX_ORB];  // IN_LDS: N <= 64
    __shared__ int piv[IN_LDS ? 64 : UPF_MAX_ORB];
    __shared__ double rmax[BS];
    __shared__ int ridx[BS];
    extern __shared__ cd mlds[];
    const int t = threadIdx.x;
    cd* G = A + (size_t)blockIdx.x * N * N;
    cd* M = G;
    if constexpr (IN_LDS) {
        for (int e = t; e < N * N; e += BS) mlds[e] = G[e];
        M = mlds;
        __syncthreads();
    }
    // the elimination's (row, column) walk: thread t keeps row t % N and takes every
    // (BS / N)-th column from t / N (N <= BS), no division per element
    const int cpp = N <= BS ? BS / N : 0;
    const int ti = N <= BS ? t % N : 0, tc = N <= BS ? t / N : 0;
    for (int j = 0; j < N; ++j) {
        if (N <= 64) {  // every candidate row in wavefront 0: a wavefront reduction, one barrier
            if (t < 64) {
                double best = -1.0;
                int bi = j;
                if (j + t < N) {
                    const cd a = M[j + t + (size_t)j * N];
                    best = fma(a.x, a.x, a.y * a.y);
                    bi = j + t;
                }
                for (int off = 32; off > 0; off >>= 1) {
                    const double o = __shfl_xor(best, off);
                    const int oi = __shfl_xor(bi, off);
                    if (o > best || (o == best && oi < bi)) { best = o; bi = oi; }
                }
                if (t == 0) {
                    rmax[0] = best;
                    ridx[0] = bi;
                }
            }
            __syncthreads();
        } else {
            double best = -1.0;
            int bi = j;
            for (int i = j + t; i < N; i += BS) {
                const cd a = M[i + (size_t)j * N];
                const double m = fma(a.x, a.x, a.y * a.y);
                if (m > best) { best = m; bi = i; }
            }
            rmax[t] = best;
            ridx[t] = bi;
            __syncthreads();
            for (int w = BS / 2; w > 0; w >>= 1) {
                if (t < w) {
                    const double o = rmax[t + w];
                    const int oi = ridx[t + w];
                    if (o > rmax[t] || (o == rmax[t] && oi < ridx[t])) { rmax[t] = o; ridx[t] = oi; }
                }
                __syncthreads();
            }
        }
        const int p = ridx[0];
        const bool zero = !(rmax[0] > 0.0);
        if (t == 0) {
            piv[j] = p;
            if (zero) *singular = 1;
        }
        if (p != j)
            for (int c = t; c < N; c += BS) {
                const cd a = M[j + (size_t)c * N];
                M[j + (size_t)c * N] = M[p + (size_t)c * N];
                M[p + (size_t)c * N] = a;
            }
        __syncthreads();
        const cd d = M[j + (size_t)j * N];
        const cd r = zero ? cd{0.0, 0.0} : crcp(d);
        for (int i = t; i < N; i += BS) col[i] = M[i + (size_t)j * N];
        __syncthreads();
        for (int c = t; c < N; c += BS) M[j + (size_t)c * N] = (c == j) ? r : cmul(M[j + (size_t)c * N], r);
        __syncthreads();
        if (N <= BS) {
            if (ti != j && tc < cpp) {
                const cd ci = col[ti];
                for (int c = tc; c < N; c += cpp) {
                    const size_t e = ti + (size_t)c * N;
                    const cd base = (c == j) ? cd{0.0, 0.0} : M[e];
                    M[e] = cfms(ci, M[j + (size_t)c * N], base);
                }
            }
        } else {
            for (size_t e = t; e < (size_t)N * N; e += BS) {
                const int i = (int)(e % N), c = (int)(e / N);
                if (i == j) continue;
                const cd base = (c == j) ? cd{0.0, 0.0} : M[e];
                M[e] = cfms(col[i], M[j + (size_t)c * N], base);
            }
        }
        __syncthreads();
    }
    for (int j = N - 1; j >= 0; --j) {
        const int p = piv[j];
        if (p != j)
            for (int i = t; i < N; i += BS) {
                const cd a = M[i + (size_t)j * N];
                M[i + (size_t)j * N] = M[i + (size_t)p * N];
                M[i + (size_t)p * N] = a;
            }
        __syncthreads();
    }
    if constexpr (IN_LDS)
        for (int e = t; e < N * N; e += BS) G[e] = mlds[e];
}

hipError_t gj_inverse(int N, int batch, cd* A, int* singular, hipStream_t s)
{
    if (N <= 64) {  // up to 64 KiB of matrix next to the static arrays; one wavefront per matrix
        const size_t shm = sizeof(cd) * (size_t)N * N;
        const hipError_t e = hipFuncSetAttribute((const void*)k_gj_inverse<true, 64>,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_gj_inverse<true, 64>), dim3((unsigned)batch), dim3(64), shm, s, N, A, singular);
    } else
        hipLaunchKernelGGL((k_gj_inverse<false, 256>), dim3((unsigned)batch), dim3(256), 0, s, N, A, singular);
    return hipGetLastError();
}

// K(a, b) of upfolding_k in the KS basis: the model block M(inv_a, inv_b)
// where both bands are in the window (inv >= 0), else the clamped Fermi
// factor on the diagonal (upfolding.cpp:15-32).
__global__ void k_upfold_assemble(const cd* __restrict__ M, const double* __restrict__ eps, const int* __restrict__ inv,
                                  int n, int n_orb, int kc, double beta, double mu, cd* __restrict__ K)
{
    const size_t oo = (size_t)n_orb * n_orb;
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= oo * kc) return;
    const int k = (int)(gid / oo);
    const int e = (int)(gid % oo);
    const int a = e % n_orb, b = e / n_orb;
    const int ia = inv[(size_t)k * n_orb + a], ib = inv[(size_t)k * n_orb + b];
    cd v = cd{0.0, 0.0};
    if (ia >= 0 && ib >= 0) {
        v = M[(size_t)k * n * n + ia + (size_t)ib * n];
    } else if (a == b) {
        const double x = beta * (eps[(size_t)k * n_orb + a] - mu);
        v.x = x > 100.0 ? 0.0 : x < -100.0 ? 1.0 : 1.0 / (1.0 + exp(x));
    }
    K[gid] = v;
}

// dmr[x][orig] += (1/n_k_global) sum_{k in chunk} Y_x(k)(e) exp(+i p_k . R):
// one workgroup per (matrix element e, up to 256 of its R); the chunk's k run
// through LDS in tiles of 256 (the element of the three matrices + the phase).
struct RItem {
    int e, r0, r1, pad;
};
__global__ __launch_bounds__(256) void k_dm_r(const cd* __restrict__ Y0, const cd* __restrict__ Y1,
                                              const cd* __restrict__ Y2, const double* __restrict__ kphase, int kc,
                                              size_t oo, const RItem* __restrict__ items, const int4* __restrict__ rs,
                                              int n_r, double inv_nk, cd* __restrict__ dmr)
{
    __shared__ cd y[3][256];
    __shared__ double ph[3][256];
    const RItem it = items[blockIdx.x];
    const int t = threadIdx.x;
    const int r = it.r0 + t;
    const bool live = r < it.r1;
    const int4 R = rs[live ? r : it.r0];
    cd a0 = cd{0.0, 0.0}, a1 = a0, a2 = a0;
    for (int k0 = 0; k0 < kc; k0 += 256) {
        const int k = k0 + t;
        if (k < kc) {
            const size_t o = (size_t)k * oo + it.e;
            y[0][t] = Y0[o];
            y[1][t] = Y1[o];
            y[2][t] = Y2[o];
            ph[0][t] = kphase[3 * (size_t)k];
            ph[1][t] = kphase[3 * (size_t)k + 1];
            ph[2][t] = kphase[3 * (size_t)k + 2];
        }
        __syncthreads();
        const int kn = min(256, kc - k0);
        for (int q = 0; q < kn; ++q) {
            double sn, cs;
            sincos(ph[0][q] * R.x + ph[1][q] * R.y + ph[2][q] * R.z, &sn, &cs);
            const cd z = cd{cs, sn};
            a0 = cfma(y[0][q], z, a0);
            a1 = cfma(y[1][q], z, a1);
            a2 = cfma(y[2][q], z, a2);
        }
        __syncthreads();
    }
    if (!live) return;
    const int o = R.w;
    cd* d0 = dmr + o;
    cd* d1 = dmr + (size_t)n_r + o;
    cd* d2 = dmr + 2 * (size_t)n_r + o;
    *d0 = cadd(*d0, cscale(a0, inv_nk));
    *d1 = cadd(*d1, cscale(a1, inv_nk));
    *d2 = cadd(*d2, cscale(a2, inv_nk));
}

// exp(+i phi_a r), |r| <= rmax, per k of the chunk: tab[k][a][r + rmax]
__global__ void k_axis_phases(const double* __restrict__ kphase, int kc, int rmax, cd* __restrict__ tab)
{
    const int nr = 2 * rmax + 1;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)kc * 3 * nr) return;
    const int k = (int)(t / (3 * nr)), q = (int)(t % (3 * nr));
    const int ax = q / nr, r = q % nr - rmax;
    double sn, cs;
    sincos(kphase[3 * (size_t)k + ax] * r, &sn, &cs);
    tab[t] = cd{cs, sn};
}

// k_dm_r with the phase taken apart per axis, exp(+i phi.R) = ex[R1] ey[R2] ez[R3]:
// a tile of kt k values stages the matrix elements and the k's axis tables in
// LDS; two complex products per (R, k) instead of a sincos.  A work item is a
// run of up to 256 consecutive entries of the R list over up to DMR_EM matrix
// elements (entry q: element slot eslot[q] of the item), so that a workgroup's
// threads are all busy when each element has few R (27 for R in {-1,0,1}^3).
// blockIdx.y takes segment y of the chunk's k (segments of kseg k, so that the
// launch has enough workgroups to fill the GPU); each thread sums its k of the
// segment in order and stores the sum to part[y][x][entry] for k_dm_r_merge.
constexpr int DMR_EM = 16;
struct RItemE {
    int q0, q1, eb, ne;  // entries [q0, q1), elements ielem[eb .. eb + ne)
};
__global__ __launch_bounds__(256) void k_dm_r_axes(const cd* __restrict__ Y0, const cd* __restrict__ Y1,
                                                   const cd* __restrict__ Y2, const cd* __restrict__ tab, int rmax,
                                                   int kc, int kt, size_t oo, const RItemE* __restrict__ items,
                                                   const int* __restrict__ ielem, const int* __restrict__ eslot,
                                                   const int4* __restrict__ rs, int n_r, int kseg,
                                                   cd* __restrict__ part)
{
    extern __shared__ cd sm[];
    __shared__ int se[DMR_EM];
    const int nr = 2 * rmax + 1, per = 3 * nr;
    cd* y = sm;                        // [3][DMR_EM][kt]
    cd* ph = sm + 3 * DMR_EM * kt;     // [kt][3][nr]
    const RItemE it = items[blockIdx.x];
    const int t = threadIdx.x;
    const int q = it.q0 + t;
    const bool live = q < it.q1;
    const int4 R = rs[live ? q : it.q0];
    const int j = eslot[live ? q : it.q0];
    if (t < it.ne) se[t] = ielem[it.eb + t];
    const int ox = R.x + rmax, oy = nr + R.y + rmax, oz = 2 * nr + R.z + rmax;
    const cd* y0 = y + (size_t)j * kt;
    const cd* y1 = y + (size_t)(DMR_EM + j) * kt;
    const cd* y2 = y + (size_t)(2 * DMR_EM + j) * kt;
    cd a0 = cd{0.0, 0.0}, a1 = a0, a2 = a0;
    __syncthreads();
    const int kb = blockIdx.y * kseg, ke = min(kc, kb + kseg);
    for (int k0 = kb; k0 < ke; k0 += kt) {
        const int kn = min(kt, ke - k0);
        for (int v = t; v < kn * it.ne; v += 256) {  // the item's elements of one k are neighbours in memory
            const int qk = v / it.ne, jj = v - qk * it.ne;
            const size_t o = (size_t)(k0 + qk) * oo + se[jj];
            y[(size_t)jj * kt + qk] = Y0[o];
            y[(size_t)(DMR_EM + jj) * kt + qk] = Y1[o];
            y[(size_t)(2 * DMR_EM + jj) * kt + qk] = Y2[o];
        }
        for (int v = t; v < kn * per; v += 256) ph[v] = tab[(size_t)k0 * per + v];
        __syncthreads();
        for (int qk = 0; qk < kn; ++qk) {
            const cd* p = ph + qk * per;
            const cd z = cmul(cmul(p[ox], p[oy]), p[oz]);
            a0 = cfma(y0[qk], z, a0);
            a1 = cfma(y1[qk], z, a1);
            a2 = cfma(y2[qk], z, a2);
        }
        __syncthreads();
    }
    if (!live) return;
    cd* pp = part + (size_t)blockIdx.y * 3 * n_r + q;
    pp[0] = a0;
    pp[n_r] = a1;
    pp[2 * (size_t)n_r] = a2;
}

// dmr[x][R.w] += (1/n_k_global) sum over the segments, in order, of part[y][x][entry]
__global__ void k_dm_r_merge(const cd* __restrict__ part, int nseg, const int4* __restrict__ rs, int n_r,
                             double inv_nk, cd* __restrict__ dmr)
{
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= 3 * (size_t)n_r) return;
    const int x = (int)(gid / n_r), q = (int)(gid % n_r);
    cd a = part[gid];
    for (int y = 1; y < nseg; ++y) a = cadd(a, part[(size_t)y * 3 * n_r + gid]);
    cd* d = dmr + (size_t)x * n_r + rs[q].w;
    *d = cadd(*d, cscale(a, inv_nk));
}

__global__ void k_zero(cd* __restrict__ p, size_t count)
{
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid < count) p[gid] = cd{0.0, 0.0};
}

__global__ void k_zero_int(int* p) { *p = 0; }

inline size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

// upfolding_density (upfolding.cpp:38-139) with upfolding_k (:10-35)
extern "C" int dmftgpu_upfold_density(dmftgpu_handle h, int n, int n_orb, int n_k, int n_k_global, double beta,
                                      double mu, const double* dm_model, const double* corr_base,
                                      const double* ks_energy, const int* val_to_ks, const double* transform,
                                      const double* ks_ortho, const double* overlap, int n_r, const int* r_idx,
                                      const double* kphase, double* dm_k, double* dmr)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (n < 1 || n_orb < n || n_k < 1 || n_k_global < n_k || n_r < 0 || !(beta > 0.0) || !std::isfinite(mu))
        return fail(h, DMFTGPU_ERR_ARG, "n=%d n_orb=%d n_k=%d n_k_global=%d n_r=%d beta=%g mu=%g", n, n_orb, n_k,
                    n_k_global, n_r, beta, mu);
    if (n_orb > UPF_MAX_ORB)
        return fail(h, DMFTGPU_ERR_UNSUPPORTED, "n_orb=%d above %d", n_orb, UPF_MAX_ORB);
    if (!dm_model || !corr_base || !ks_energy || !val_to_ks || !transform || !ks_ortho || !kphase)
        return fail(h, DMFTGPU_ERR_ARG, "null argument");
    if (n_r > 0 && (!r_idx || !dmr)) return fail(h, DMFTGPU_ERR_ARG, "n_r > 0 needs r_idx and dmr");
    if (n_r == 0 && !dm_k) return fail(h, DMFTGPU_ERR_ARG, "nothing to compute: n_r = 0 and dm_k = NULL");
    // FromValToKS -> band-to-window map, checked: each k's window bands are distinct
    std::vector<int> inv((size_t)n_k * n_orb, -1);
    for (int k = 0; k < n_k; ++k)
        for (int i = 0; i < n; ++i) {
            const int b = val_to_ks[(size_t)k * n + i];
            if (b == -1) continue;  // a padded band (NBAND[k] < n): no KS band
            if (b < 0 || b >= n_orb) return fail(h, DMFTGPU_ERR_ARG, "val_to_ks[%d][%d]=%d outside 0..%d", k, i, b, n_orb - 1);
            if (inv[(size_t)k * n_orb + b] >= 0)
                return fail(h, DMFTGPU_ERR_ARG, "val_to_ks[%d] names band %d twice", k, b);
            inv[(size_t)k * n_orb + b] = i;
        }
    // R list grouped by matrix element (stable), work items of <= 256 R
    const size_t oo = (size_t)n_orb * n_orb, nn = (size_t)n * n;
    std::vector<int> order(n_r);
    for (int r = 0; r < n_r; ++r) {
        const int i0 = r_idx[5 * (size_t)r + 3], m0 = r_idx[5 * (size_t)r + 4];
        if (i0 < 0 || i0 >= n_orb || m0 < 0 || m0 >= n_orb)
            return fail(h, DMFTGPU_ERR_ARG, "r_idx[%d]: orbital pair (%d,%d) outside 0..%d", r, i0, m0, n_orb - 1);
        order[r] = r;
    }
    auto elem = [&](int r) { return (size_t)r_idx[5 * (size_t)r + 3] + (size_t)r_idx[5 * (size_t)r + 4] * n_orb; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return elem(a) < elem(b); });
    std::vector<int> rs(4 * (size_t)std::max(n_r, 1), 0);
    std::vector<RItem> items;
    for (int q = 0; q < n_r;) {
        const size_t e = elem(order[q]);
        int q1 = q;
        while (q1 < n_r && elem(order[q1]) == e) ++q1;
        for (int a = q; a < q1; a += 256) items.push_back(RItem{(int)e, a, std::min(q1, a + 256), 0});
        q = q1;
    }
    // the axes kernel's items: runs of <= 256 entries over <= DMR_EM elements
    std::vector<RItemE> items_e;
    std::vector<int> ielem, eslot(std::max(n_r, 1), 0);
    {
        RItemE cur{0, 0, 0, 0};
        bool open = false;
        for (int q = 0; q < n_r;) {
            const size_t e = elem(order[q]);
            int q1 = q;
            while (q1 < n_r && elem(order[q1]) == e) ++q1;
            for (int a = q; a < q1;) {
                if (!open || cur.q1 - cur.q0 == 256 || cur.ne == DMR_EM) {
                    if (open) items_e.push_back(cur);
                    cur = RItemE{a, a, (int)ielem.size(), 0};
                    open = true;
                }
                const int take = std::min(q1 - a, 256 - (cur.q1 - cur.q0));
                ielem.push_back((int)e);
                for (int b = a; b < a + take; ++b) eslot[b] = cur.ne;
                cur.ne += 1;
                cur.q1 += take;
                a += take;
            }
            q = q1;
        }
        if (open) items_e.push_back(cur);
    }
    int rmax = 0;  // the R sums take the per-axis phase tables when |R_a| <= 16
    for (int r = 0; r < n_r; ++r)
        for (int c = 0; c < 3; ++c) rmax = std::max(rmax, std::abs(r_idx[5 * (size_t)r + c]));
    const bool axes = n_r > 0 && rmax <= 16;
    const int nrr = 2 * rmax + 1;
    const int kt = std::max(16, std::min(256, (48 * 1024 / (int)sizeof(cd)) / (3 * DMR_EM + 3 * nrr)));
    // k segments of the R sums: about 1024 workgroups per launch, each segment >= 2 tiles
    const int nseg_max = axes ? std::max(1, std::min(16, (1024 + (int)items_e.size() - 1) / (int)items_e.size())) : 1;
    for (int q = 0; q < n_r; ++q) {
        const int r = order[q];
        rs[4 * (size_t)q] = r_idx[5 * (size_t)r];
        rs[4 * (size_t)q + 1] = r_idx[5 * (size_t)r + 1];
        rs[4 * (size_t)q + 2] = r_idx[5 * (size_t)r + 2];
        rs[4 * (size_t)q + 3] = r;
    }
    // scratch: maps + per-chunk W1, W2, E (+ D) [oo], T1, M [nn] and the axis phases per k
    const size_t per_k = (4 * oo + 2 * nn + (axes ? 3 * (size_t)nrr : 0)) * sizeof(cd);
    size_t kc_cap = std::min<size_t>({(size_t)n_k, UPF_CHUNK_BYTES / per_k, 65535});
    if (h->upfold_chunk > 0) kc_cap = std::min<size_t>(kc_cap, (size_t)h->upfold_chunk);
    const int kc_max = (int)std::max<size_t>(1, kc_cap);
    const size_t b_inv = align256(inv.size() * sizeof(int));
    const size_t b_rs = align256(rs.size() * sizeof(int));
    const size_t b_items = align256(std::max<size_t>(items.size(), 1) * sizeof(RItem));
    const size_t b_items_e = align256(std::max<size_t>(items_e.size(), 1) * sizeof(RItemE));
    const size_t b_ielem = align256(std::max<size_t>(ielem.size(), 1) * sizeof(int));
    const size_t b_eslot = align256(eslot.size() * sizeof(int));
    const size_t b_part = axes ? align256((size_t)nseg_max * 3 * n_r * sizeof(cd)) : 0;
    const size_t b_lists = b_items + b_items_e + b_ielem + b_eslot + b_part;
    const size_t b_flag = 256;
    const size_t b_chunk = align256(per_k * kc_max);
    hipError_t e;
    char* base = (char*)h->ws.get(b_inv + b_rs + b_lists + b_flag + b_chunk, &e);
    if (!base) return fail(h, DMFTGPU_ERR_HIP, "workspace: %s", hipGetErrorString(e));
    int* d_inv = (int*)base;
    int4* d_rs = (int4*)(base + b_inv);
    RItem* d_items = (RItem*)(base + b_inv + b_rs);
    RItemE* d_items_e = (RItemE*)(base + b_inv + b_rs + b_items);
    int* d_ielem = (int*)(base + b_inv + b_rs + b_items + b_items_e);
    int* d_eslot = (int*)(base + b_inv + b_rs + b_items + b_items_e + b_ielem);
    cd* d_part = (cd*)(base + b_inv + b_rs + b_items + b_items_e + b_ielem + b_eslot);
    int* d_flag = (int*)(base + b_inv + b_rs + b_lists);
    cd* W1 = (cd*)(base + b_inv + b_rs + b_lists + b_flag);
    cd* W2 = W1 + oo * kc_max;
    cd* Eb = W2 + oo * kc_max;
    cd* Dw = Eb + oo * kc_max;
    cd* T1 = Dw + oo * kc_max;
    cd* Mb = T1 + nn * kc_max;
    cd* Ptab = Mb + nn * kc_max;  // axes: [kc][3][2 rmax + 1]
    hipStream_t s = h->stream;
    HIPCHK(hipMemcpyAsync(d_inv, inv.data(), inv.size() * sizeof(int), hipMemcpyHostToDevice, s));
    if (n_r > 0) {
        HIPCHK(hipMemcpyAsync(d_rs, rs.data(), rs.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(RItem), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_items_e, items_e.data(), items_e.size() * sizeof(RItemE), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_ielem, ielem.data(), ielem.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d_eslot, eslot.data(), eslot.size() * sizeof(int), hipMemcpyHostToDevice, s));
        const size_t cnt = 3 * (size_t)n_r;
        hipLaunchKernelGGL(k_zero, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, (cd*)dmr, cnt);
        HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, s, d_flag);
    HIPCHK(hipGetLastError());
    const cd *X = (const cd*)dm_model, *Cb = (const cd*)corr_base, *Tm = (const cd*)transform,
             *Ko = (const cd*)ks_ortho, *Sm = (const cd*)overlap;
    const double inv_nk = 1.0 / (double)n_k_global;
    for (int k0 = 0; k0 < n_k; k0 += kc_max) {
        const int kc = std::min(kc_max, n_k - k0);
        const size_t ko = (size_t)k0 * oo, kn = (size_t)k0 * nn;
        cd* D = dm_k ? (cd*)dm_k + ko : Dw;
        // E = (T^{-1})^H K_o
        HIPCHK(hipMemcpyAsync(W1, Tm + ko, oo * kc * sizeof(cd), hipMemcpyDeviceToDevice, s));
        HIPCHK(gj_inverse(n_orb, kc, W1, d_flag, s));
        HIPCHK((zgemm_sq<true, false>(n_orb, kc, W1, Ko + ko, Eb, s)));
        // M = C^H (X C)
        HIPCHK((zgemm_sq<false, false>(n, kc, X + kn, Cb + kn, T1, s)));
        HIPCHK((zgemm_sq<true, false>(n, kc, Cb + kn, T1, Mb, s)));
        // K, then D = (E K) E^H
        const size_t tot = oo * kc;
        hipLaunchKernelGGL(k_upfold_assemble, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, Mb,
                           ks_energy + (size_t)k0 * n_orb, d_inv + (size_t)k0 * n_orb, n, n_orb, kc, beta, mu, W1);
        HIPCHK(hipGetLastError());
        HIPCHK((zgemm_sq<false, false>(n_orb, kc, Eb, W1, W2, s)));
        HIPCHK((zgemm_sq<false, true>(n_orb, kc, W2, Eb, D, s)));
        const cd *DD = D, *DDir = D;
        if (Sm) {  // DDir = S^{-1} D (W2), DD = DDir S^{-1} (Eb)
            HIPCHK(hipMemcpyAsync(W1, Sm + ko, oo * kc * sizeof(cd), hipMemcpyDeviceToDevice, s));
            HIPCHK(gj_inverse(n_orb, kc, W1, d_flag, s));
            HIPCHK((zgemm_sq<false, false>(n_orb, kc, W1, D, W2, s)));
            HIPCHK((zgemm_sq<false, false>(n_orb, kc, W2, W1, Eb, s)));
            DDir = W2;
            DD = Eb;
        }
        if (axes) {
            const size_t np = (size_t)kc * 3 * nrr;
            hipLaunchKernelGGL(k_axis_phases, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s,
                               kphase + 3 * (size_t)k0, kc, rmax, Ptab);
            HIPCHK(hipGetLastError());
            const size_t shm = sizeof(cd) * ((size_t)3 * DMR_EM * kt + (size_t)kt * 3 * nrr);
            const int nseg = std::max(1, std::min(nseg_max, kc / (2 * kt)));
            const int kseg = (kc + nseg - 1) / nseg;
            hipLaunchKernelGGL(k_dm_r_axes, dim3((unsigned)items_e.size(), (unsigned)nseg), dim3(256), shm, s, DD,
                               DDir, (const cd*)D, (const cd*)Ptab, rmax, kc, kt, oo, d_items_e, d_ielem, d_eslot,
                               d_rs, n_r, kseg, d_part);
            HIPCHK(hipGetLastError());
            const size_t cnt = 3 * (size_t)n_r;
            hipLaunchKernelGGL(k_dm_r_merge, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, (const cd*)d_part,
                               nseg, d_rs, n_r, inv_nk, (cd*)dmr);
            HIPCHK(hipGetLastError());
        } else if (n_r > 0) {
            hipLaunchKernelGGL(k_dm_r, dim3((unsigned)items.size()), dim3(256), 0, s, DD, DDir, (const cd*)D,
                               kphase + 3 * (size_t)k0, kc, oo, d_items, d_rs, n_r, inv_nk, (cd*)dmr);
            HIPCHK(hipGetLastError());
        }
    }
    int flag = 0;
    HIPCHK(hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));  // also: the host maps die here
    if (flag) return fail(h, DMFTGPU_ERR_NUMERIC, "singular transform or overlap matrix");
    return DMFTGPU_OK;
}
