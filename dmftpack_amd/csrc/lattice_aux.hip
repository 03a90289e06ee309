// Static-part densities, GwHF and the deterministic reductions of the lattice
// path.
//
// Static part (GreenFtn.cpp:311-340): with H(k) + P m0 P^H = U diag(eps) U^H,
// the G0 subtraction and the Fermi add-back of the density collapse per level
// to the weight
//     w_kl = f(eps_kl - mu) - (1/beta) sum_{n < high*N_freq} 2 Re 1/(iw_n + mu - eps_kl)
// so that NumMatrix_static = sum_k (U diag(w) U^H)_cc and N_static = sum w;
// dN/dmu gets beta f (1-f) + (1/beta) sum_n 2 Re 1/(iw_n+mu-eps)^2 per level.
//
// Closed form (M = high*N_freq frequencies, x = mu - eps, a = pi/beta): the
// full Matsubara sums are sum_{n>=0} 2x/(x^2+w_n^2) = (beta/2) tanh(beta x/2)
// and, its x-derivative, sum_{n>=0} 2 Re (x+iw_n)^-2 = -(beta^2/4) sech^2, while
// f = (1 + tanh(beta x/2))/2 and beta f (1-f) = (beta/4) sech^2, so
//     w_kl = 1/2 + T/beta,   dN/dmu term = -T2/beta
// with the tails T = sum_{n>=M} 2 Re 1/(x+iw_n), T2 = sum_{n>=M} 2 Re 1/(x+iw_n)^2.
// The tails follow from the Euler-Maclaurin formula with v = a(2M+1),
// z = x + iv, w_j = z^-j (terms through the fifth derivative; the remainder is
// ~ (2M+1)^-7 of the tail):
//     T  = atan(x/v)/a + Re w1 - (a/3) Im w2 - (2/15) a^3 Im w4 - (16/63) a^5 Im w6
//     T2 = Im w1 / a   + Re w2 - (2a/3) Im w3 - (8/15) a^3 Im w5 - (32/21) a^5 Im w7
// The reference's clamp of f at |beta(eps - mu)| > 100 changes the result by
// < e^-100.  Below M = 256 the finite sums are taken term by term.
#include "lattice_kernels.hpp"

namespace dg {

__device__ __forceinline__ void matsubara_tails(double x, double beta, int M, double& T, double& T2)
{
    const double a = kPi / beta;
    const double v = a * (2.0 * M + 1.0);
    const cd w1 = crcp(cd{x, v});
    const cd w2 = cmul(w1, w1), w3 = cmul(w2, w1), w4 = cmul(w2, w2);
    const cd w5 = cmul(w4, w1), w6 = cmul(w4, w2), w7 = cmul(w6, w1);
    const double a3 = a * a * a, a5 = a3 * a * a;
    T = atan(x / v) / a + w1.x - (a / 3.0) * w2.y - (2.0 / 15.0) * a3 * w4.y - (16.0 / 63.0) * a5 * w6.y;
    T2 = w1.y / a + w2.x - (2.0 * a / 3.0) * w3.y - (8.0 / 15.0) * a3 * w5.y - (32.0 / 21.0) * a5 * w7.y;
}

__global__ __launch_bounds__(256) void k_static_terms(const double* __restrict__ ev, int total, int nw_tot,
                                                      double beta, double mu, double* __restrict__ wt,
                                                      double* __restrict__ dt)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const double e = ev[t];
    const double x = mu - e;
    if (nw_tot >= 256) {
        double T, T2;
        matsubara_tails(x, beta, nw_tot, T, T2);
        wt[t] = 0.5 + T / beta;
        dt[t] = -T2 / beta;
        return;
    }
    const double x2 = x * x;
    double s = 0.0, s2 = 0.0;
    for (int nw = 0; nw < nw_tot; ++nw) {
        const double w = matsubara(nw, beta);
        const double q = fma(w, w, x2);
        const double den = rcp_pos(q);
        s = fma(2.0 * x, den, s);                     // 2 Re 1/(x + iw)
        s2 = fma(2.0 * (x2 - w * w), den * den, s2);  // 2 Re 1/(x + iw)^2
    }
    const double xb = beta * (e - mu);
    double f;
    if (xb > 100.0) f = 0.0;  // clamp of GreenFtn.cpp:314-316
    else if (xb < -100.0) f = 1.0;
    else f = 1.0 / (1.0 + exp(xb));
    wt[t] = f - s / beta;
    dt[t] = beta * f * (1.0 - f) + s2 / beta;
}

hipError_t launch_static_terms(const double* ev, int n, int nk, int nw_tot, double beta, double mu, double* wt,
                               double* dt, hipStream_t s)
{
    const int total = n * nk;
    hipLaunchKernelGGL(k_static_terms, dim3((total + 255) / 256), dim3(256), 0, s, ev, total, nw_tot, beta, mu, wt,
                       dt);
    return hipGetLastError();
}

// One workgroup per k chunk.  Thread e < m2 accumulates element e of
// sum_k U diag(w) U^H over the chunk; the level sums are reduced by the block.
__global__ __launch_bounds__(256) void k_static_reduce(const double* __restrict__ wt, const double* __restrict__ dt,
                                                       const cd* __restrict__ U, int n, int nc, int nk,
                                                       int k_per_chunk, cd* __restrict__ pNum,
                                                       double* __restrict__ pN, double* __restrict__ pdN)
{
    __shared__ double red[2][256];
    const int ch = blockIdx.x;
    const int k0 = ch * k_per_chunk, k1 = min(nk, k0 + k_per_chunk);
    const int m2 = nc * nc;
    // groups of m2 threads take every ng-th k of the chunk; the group partials
    // are added in group order (deterministic)
    // (m2 > 256: one element per thread, strided)
    __shared__ cd part[256];
    const int ng = m2 <= 256 ? 256 / m2 : 1;
    if (m2 <= 256) {
        const int grp = threadIdx.x / m2, e = threadIdx.x % m2;
        if (grp < ng) {
            const int i = e % nc, j = e / nc;
            cd acc = cd{0.0, 0.0};
            for (int k = k0 + grp; k < k1; k += ng) {
                const cd* Uk = U + (size_t)k * n * n;
                const double* w = wt + (size_t)k * n;
                for (int l = 0; l < n; ++l) acc = cadd(acc, cscale(cmulc(Uk[i + l * n], Uk[j + l * n]), w[l]));
            }
            part[threadIdx.x] = acc;
        }
        __syncthreads();
        for (int t = threadIdx.x; t < m2; t += blockDim.x) {
            cd acc = cd{0.0, 0.0};
            for (int g = 0; g < ng; ++g) acc = cadd(acc, part[g * m2 + t]);
            pNum[(size_t)ch * m2 + t] = acc;
        }
    } else {
        for (int e = threadIdx.x; e < m2; e += blockDim.x) {
            const int i = e % nc, j = e / nc;
            cd acc = cd{0.0, 0.0};
            for (int k = k0; k < k1; ++k) {
                const cd* Uk = U + (size_t)k * n * n;
                const double* w = wt + (size_t)k * n;
                for (int l = 0; l < n; ++l) acc = cadd(acc, cscale(cmulc(Uk[i + l * n], Uk[j + l * n]), w[l]));
            }
            pNum[(size_t)ch * m2 + e] = acc;
        }
    }
    double sN = 0.0, sD = 0.0;
    for (int t = k0 * n + threadIdx.x; t < k1 * n; t += blockDim.x) {
        sN += wt[t];
        sD += dt[t];
    }
    red[0][threadIdx.x] = sN;
    red[1][threadIdx.x] = sD;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            red[0][threadIdx.x] += red[0][threadIdx.x + o];
            red[1][threadIdx.x] += red[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pN[ch] = red[0][0];
        pdN[ch] = red[1][0];
    }
}

hipError_t launch_static_reduce(const double* wt, const double* dt, const cd* U, int n, int nc, int nk,
                                int k_per_chunk, int n_chunks, cd* pNum, double* pN, double* pdN, hipStream_t s)
{
    hipLaunchKernelGGL(k_static_reduce, dim3(n_chunks), dim3(256), 0, s, wt, dt, U, n, nc, nk, k_per_chunk, pNum, pN,
                       pdN);
    return hipGetLastError();
}

// Core-band term of GiwToNele (GreenFtn.cpp:622-631): the reference adds
// f(eps_kl - mu) over all NumOrbit Kohn-Sham bands and subtracts it over the
// k point's window bands, i.e. it counts the bands outside the window.  Here
// those bands are flagged (out[k][l] = 1) and summed directly, f unclamped as
// in the reference, with the mu-derivative beta f (1 - f).  One workgroup per
// chunk of k_per_chunk k points; fixed-order block sums into pN[ch], pdN[ch].
__global__ __launch_bounds__(256) void k_core_bands(const double* __restrict__ eps,
                                                    const unsigned char* __restrict__ out, int n_orb, int nk,
                                                    int k_per_chunk, double beta, double mu, double* __restrict__ pN,
                                                    double* __restrict__ pdN)
{
    __shared__ double red[2][256];
    const int ch = blockIdx.x;
    const size_t t0 = (size_t)ch * k_per_chunk * n_orb;
    const size_t t1 = (size_t)min(nk, (ch + 1) * k_per_chunk) * n_orb;
    double sN = 0.0, sD = 0.0;
    for (size_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
        if (!out[t]) continue;
        const double f = 1.0 / (1.0 + exp(beta * (eps[t] - mu)));
        sN += f;
        sD += beta * f * (1.0 - f);
    }
    red[0][threadIdx.x] = sN;
    red[1][threadIdx.x] = sD;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            red[0][threadIdx.x] += red[0][threadIdx.x + o];
            red[1][threadIdx.x] += red[1][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pN[ch] = red[0][0];
        pdN[ch] = red[1][0];
    }
}

hipError_t launch_core_bands(const double* eps, const unsigned char* out, int n_orb, int nk, int k_per_chunk,
                             int n_chunks, double beta, double mu, double* pN, double* pdN, hipStream_t s)
{
    hipLaunchKernelGGL(k_core_bands, dim3(n_chunks), dim3(256), 0, s, eps, out, n_orb, nk, k_per_chunk, beta, mu, pN,
                       pdN);
    return hipGetLastError();
}

// GwHF: thread per (nw, e); loops over all k and levels.  Optional output,
// written only to files by the reference (GreenFtn.cpp:394-407).
__global__ __launch_bounds__(256) void k_gloc_hf(const double* __restrict__ ev, const cd* __restrict__ U, int n,
                                                 int nc, int nk, int n_freq, double beta, double mu, double inv_nkg,
                                                 int npc, cd* __restrict__ out)
{
    const int m2 = nc * nc;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_freq * m2) return;
    const int nw = t / m2, e = t % m2, i = e % nc, j = e / nc;
    if (i / npc != j / npc) {
        out[t] = cd{0.0, 0.0};
        return;
    }
    const double w = matsubara(nw, beta);
    cd acc = cd{0.0, 0.0};
    for (int k = 0; k < nk; ++k) {
        const cd* Uk = U + (size_t)k * n * n;
        const double* ek = ev + (size_t)k * n;
        for (int l = 0; l < n; ++l) {
            const cd g = crcp(cd{mu - ek[l], w});
            acc = cfma(cmulc(Uk[i + l * n], Uk[j + l * n]), g, acc);
        }
    }
    out[t] = cscale(acc, inv_nkg);
}

hipError_t launch_gloc_hf(const double* ev, const cd* U, int n, int nc, int nk, int n_freq, double beta, double mu,
                          double inv_nkg, int n_cluster, cd* out, hipStream_t s)
{
    const int total = n_freq * nc * nc;
    hipLaunchKernelGGL(k_gloc_hf, dim3((total + 255) / 256), dim3(256), 0, s, ev, U, n, nc, nk, n_freq, beta, mu,
                       inv_nkg, nc / n_cluster, out);
    return hipGetLastError();
}

// Sum of n_chunks partial arrays part[c * len + i] (one per k chunk).  A
// workgroup covers 32 consecutive entries; each half-wavefront sums one of 8
// contiguous ranges of chunks, in chunk order, and the 8 range sums are added
// in range order.  The result is deterministic for a given (n_chunks, len),
// and the chunk loop is 8 times shorter than a one-thread-per-entry sum, which
// matters for the short per-frequency arrays (traces, dL/dmu) whose grids
// would otherwise hold a few workgroups walking hundreds of chunks.
__device__ __forceinline__ double radd(double a, double b) { return a + b; }
__device__ __forceinline__ cd radd(cd a, cd b) { return cadd(a, b); }

constexpr int kRedX = 32, kRedY = 8;

template <class T>
__global__ __launch_bounds__(kRedX* kRedY) void k_reduce_chunks(const T* __restrict__ part, int n_chunks, size_t len,
                                                                 T* __restrict__ out)
{
    __shared__ T red[kRedY][kRedX];
    const int tx = threadIdx.x % kRedX, ty = threadIdx.x / kRedX;
    const size_t i = (size_t)blockIdx.x * kRedX + tx;
    const int c0 = (int)((long long)n_chunks * ty / kRedY);
    const int c1 = (int)((long long)n_chunks * (ty + 1) / kRedY);
    T acc{};
    if (i < len) {
#pragma unroll 4
        for (int c = c0; c < c1; ++c) acc = radd(acc, part[(size_t)c * len + i]);
    }
    red[ty][tx] = acc;
    __syncthreads();
    if (ty == 0 && i < len) {
        T sum = red[0][tx];
        for (int y = 1; y < kRedY; ++y) sum = radd(sum, red[y][tx]);
        out[i] = sum;
    }
}

template <class T>
static hipError_t launch_reduce(const T* part, int n_chunks, size_t len, T* out, hipStream_t s)
{
    if (len == 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_chunks<T>, dim3((unsigned)((len + kRedX - 1) / kRedX)), dim3(kRedX * kRedY), 0, s,
                       part, n_chunks, len, out);
    return hipGetLastError();
}

hipError_t launch_reduce_chunks(const cd* part, int n_chunks, size_t len, cd* out, hipStream_t s)
{
    return launch_reduce(part, n_chunks, len, out, s);
}

hipError_t launch_reduce_chunks_d(const double* part, int n_chunks, size_t len, double* out, hipStream_t s)
{
    return launch_reduce(part, n_chunks, len, out, s);
}

// One workgroup per output quantity q: q < m2 -> element e=q of
// sum_nw (G + G^H)[e]; q = m2 -> sum 2 Re tr; m2+1 -> sum 2 Re tr2;
// m2+2 -> sum Re mu.  Fixed-order block reduction.
__global__ __launch_bounds__(256) void k_freq_totals(const cd* __restrict__ G, int m2, int nc,
                                                     const cd* __restrict__ tr, const cd* __restrict__ tr2, int nw_tot,
                                                     const cd* __restrict__ mu, int n_mu, cd* __restrict__ out)
{
    __shared__ double rx[256], ry[256];
    const int q = blockIdx.x;
    double ax = 0.0, ay = 0.0;
    if (q < m2) {
        if (G) {
            const int i = q % nc, j = q / nc, et = j + i * nc;
            for (int nw = threadIdx.x; nw < nw_tot; nw += 256) {
                const cd g = G[(size_t)nw * m2 + q], gt = G[(size_t)nw * m2 + et];
                ax += g.x + gt.x;
                ay += g.y - gt.y;
            }
        }
    } else if (q == m2) {
        if (tr)
            for (int nw = threadIdx.x; nw < nw_tot; nw += 256) ax += 2.0 * tr[nw].x;
    } else if (q == m2 + 1) {
        if (tr2)
            for (int nw = threadIdx.x; nw < nw_tot; nw += 256) ax += 2.0 * tr2[nw].x;
    } else {
        if (mu)
            for (int nw = threadIdx.x; nw < n_mu; nw += 256) ax += mu[nw].x;
    }
    rx[threadIdx.x] = ax;
    ry[threadIdx.x] = ay;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            rx[threadIdx.x] += rx[threadIdx.x + o];
            ry[threadIdx.x] += ry[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = cd{rx[0], ry[0]};
}

hipError_t launch_freq_totals(const cd* G, int m2, int nc, const cd* tr, const cd* tr2, int nw_tot, const cd* mu,
                              int n_mu, cd* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_freq_totals, dim3(m2 + 3), dim3(256), 0, s, G, m2, nc, tr, tr2, nw_tot, mu, n_mu, out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_gloc_finalize(const cd* __restrict__ Gsum, int n_freq, int nc, int npc,
                                                       double scale, cd* __restrict__ gloc)
{
    const int m2 = nc * nc;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_freq * m2) return;
    const int e = t % m2, i = e % nc, j = e / nc;
    gloc[t] = (i / npc == j / npc) ? cscale(Gsum[t], scale) : cd{0.0, 0.0};
}

hipError_t launch_gloc_finalize(const cd* Gsum, int n_freq, int nc, int n_cluster, double scale, cd* gloc,
                                hipStream_t s)
{
    const int total = n_freq * nc * nc;
    hipLaunchKernelGGL(k_gloc_finalize, dim3((total + 255) / 256), dim3(256), 0, s, Gsum, n_freq, nc, nc / n_cluster,
                       scale, gloc);
    return hipGetLastError();
}

// out[m] = scale * in[m]^H for a stack of d x d matrices (column-major)
__global__ __launch_bounds__(256) void k_adjoint_stack(const cd* __restrict__ in, size_t count, int d, double scale,
                                                       cd* __restrict__ out)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t m2 = (size_t)d * d;
    if (t >= count * m2) return;
    const size_t m = t / m2;
    const int e = (int)(t % m2), i = e % d, j = e / d;
    out[t] = cscale(cconj(in[m * m2 + j + (size_t)i * d]), scale);
}

hipError_t launch_adjoint_stack(const cd* in, size_t count, int d, double scale, cd* out, hipStream_t s)
{
    const size_t total = count * d * d;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_adjoint_stack, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, count, d, scale,
                       out);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_scale(const cd* __restrict__ in, size_t len, double scale,
                                               cd* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) out[i] = cscale(in[i], scale);
}

hipError_t launch_scale(const cd* in, size_t len, double scale, cd* out, hipStream_t s)
{
    if (len == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, in, len, scale, out);
    return hipGetLastError();
}

// Loss cotangents of L = sum ||G_loc - G_tgt||^2: gbar = 2 (G_loc - G_tgt) on the
// entries inside a cluster block, 0 across clusters (G_loc is returned
// block-diagonal, GreenFtn.cpp:371-383, so those entries do not depend on Sigma);
// *loss = sum |G_loc - G_tgt|^2 over all entries.  One workgroup of 1024
// threads (the whole array is n_freq * nc^2 entries, tens of thousands),
// fixed-order tree sum.
constexpr int kLossThreads = 1024;

__global__ __launch_bounds__(kLossThreads) void k_loss_cotangent(const cd* __restrict__ gloc,
                                                                 const cd* __restrict__ gtgt, int n_freq, int nc,
                                                                 int npc, cd* __restrict__ gbar,
                                                                 double* __restrict__ loss)
{
    __shared__ double red[kLossThreads];
    const int m2 = nc * nc;
    const size_t len = (size_t)n_freq * m2;
    double acc = 0.0;
    for (size_t t = threadIdx.x; t < len; t += kLossThreads) {
        const cd d = csub(gloc[t], gtgt[t]);
        acc = fma(d.x, d.x, fma(d.y, d.y, acc));
        const int e = (int)(t % m2), i = e % nc, j = e / nc;
        gbar[t] = (i / npc == j / npc) ? cscale(d, 2.0) : cd{0.0, 0.0};
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = kLossThreads / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = red[0];
}

hipError_t launch_loss_cotangent(const cd* gloc, const cd* gtgt, int n_freq, int nc, int npc, cd* gbar, double* loss,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(k_loss_cotangent, dim3(1), dim3(kLossThreads), 0, s, gloc, gtgt, n_freq, nc, npc, gbar, loss);
    return hipGetLastError();
}

}  // namespace dg
