// The callers either side of the lattice path:
//   * dmftgpu_hyb_delta  -- hybridisation function of one impurity from G_loc
//     (Construct_hyb_delta, Construct_hyb_delta.cpp:47-119)
//   * dmftgpu_hk_from_hr -- H(k) from a hopping list H(R)
//     (ConstructModelHamiltonian.cpp:58-77, hopping list of read_HR,
//     read_TB_parm.cpp:6-76)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <vector>

#include "../../include/dmftgpu.h"
#include "api_internal.hpp"
#include "gj.hpp"

using dg::cd;
using namespace dgi;

#define HIPCHK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(h, DMFTGPU_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace {

using namespace dg;

struct Orbits {
    int o[32];
};

// Delta(iw_n) = (iw_n + mu) I - H_imp - Sigma_imp(iw_n) - G_imp(iw_n)^{-1}:
// a slot of N lanes per frequency, lane c gathers column c of G_imp (the
// orbit sub-block of G_loc), inverts it with the column Gauss-Jordan of
// gj.hpp (G_loc of a causal lattice has -Im G_loc > 0, so -i G_imp has a
// positive definite Hermitian part and needs no pivoting), and writes column c
// of Delta.
template <int N>
__global__ __launch_bounds__(64) void k_hyb_delta(const cd* __restrict__ gloc, const cd* __restrict__ sigma,
                                                  const cd* __restrict__ himp, int nc, int n_freq, double beta,
                                                  double mu, int dim, Orbits orb, cd* __restrict__ delta)
{
    constexpr int SPW = 64 / N;
    __shared__ cd sb[(SPW + 1) * 2 * N];
    const int lane = threadIdx.x;
    const int sl = lane / N, c = lane - sl * N;
    const bool active = sl < SPW;
    const int nw = blockIdx.x * SPW + (active ? sl : 0);
    const bool valid = active && nw < n_freq;
    const int nwc = nw < n_freq ? nw : n_freq - 1;
    const bool live = valid && c < dim;
    const int oc = orb.o[live ? c : 0];
    const size_t m2 = (size_t)nc * nc;
    cd col[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        cd v = (i == c) ? cd{1.0, 0.0} : cd{0.0, 0.0};
        if (live) v = (i < dim) ? gloc[(size_t)nwc * m2 + orb.o[i < dim ? i : 0] + (size_t)oc * nc] : cd{0.0, 0.0};
        col[i] = v;
    }
    gj_columns<N>(col, sb + (size_t)sl * 2 * N, c, dim);
    if (!live) return;
    const double w = matsubara(nwc, beta);
    const size_t d2 = (size_t)dim * dim;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (i < dim) {
            const int oi = orb.o[i];
            cd v = cneg(cadd(himp[oi + (size_t)oc * nc], sigma[(size_t)nwc * m2 + oi + (size_t)oc * nc]));
            v = csub(v, col[i]);
            if (i == c) v = cadd(v, cd{mu, w});
            delta[(size_t)nw * d2 + i + (size_t)c * dim] = v;
        }
    }
}

template <int N>
hipError_t launch_hyb(const cd* gloc, const cd* sigma, const cd* himp, int nc, int n_freq, double beta, double mu,
                      int dim, const Orbits& orb, cd* delta, hipStream_t s)
{
    constexpr int SPW = 64 / N;
    hipLaunchKernelGGL(k_hyb_delta<N>, dim3((n_freq + SPW - 1) / SPW), dim3(64), 0, s, gloc, sigma, himp, nc, n_freq,
                       beta, mu, dim, orb, delta);
    return hipGetLastError();
}

// H(k)_e = sum over the hops of element e of t exp(-i (phi_x R1 + phi_y R2 + phi_z R3)),
// phi = k.a per axis.  One thread per (k, element); the hops are grouped by
// element (CSR), so each output is a plain sum in the order of the input list.
__global__ void k_hk_from_hr(const int* __restrict__ start, const int4* __restrict__ hr, const cd* __restrict__ t,
                             const double* __restrict__ kphase, int n, int nk, cd* __restrict__ Hk)
{
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nn = (size_t)n * n;
    if (gid >= nn * nk) return;
    const int k = (int)(gid / nn);
    const int e = (int)(gid % nn);
    const double px = kphase[3 * (size_t)k], py = kphase[3 * (size_t)k + 1], pz = kphase[3 * (size_t)k + 2];
    cd acc = cd{0.0, 0.0};
    for (int q = start[e]; q < start[e + 1]; ++q) {
        const int4 r = hr[q];
        double sn, cs;
        sincos(px * r.x + py * r.y + pz * r.z, &sn, &cs);
        acc = cfma(t[q], cd{cs, -sn}, acc);
    }
    Hk[gid] = acc;
}

// The same sum with the phase taken apart per axis, exp(-i phi.R) =
// ex[R1] ey[R2] ez[R3] with ea[r] = exp(-i phi_a r), |r| <= rmax, and a tile of
// KT k points per workgroup: the tile's axis tables are formed once into LDS,
// and each thread takes one matrix element for all KT k, so each hop is read
// once per tile instead of once per k.
template <int KT>
__global__ __launch_bounds__(256) void k_hk_from_hr_axes(const int* __restrict__ start, const int4* __restrict__ hr,
                                                         const cd* __restrict__ t, const double* __restrict__ kphase,
                                                         int n, int nk, int rmax, cd* __restrict__ Hk)
{
    extern __shared__ cd tab[];  // [KT][3][2 rmax + 1]
    const int k0 = blockIdx.y * KT;
    const int nr = 2 * rmax + 1, per = 3 * nr;
    for (int q = threadIdx.x; q < KT * per; q += blockDim.x) {
        const int kk = q / per, ax = (q % per) / nr, r = q % nr - rmax;
        const int k = min(k0 + kk, nk - 1);
        double sn, cs;
        sincos(kphase[3 * (size_t)k + ax] * r, &sn, &cs);
        tab[q] = cd{cs, -sn};
    }
    __syncthreads();
    const int nn = n * n;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nn) return;
    cd acc[KT];
#pragma unroll
    for (int kk = 0; kk < KT; ++kk) acc[kk] = cd{0.0, 0.0};
    for (int q = start[e]; q < start[e + 1]; ++q) {
        const int4 r = hr[q];
        const cd tq = t[q];
#pragma unroll
        for (int kk = 0; kk < KT; ++kk) {
            const cd* p = tab + kk * per + rmax;
            acc[kk] = cfma(tq, cmul(cmul(p[r.x], p[nr + r.y]), p[2 * nr + r.z]), acc[kk]);
        }
    }
#pragma unroll
    for (int kk = 0; kk < KT; ++kk)
        if (k0 + kk < nk) Hk[(size_t)(k0 + kk) * nn + e] = acc[kk];
}

}  // namespace

// Construct_hyb_delta (Construct_hyb_delta.cpp:47-119), off-diagonal form (:96-117)
extern "C" int dmftgpu_hyb_delta(dmftgpu_handle h, int nc, int n_freq, double beta, double mu, const double* gloc,
                                 const double* sigma, const double* h_imp, int dim, const int* orbit, double* delta)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (nc < 1 || nc > 32 || n_freq < 1 || !(beta > 0.0) || !std::isfinite(mu))
        return fail(h, DMFTGPU_ERR_ARG, "nc=%d n_freq=%d beta=%g mu=%g", nc, n_freq, beta, mu);
    if (dim < 0 || dim > nc) return fail(h, DMFTGPU_ERR_ARG, "dim=%d outside 0..nc", dim);
    if (dim == 0) return DMFTGPU_OK;  // the reference skips the inverse for an empty impurity (:98)
    if (!gloc || !sigma || !h_imp || !orbit || !delta) return fail(h, DMFTGPU_ERR_ARG, "null argument");
    Orbits orb{};
    for (int a = 0; a < dim; ++a) {
        if (orbit[a] < 0 || orbit[a] >= nc) return fail(h, DMFTGPU_ERR_ARG, "orbit[%d]=%d outside 0..nc", a, orbit[a]);
        orb.o[a] = orbit[a];
    }
    const cd *g = (const cd*)gloc, *sg = (const cd*)sigma, *hi = (const cd*)h_imp;
    cd* out = (cd*)delta;
    if (dim <= 4) HIPCHK(launch_hyb<4>(g, sg, hi, nc, n_freq, beta, mu, dim, orb, out, h->stream));
    else if (dim <= 8) HIPCHK(launch_hyb<8>(g, sg, hi, nc, n_freq, beta, mu, dim, orb, out, h->stream));
    else if (dim <= 16) HIPCHK(launch_hyb<16>(g, sg, hi, nc, n_freq, beta, mu, dim, orb, out, h->stream));
    else HIPCHK(launch_hyb<32>(g, sg, hi, nc, n_freq, beta, mu, dim, orb, out, h->stream));
    return DMFTGPU_OK;
}

// H(k) = sum_R H(R) exp(-i k.a.R)  (ConstructModelHamiltonian.cpp:58-77)
extern "C" int dmftgpu_hk_from_hr(dmftgpu_handle h, int n, int n_hop, const int* hop_idx, const double* hop, int n_k,
                                  const double* kphase, double* Hk)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (n < 1 || n > 4096 || n_hop < 0 || n_k < 1) return fail(h, DMFTGPU_ERR_ARG, "n=%d n_hop=%d n_k=%d", n, n_hop, n_k);
    if ((n_hop > 0 && (!hop_idx || !hop)) || !kphase || !Hk) return fail(h, DMFTGPU_ERR_ARG, "null argument");
    // group the hops by matrix element, keeping their input order
    const size_t nn = (size_t)n * n;
    std::vector<int> start(nn + 1, 0);
    for (int q = 0; q < n_hop; ++q) {
        const int i = hop_idx[5 * q + 3], j = hop_idx[5 * q + 4];
        if (i < 0 || i >= n || j < 0 || j >= n)
            return fail(h, DMFTGPU_ERR_ARG, "hop %d: orbital pair (%d,%d) outside 0..%d", q, i, j, n - 1);
        ++start[i + (size_t)j * n + 1];
    }
    for (size_t e = 0; e < nn; ++e) start[e + 1] += start[e];
    std::vector<int> fill(start.begin(), start.end() - 1);
    int rmax = 0;
    for (int q = 0; q < n_hop; ++q)
        for (int c = 0; c < 3; ++c) rmax = std::max(rmax, std::abs(hop_idx[5 * q + c]));
    std::vector<int> rq(4 * (size_t)(n_hop > 0 ? n_hop : 1), 0);
    std::vector<double> tq(2 * (size_t)(n_hop > 0 ? n_hop : 1), 0.0);
    for (int q = 0; q < n_hop; ++q) {
        const size_t e = hop_idx[5 * q + 3] + (size_t)hop_idx[5 * q + 4] * n;
        const int p = fill[e]++;
        rq[4 * (size_t)p] = hop_idx[5 * q];
        rq[4 * (size_t)p + 1] = hop_idx[5 * q + 1];
        rq[4 * (size_t)p + 2] = hop_idx[5 * q + 2];
        tq[2 * (size_t)p] = hop[2 * q];
        tq[2 * (size_t)p + 1] = hop[2 * q + 1];
    }
    const size_t b_start = ((nn + 1) * sizeof(int) + 255) / 256 * 256;
    const size_t b_r = (rq.size() * sizeof(int) + 255) / 256 * 256;
    const size_t b_t = tq.size() * sizeof(double);
    hipError_t e;
    char* base = (char*)h->ws.get(b_start + b_r + b_t, &e);
    if (!base) return fail(h, DMFTGPU_ERR_HIP, "workspace: %s", hipGetErrorString(e));
    int* dstart = (int*)base;
    int4* dr = (int4*)(base + b_start);
    cd* dt = (cd*)(base + b_start + b_r);
    HIPCHK(hipMemcpyAsync(dstart, start.data(), (nn + 1) * sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(dr, rq.data(), rq.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(dt, tq.data(), b_t, hipMemcpyHostToDevice, h->stream));
    if (rmax <= 64 && (n_k + 7) / 8 <= 65535) {  // per-axis phase tables in LDS, 8 k points per workgroup
        constexpr int KT = 8;
        const unsigned wg = nn >= 256 ? 256u : (unsigned)((nn + 63) / 64 * 64);
        const dim3 grid((unsigned)((nn + wg - 1) / wg), (unsigned)((n_k + KT - 1) / KT));
        hipLaunchKernelGGL(k_hk_from_hr_axes<KT>, grid, dim3(wg), sizeof(cd) * KT * 3 * (2 * rmax + 1), h->stream,
                           dstart, dr, dt, kphase, n, n_k, rmax, (cd*)Hk);
    } else {
        const size_t total = nn * (size_t)n_k;
        hipLaunchKernelGGL(k_hk_from_hr, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, h->stream, dstart, dr, dt,
                           kphase, n, n_k, (cd*)Hk);
    }
    HIPCHK(hipGetLastError());
    // the host vectors die here: wait for the copies that read them
    HIPCHK(hipStreamSynchronize(h->stream));
    return DMFTGPU_OK;
}
