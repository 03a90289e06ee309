// Matsubara tail moments and the iw <-> tau transforms.
//
//   dmftgpu_moments   getAsymto_moments   (FourierTransform.cpp:184-465)
//   dmftgpu_ft_w2tau  FourierTransform_ + FourierTransform_tau (:126-178, :16-118)
//   dmftgpu_ft_tau2w  FT_t_to_w           (:521-561)
//
// GPU work: the sliding-window least-squares fits (one thread per window start
// and matrix element, summing in the reference's ascending order), the window
// averages, the batched positivity tests, and both transforms as tiled complex
// matrix products over frequency (tau) with the twiddles computed from the
// same FP64 phase w*tau the reference uses.  The tail of FourierTransform_tau
// does not depend on tau, so it is subtracted once per frequency, not once
// per (tau, frequency) as in the reference loop (:55-61).
// Host work: the small d x d set-up (square roots, inverses, eigenbases) and
// the sequential selection / patch logic, restated from the reference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/dmftgpu.h"
#include "api_internal.hpp"
#include "host_linalg.hpp"
#include "lattice_kernels.hpp"

using dg::cd;
using dgi::fail;

namespace {

using namespace dg;

// ---------------------------------------------------------------------------
// getAsymto_moments kernels
// ---------------------------------------------------------------------------
// Partial-pivot LU inverse of the real n x n (n <= 4) normal matrix, the
// steps of Eigen's PartialPivLU::inverse: pivot = the first largest |entry|
// of the column, multipliers by the reciprocal pivot, fused updates, forward
// then backward substitution per unit column.  Every multiply-add is an
// explicit fma and nothing else is contracted, so the oracle's restatement
// (which does the same) gives the same bits.
__device__ void lu_inverse4(const double (&A)[16], int n, double (&X)[16])
{
#pragma clang fp contract(off)
    double lu[16];
    int perm[4];
    for (int e = 0; e < 16; ++e) lu[e] = A[e];
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(lu[k + k * 4]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(lu[i + k * 4]) > best) {
                best = fabs(lu[i + k * 4]);
                p = i;
            }
        if (p != k) {
            for (int j = 0; j < n; ++j) {
                double t = lu[k + j * 4];
                lu[k + j * 4] = lu[p + j * 4];
                lu[p + j * 4] = t;
            }
            int t = perm[k];
            perm[k] = perm[p];
            perm[p] = t;
        }
        if (best != 0.0) {
            const double inv = 1.0 / lu[k + k * 4];
            for (int i = k + 1; i < n; ++i) lu[i + k * 4] *= inv;
        }
        for (int j = k + 1; j < n; ++j) {
            const double u = lu[k + j * 4];
            for (int i = k + 1; i < n; ++i) lu[i + j * 4] = fma(-lu[i + k * 4], u, lu[i + j * 4]);
        }
    }
    for (int c = 0; c < n; ++c) {
        double x[4];
        for (int i = 0; i < n; ++i) x[i] = (perm[i] == c) ? 1.0 : 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < i; ++j) x[i] = fma(-lu[i + j * 4], x[j], x[i]);
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) x[i] = fma(-lu[i + j * 4], x[j], x[i]);
            x[i] /= lu[i + i * 4];
        }
        for (int i = 0; i < n; ++i) X[i + c * 4] = x[i];
    }
}

// columns of the design matrix for row pair jj (Re row 2jj, Im row 2jj+1)
// non-G: Re: [1, 0, -1/z^2, 0]   Im: [0, -1/z, 0, 1/z^3]   (:227-230)
// G    : Re: [-1/z^2, 0]          Im: [0, 1/z^3]             (:366-367)
// basis values per frequency, tab[wn] = {-1/z, -1/z^2, 1/z^3, 1/z}, formed on
// the host with std::pow as the reference forms them (dmftgpu_moments)
// K^T K of window w, entry (a, b) with a <= b (the matrix is symmetric): one
// thread per (window, entry), summing the rows in order, Re row then Im row of
// each frequency, one fma per row (zero entries included).
__global__ void k_mom_ktk(int n_freq, int range, int green, const double4* __restrict__ tab, double* __restrict__ ktk)
{
#pragma clang fp contract(off)
    const int nm = green ? 2 : 4;
    const int npair = nm * (nm + 1) / 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= range * npair) return;
    const int w = t / npair;
    int q = t % npair, b = 0;
    while (q > b) {  // pair index q -> (a, b), a <= b, column by column
        q -= b + 1;
        ++b;
    }
    const int a = q;
    double K = 0.0;
    for (int wn = w; wn < n_freq; ++wn) {
        const double4 c = tab[wn];
        double re[4] = {0, 0, 0, 0}, im[4] = {0, 0, 0, 0};
        if (!green) {
            re[0] = 1.0;
            im[1] = c.x;
            re[2] = c.y;
            im[3] = c.z;
        } else {
            re[0] = c.y;
            im[1] = c.z;
        }
        K = fma(re[a], re[b], K);
        K = fma(im[a], im[b], K);
    }
    ktk[(size_t)w * 16 + a + b * 4] = K;
    ktk[(size_t)w * 16 + b + a * 4] = K;
}

// inverse of each window's normal matrix, in place (one thread per window)
__global__ void k_mom_kinv(int range, int green, double* __restrict__ kinv)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= range) return;
    const int nm = green ? 2 : 4;
    double K[16], X[16];
    for (int e = 0; e < 16; ++e) K[e] = kinv[(size_t)w * 16 + e];
    lu_inverse4(K, nm, X);
    for (int e = 0; e < 16; ++e) kinv[(size_t)w * 16 + e] = X[e];
}

// mt[a][w][e] = sum_rows KM_solve[w][a][row] g_e[row], with the rows of
// KM_solve = (K^T K)^{-1} K^T formed first, as the reference associates the
// product (FourierTransform.cpp:239-243): per frequency the Re row of the
// design matrix contributes sum_b Kinv[a][b] Re_b and the Im row
// sum_b Kinv[a][b] Im_b (the rows' zero entries left out), each taken times
// its half of g and added, Re row first, in ascending frequency order.
__global__ void k_mom_fit(const cd* __restrict__ fw, int d, int n_freq, int range, const double4* __restrict__ tab,
                          int green, const double* __restrict__ kinv, cd* __restrict__ mt)
{
#pragma clang fp contract(off)
    const int m2 = d * d;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)range * m2) return;
    const int w = (int)(t / m2), e = (int)(t % m2);
    const int al = e % d, be = e / d;
    const int nm = green ? 2 : 4;
    const double* X = kinv + (size_t)w * 16;  // Kinv(a, b) at X[a + 4 b]
    double xk[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) xk[a][b] = X[a + 4 * b];
    cd s[4] = {cd{0, 0}, cd{0, 0}, cd{0, 0}, cd{0, 0}};
    for (int wn = w; wn < n_freq; ++wn) {
        const double4 c = tab[wn];
        const cd fab = fw[(size_t)wn * m2 + al + be * d];
        const cd fba = fw[(size_t)wn * m2 + be + al * d];
        // Re = (F_ab + conj(F_ba))/2 ; Im = (F_ab - conj(F_ba))/(2i)
        const cd re = cd{0.5 * (fab.x + fba.x), 0.5 * (fab.y - fba.y)};
        cd im = cd{0.5 * (fab.y + fba.y), -0.5 * (fab.x - fba.x)};
        if (!green) {
            // Re row [1, 0, -1/z^2, 0], Im row [0, -1/z, 0, 1/z^3]   (:227-230)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const double kr = fma(xk[a][2], c.y, xk[a][0]);
                const double ki = fma(xk[a][3], c.z, xk[a][1] * c.x);
                s[a] = cadd(s[a], cd{fma(ki, im.x, kr * re.x), fma(ki, im.y, kr * re.y)});
            }
        } else {
            // Re row [-1/z^2, 0], Im row [0, 1/z^3]   (:366-367)
            if (al == be) im.x += c.w;  // (:373)
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const double kr = xk[a][0] * c.y;
                const double ki = xk[a][1] * c.z;
                s[a] = cadd(s[a], cd{fma(ki, im.x, kr * re.x), fma(ki, im.y, kr * re.y)});
            }
        }
    }
    for (int a = 0; a < nm; ++a) mt[((size_t)a * range + w) * m2 + e] = s[a];
}

// trace[a][w] = sum_alpha Re mt[a][w][alpha alpha]     (:244-247)
__global__ void k_mom_trace(const cd* __restrict__ mt, int d, int range, int nm, double* __restrict__ trace)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= range * nm) return;
    const int a = t / range, w = t % range;
    double s = 0.0;
    for (int al = 0; al < d; ++al) s += mt[((size_t)a * range + w) * d * d + al + al * d].x;
    trace[t] = s;
}

// avg[a][c][e] = mean_{iw=c..c+2Nv} mt[a][iw][e]     (:259-283)
__global__ void k_mom_avg(const cd* __restrict__ mt, int m2, int range, int nm, int Nv, cd* __restrict__ avg)
{
    const int ncent = range - 2 * Nv;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nm * ncent * m2) return;
    const int e = (int)(t % m2);
    const int c = (int)((t / m2) % ncent);
    const int a = (int)(t / ((size_t)m2 * ncent));
    cd s = cd{0, 0};
    for (int iw = c; iw <= c + 2 * Nv; ++iw) s = cadd(s, mt[((size_t)a * range + iw) * m2 + e]);
    const double cnt = 2 * Nv + 1;
    avg[t] = cd{s.x / cnt, s.y / cnt};
}

// positivity matrix per window centre:
//   non-G: T = Vi m3 Vi - (Vi m2 Vi)^2, Vi = (m1)^{-1/2}   (:298-308)
//   G    : T = m1 - m0^2                                   (:423-424)
// (w1, U1) = eig(m1avg) for the non-G branch.  scratch: 3 d*d per centre
// (Vi, P2 = Vi m2 Vi, tmp).  One thread per (centre, entry); the non-G branch
// runs as five launches (stage 0: Vi, 1: tmp = Vi m2, 2: P2 = tmp Vi,
// 3: tmp = Vi m3, 4: T = tmp Vi - P2 P2), stage -1 is the G branch.  A centre
// whose m1 is not positive gets T = -1 (the window is rejected anyway).
__global__ void k_mom_posmat(const cd* __restrict__ avg, int d, int ncent, int stage, const double* __restrict__ w1,
                             const cd* __restrict__ U1, cd* __restrict__ T, cd* __restrict__ scr)
{
    const int m2 = d * d;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)ncent * m2) return;
    const int c = (int)(t / m2), e = (int)(t % m2), i = e % d, j = e / d;
    auto A = [&](int a) { return avg + ((size_t)a * ncent + c) * m2; };
    cd* out = T + (size_t)c * m2;
    if (stage < 0) {
        const cd* m0 = A(0);
        cd s = A(1)[e];
        for (int l = 0; l < d; ++l) s = cfms(m0[i + l * d], m0[l + j * d], s);
        out[e] = s;
        return;
    }
    const double* w = w1 + (size_t)c * d;
    if (!(w[0] > 0.0)) {
        if (stage == 4) out[e] = cd{-1.0, 0.0};
        return;
    }
    cd* Vi = scr + (size_t)c * 3 * m2;
    cd* P2 = Vi + m2;
    cd* tmp = P2 + m2;
    cd s = cd{0, 0};
    switch (stage) {
    case 0: {
        const cd* U = U1 + (size_t)c * m2;
        for (int l = 0; l < d; ++l) s = cadd(s, cscale(cmulc(U[i + l * d], U[j + l * d]), 1.0 / sqrt(w[l])));
        Vi[e] = s;
        break;
    }
    case 1: {
        const cd* m2a = A(2);
        for (int l = 0; l < d; ++l) s = cfma(Vi[i + l * d], m2a[l + j * d], s);
        tmp[e] = s;
        break;
    }
    case 2:
        for (int l = 0; l < d; ++l) s = cfma(tmp[i + l * d], Vi[l + j * d], s);
        P2[e] = s;
        break;
    case 3: {
        const cd* m3a = A(3);
        for (int l = 0; l < d; ++l) s = cfma(Vi[i + l * d], m3a[l + j * d], s);
        tmp[e] = s;
        break;
    }
    default:
        for (int l = 0; l < d; ++l) s = cfma(tmp[i + l * d], Vi[l + j * d], s);
        for (int l = 0; l < d; ++l) s = cfms(P2[i + l * d], P2[l + j * d], s);
        out[e] = s;
        break;
    }
}

// threads of the window-selection workgroup (one workgroup per moment, the
// window centres spread over its threads)
constexpr int kSelThreads = 1024;

// Window choice of getAsymto_moments (:289-315) for moment a = blockIdx.x:
// the spread of the moment's trace over the 2 Nv + 1 windows around each
// centre (mean, then the squared deviations, in the reference's order), and
// the scan "minPos = ncent-1; for i: if ok[i] && var[i] < var[minPos] then
// minPos = i", i.e. the first admissible centre of least spread if that
// spread is below the last centre's, else the last centre.  ok[i] is the
// positivity test on the window-averaged moments (:301-312, :432).
// cen[target] = minPos + Nv, target = a (a + 2 for the G branch).
__global__ __launch_bounds__(kSelThreads) void k_mom_select(const double* __restrict__ trace, const double* __restrict__ w1,
                                                    const double* __restrict__ wT, int d, int range, int ncent, int Nv,
                                                    int green, int* __restrict__ cen)
{
    __shared__ double bv[kSelThreads];
    __shared__ int bi[kSelThreads];
    const int a = blockIdx.x, t = threadIdx.x;
    auto spread = [&](int i) {
#pragma clang fp contract(off)
        const double* v = trace + (size_t)a * range + i;  // centre i + Nv, window [i, i + 2 Nv]
        double mean = 0.0;
        for (int j = 0; j < 2 * Nv + 1; ++j) mean += v[j];
        mean /= (2 * Nv + 1);
        double ss = 0.0;
        for (int j = 0; j < 2 * Nv + 1; ++j) {
            const double dv = v[j] - mean;
            ss += dv * dv;
        }
        return sqrt(ss / (2 * Nv + 1));
    };
    double best = 0.0;
    int besti = -1;
    for (int i = t; i < ncent; i += kSelThreads) {
        const bool ok = green ? wT[(size_t)i * d] > 1 : (w1[(size_t)i * d] > 0 && wT[(size_t)i * d] > 0);
        if (!ok) continue;
        const double v = spread(i);
        if (besti < 0 || v < best) {  // ascending i within a thread: the first of equals stays
            best = v;
            besti = i;
        }
    }
    bv[t] = best;
    bi[t] = besti;
    __syncthreads();
    for (int w = kSelThreads / 2; w > 0; w >>= 1) {
        if (t < w) {
            const int oi = bi[t + w];
            const double ov = bv[t + w];
            if (oi >= 0 && (bi[t] < 0 || ov < bv[t] || (ov == bv[t] && oi < bi[t]))) {
                bv[t] = ov;
                bi[t] = oi;
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        const int last = ncent - 1;
        const int minPos = (bi[0] >= 0 && bv[0] < spread(last)) ? bi[0] : last;
        cen[green ? a + 2 : a] = minPos + Nv;
    }
}

// moments[4][d][d] = 0 and windows = -1, then m1 = I on the G branch (:451-452)
__global__ void k_mom_init(cd* __restrict__ mom, int d, int green, int* __restrict__ cen)
{
    const int m2 = d * d;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 4) cen[t] = -1;
    if (t >= 4 * m2) return;
    const int q = t / m2, e = t % m2;
    mom[t] = cd{(green && q == 1 && e % d == e / d) ? 1.0 : 0.0, 0.0};
}

// final moment of the chosen window: mean over the window of (mt_ab + conj(mt_ba))/2,
// then Hermitised (:319-330, :458-461); blockIdx.y = moment index
__global__ void k_mom_final(const cd* __restrict__ mt, int d, int range, int green, const int* __restrict__ cen_all,
                            int Nv, cd* __restrict__ mom)
{
    const int m2 = d * d;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m2) return;
    const int mom_idx = blockIdx.y;
    const int target = green ? mom_idx + 2 : mom_idx;
    const int cen = cen_all[target];
    cd* out = mom + (size_t)target * m2;
    const int al = e % d, be = e / d;
    cd s = cd{0, 0}, st = cd{0, 0};
    for (int w = cen - Nv; w <= cen + Nv; ++w) {
        const cd* m = mt + ((size_t)mom_idx * range + w) * m2;
        s = cadd(s, cscale(cadd(m[al + be * d], cconj(m[be + al * d])), 0.5));
        st = cadd(st, cscale(cadd(m[be + al * d], cconj(m[al + be * d])), 0.5));
    }
    const double cnt = 2 * Nv + 1;
    s = cd{s.x / cnt, s.y / cnt};
    st = cd{st.x / cnt, st.y / cnt};
    out[e] = cscale(cadd(s, cconj(st)), 0.5);
}

// ---------------------------------------------------------------------------
// iw -> tau
// ---------------------------------------------------------------------------
// D_i = F_i - tail_i,  tail_i = 1/2 sum_h A_h diag(1/(i w_i - e_h)) A_h^H   (:55-61)
__global__ void k_ft_subtract_tail(const cd* __restrict__ fw, int d, int n_freq, double beta,
                                   const cd* __restrict__ A1, const double* __restrict__ e1,
                                   const cd* __restrict__ A2, const double* __restrict__ e2, cd* __restrict__ D)
{
    const int m2 = d * d;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n_freq * m2) return;
    const int i = (int)(t / m2), e = (int)(t % m2), a = e % d, b = e / d;
    const double w = (kPi / beta) + i * ((2 * kPi) / beta);
    cd tail = cd{0, 0};
    for (int l = 0; l < d; ++l) {
        const cd g1 = crcp(cd{-e1[l], w});
        const cd g2 = crcp(cd{-e2[l], w});
        tail = cfma(cmulc(A1[a + l * d], A1[b + l * d]), g1, tail);
        tail = cfma(cmulc(A2[a + l * d], A2[b + l * d]), g2, tail);
    }
    D[t] = csub(fw[t], cscale(tail, 0.5));
}

// out[j][e] = sum_i coef(j,i) in[i][e]  (NIN = 1)  or  sum_i ca in0 + cb in1 (NIN = 2)
// workgroup: TJ consecutive rows j, all elements; inputs staged in LDS TI rows at a time.
// Phase recurrence of the matrix-core transform (k_dft_mfma, iw -> tau): a lane
// walks its row j through i, i + 4, i + 8, ...; the phase factor at i + 4 is the
// one at i times a per-row constant, and every 16th step it is formed again
// exactly from the FP64 phase, so the rounding of the products stays at ~1e-15.
struct PhaseWalk {
    cd p;   // phase factor at the current i
    cd m4;  // factor from i to i + 4
};

struct CoefW2Tau {
    static constexpr bool PHASE_ROW = false;
    static constexpr bool WALK = true;  // coefficients by the phase recurrence in k_dft_mfma
    double beta, w0, dw, tstep;
    const double* taus;  // explicit tau values (FourierTransform_tau), or null for tau_j = j*tstep
    __device__ void operator()(int j, int i, cd* c) const
    {
        const double w = w0 + i * dw;
        const double tau = taus ? taus[j] : j * tstep;
        double s, co;
        sincos(w * tau, &s, &co);
        c[0] = cd{co, -s};  // exp(-i w tau)
    }
    __device__ void walk_start(int j, int i, PhaseWalk& st) const
    {
        const double tau = taus ? taus[j] : j * tstep;
        double s, co;
        sincos((w0 + i * dw) * tau, &s, &co);
        st.p = cd{co, -s};
        sincos(4.0 * dw * tau, &s, &co);
        st.m4 = cd{co, -s};
    }
    __device__ void walk_coef(const PhaseWalk& st, int, int, cd* c) const { c[0] = st.p; }
};

struct CoefTau2W {
    // the phases exp(i w_n tau_t) of a tile's rows are formed once per (n, t)
    // (k_dft stages them in LDS); segment t takes those at t and t + 1
    static constexpr bool PHASE_ROW = true;
    double beta, dt;
    __device__ cd phase(int n, int t) const
    {
        const double wn = (2 * n + 1) * kPi / beta;
        const double tau = t * dt;
        double s, c;
        sincos(tau * wn, &s, &c);
        return cd{c, s};
    }
    // The coefficients take differences of neighbouring phase factors, which cancel
    // at low frequency, so every phase is formed exactly (no recurrence): rounding
    // in them would be amplified by 1/(w dt).
    static constexpr bool WALK = false;
    __device__ void operator()(int n, int t, cd* c) const { from_phases(n, t, phase(n, t), phase(n, t + 1), c); }
    __device__ void walk_start(int, int, PhaseWalk&) const {}
    __device__ void walk_coef(const PhaseWalk&, int n, int t, cd* c) const { (*this)(n, t, c); }
    __device__ void from_phases(int n, int t, cd e0, cd e1, cd* c) const
    {
        const double wn = (2 * n + 1) * kPi / beta;
        const double tau0 = t * dt, tau1 = (t + 1) * dt;
        const double s0 = e0.y, c0 = e0.x, s1 = e1.y, c1 = e1.x;
        const cd pI = cd{c1 - c0, s1 - s0};
        const cd pI2 = cd{tau1 * c1 - tau0 * c0, tau1 * s1 - tau0 * s0};
        // ca = pI/w^2 + pI2/(i w) ; cb = pI/(i w);  1/(i w) = -i/w
        const double r = 1.0 / wn;
        const cd a1 = cscale(pI, r * r);
        const cd a2 = cd{pI2.y * r, -pI2.x * r};
        c[0] = cadd(a1, a2);
        c[1] = cd{pI.y * r, -pI.x * r};
    }
};

// out[j][e] = sum_i coef(j, i) in[i][e] over this workgroup's share of i:
// blockIdx.y selects the i range [y*ichunk, min(ni, (y+1)*ichunk)); with more
// than one range the results are partial sums, out + y*nj*m2, which
// k_dft_sum_parts adds in range order.
template <class Coef, int NIN, int OPT>
__global__ __launch_bounds__(256) void k_dft(int nj, int ni, int m2, int TJ, int TI, int ichunk,
                                             const cd* __restrict__ in0, const cd* __restrict__ in1,
                                             cd* __restrict__ out, Coef coef)
{
    extern __shared__ cd sm[];
    cd* sc = sm;                        // [NIN][TJ][TI]
    cd* sx = sm + NIN * TJ * TI;        // [NIN][TI][m2]
    cd* sp = sx + NIN * TI * m2;        // PHASE_ROW: [TJ][TI + 1]
    const int j0 = blockIdx.x * TJ;
    const int ib = blockIdx.y * ichunk;
    const int ie = min(ni, ib + ichunk);
    out += (size_t)blockIdx.y * nj * m2;
    const int nout = TJ * m2;
    cd acc[OPT];
#pragma unroll
    for (int q = 0; q < OPT; ++q) acc[q] = cd{0, 0};
    for (int i0 = ib; i0 < ie; i0 += TI) {
        __syncthreads();
        if constexpr (Coef::PHASE_ROW) {
            for (int t = threadIdx.x; t < TJ * (TI + 1); t += 256) {
                const int jj = t / (TI + 1), ii = t % (TI + 1);
                sp[t] = (j0 + jj < nj && i0 + ii <= ie) ? coef.phase(j0 + jj, i0 + ii) : cd{0, 0};
            }
            __syncthreads();
        }
        for (int t = threadIdx.x; t < TJ * TI; t += 256) {
            const int jj = t / TI, ii = t % TI;
            cd c[2] = {cd{0, 0}, cd{0, 0}};
            if (j0 + jj < nj && i0 + ii < ie) {
                if constexpr (Coef::PHASE_ROW)
                    coef.from_phases(j0 + jj, i0 + ii, sp[jj * (TI + 1) + ii], sp[jj * (TI + 1) + ii + 1], c);
                else
                    coef(j0 + jj, i0 + ii, c);
            }
            for (int q = 0; q < NIN; ++q) sc[(q * TJ + jj) * TI + ii] = c[q];
        }
        for (int t = threadIdx.x; t < TI * m2; t += 256) {
            const int ii = t / m2, e = t % m2;
            const bool ok = i0 + ii < ie;
            sx[ii * m2 + e] = ok ? in0[(size_t)(i0 + ii) * m2 + e] : cd{0, 0};
            if (NIN == 2) sx[(TI + ii) * m2 + e] = ok ? in1[(size_t)(i0 + ii) * m2 + e] : cd{0, 0};
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < OPT; ++q) {
            const int o = threadIdx.x + q * 256;
            if (o >= nout) break;
            const int jj = o / m2, e = o % m2;
            cd s = acc[q];
            for (int ii = 0; ii < TI; ++ii) {
                s = cfma(sc[jj * TI + ii], sx[ii * m2 + e], s);
                if (NIN == 2) s = cfma(sc[(TJ + jj) * TI + ii], sx[(TI + ii) * m2 + e], s);
            }
            acc[q] = s;
        }
    }
#pragma unroll
    for (int q = 0; q < OPT; ++q) {
        const int o = threadIdx.x + q * 256;
        if (o >= nout) break;
        const int jj = o / m2, e = o % m2;
        if (j0 + jj < nj) out[(size_t)(j0 + jj) * m2 + e] = acc[q];
    }
}

// G(tau_j) = Herm[ (X + X^H)/beta + 1/2 sum_h A_h diag(-1/(e^{e tau} + e^{e (tau-beta)})) A_h^H ]   (:62-90)
__global__ void k_ft_finish(const cd* __restrict__ X, int d, int ntau1, double beta, double tstep,
                            const double* __restrict__ taus, const cd* __restrict__ A1,
                            const double* __restrict__ e1, const cd* __restrict__ A2, const double* __restrict__ e2,
                            cd* __restrict__ out)
{
    const int m2 = d * d;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)ntau1 * m2) return;
    const int j = (int)(t / m2), e = (int)(t % m2), a = e % d, b = e / d;
    const double tau = taus ? taus[j] : j * tstep;
    const cd* Xj = X + (size_t)j * m2;
    auto elem = [&](int p, int q) {
        cd v = cadd(Xj[p + q * d], cconj(Xj[q + p * d]));
        v = cscale(v, 1.0 / beta);
        cd s = cd{0, 0};
        for (int l = 0; l < d; ++l) {
            const double g1 = -1.0 / (exp(e1[l] * tau) + exp(e1[l] * (tau - beta)));
            const double g2 = -1.0 / (exp(e2[l] * tau) + exp(e2[l] * (tau - beta)));
            s = cadd(s, cscale(cmulc(A1[p + l * d], A1[q + l * d]), g1));
            s = cadd(s, cscale(cmulc(A2[p + l * d], A2[q + l * d]), g2));
        }
        return cadd(v, cscale(s, 0.5));
    };
    const cd gab = elem(a, b), gba = elem(b, a);
    out[t] = cscale(cadd(gab, cconj(gba)), 0.5);
}

// FT_t_to_w segment slopes (:531-536)
__global__ void k_ft_segments(const cd* __restrict__ gt, int m2, int nseg, double dt, cd* __restrict__ at,
                              cd* __restrict__ bt)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nseg * m2) return;
    const int s = (int)(t / m2);
    const double tau0 = s * dt, tau1 = (s + 1) * dt;
    const cd a = cscale(csub(gt[t + m2], gt[t]), 1.0 / (tau1 - tau0));
    at[t] = a;
    bt[t] = cadd(cscale(cneg(a), tau0), gt[t]);
}

// (F - F^H)/(2i), for the causality test of FT_t_to_w (:552)
__global__ void k_antiherm(const cd* __restrict__ f, int d, int count, cd* __restrict__ out)
{
    const int m2 = d * d;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)count * m2) return;
    const int e = (int)(t % m2), a = e % d, b = e / d;
    const size_t base = t - e;
    const cd x = f[t], y = f[base + b + a * d];
    const cd dlt = cd{x.x - y.x, x.y + y.y};  // x - conj(y)
    out[t] = cd{0.5 * dlt.y, -0.5 * dlt.x};  // dlt / (2i)
}

// Register-tiled form for m2 <= 256 (used for tau -> iw, two inputs per i):
// thread (jg, e), jg < jgroups, keeps output
// element e of the rj consecutive rows jg*rj .. jg*rj + rj - 1, so one read of
// in[i][e] from LDS serves rj complex FMAs and the coefficient reads are the
// same address across the lanes of a row group (broadcast).  Tiles, i ranges
// and the order of the sum over i are those of k_dft.
template <class Coef, int NIN, int RJ>
__global__ __launch_bounds__(256) void k_dft_rt(int nj, int ni, int m2, int TJ, int TI, int ichunk, int jgroups,
                                                int rj, const cd* __restrict__ in0, const cd* __restrict__ in1,
                                                cd* __restrict__ out, Coef coef)
{
    extern __shared__ cd sm[];
    cd* sc = sm;                        // [NIN][TJ][TI]
    cd* sx = sm + NIN * TJ * TI;        // [NIN][TI][m2]
    cd* sp = sx + NIN * TI * m2;        // PHASE_ROW: [TJ][TI + 1]
    const int j0 = blockIdx.x * TJ;
    const int ib = blockIdx.y * ichunk;
    const int ie = min(ni, ib + ichunk);
    out += (size_t)blockIdx.y * nj * m2;
    const int e = threadIdx.x % m2, jg = threadIdx.x / m2;
    const bool act = jg < jgroups;
    cd acc[RJ];
#pragma unroll
    for (int r = 0; r < RJ; ++r) acc[r] = cd{0, 0};
    for (int i0 = ib; i0 < ie; i0 += TI) {
        __syncthreads();
        if constexpr (Coef::PHASE_ROW) {
            for (int t = threadIdx.x; t < TJ * (TI + 1); t += 256) {
                const int jj = t / (TI + 1), ii = t % (TI + 1);
                sp[t] = (j0 + jj < nj && i0 + ii <= ie) ? coef.phase(j0 + jj, i0 + ii) : cd{0, 0};
            }
            __syncthreads();
        }
        for (int t = threadIdx.x; t < TJ * TI; t += 256) {
            const int jj = t / TI, ii = t % TI;
            cd c[2] = {cd{0, 0}, cd{0, 0}};
            if (j0 + jj < nj && i0 + ii < ie) {
                if constexpr (Coef::PHASE_ROW)
                    coef.from_phases(j0 + jj, i0 + ii, sp[jj * (TI + 1) + ii], sp[jj * (TI + 1) + ii + 1], c);
                else
                    coef(j0 + jj, i0 + ii, c);
            }
            for (int q = 0; q < NIN; ++q) sc[(q * TJ + jj) * TI + ii] = c[q];
        }
        for (int t = threadIdx.x; t < TI * m2; t += 256) {
            const int ii = t / m2, ee = t % m2;
            const bool ok = i0 + ii < ie;
            sx[ii * m2 + ee] = ok ? in0[(size_t)(i0 + ii) * m2 + ee] : cd{0, 0};
            if (NIN == 2) sx[(TI + ii) * m2 + ee] = ok ? in1[(size_t)(i0 + ii) * m2 + ee] : cd{0, 0};
        }
        __syncthreads();
        if (act) {
            const int jr = jg * rj;
            for (int ii = 0; ii < TI; ++ii) {
                const cd x0 = sx[ii * m2 + e];
                const cd x1 = NIN == 2 ? sx[(TI + ii) * m2 + e] : cd{0, 0};
#pragma unroll
                for (int r = 0; r < RJ; ++r) {
                    if (r < rj) {
                        acc[r] = cfma(sc[(jr + r) * TI + ii], x0, acc[r]);
                        if (NIN == 2) acc[r] = cfma(sc[(TJ + jr + r) * TI + ii], x1, acc[r]);
                    }
                }
            }
        }
    }
    if (!act) return;
#pragma unroll
    for (int r = 0; r < RJ; ++r) {
        const int jj = jg * rj + r;
        if (r < rj && j0 + jj < nj) out[(size_t)(j0 + jj) * m2 + e] = acc[r];
    }
}

// parts[y][j][e] -> out[j][e], y ascending
__global__ void k_dft_sum_parts(const cd* __restrict__ parts, int nparts, size_t count, cd* __restrict__ out)
{
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    cd s = parts[t];
    for (int y = 1; y < nparts; ++y) s = cadd(s, parts[(size_t)y * count + t]);
    out[t] = s;
}

// The same transform on the matrix cores (FP64 MFMA), for m2 <= 16 * NE:
// out[j][e] = sum_i sum_q coef_q(j, i) in_q[i][e] as complex products split
// into real v_mfma_f64_16x16x4 (D = A B + D, A 16 x 4, B 4 x 16).  A wavefront
// owns 16 consecutive rows j and every e tile (NE tiles of 16 elements), so
// each coefficient it forms serves NE tiles.  Per step of 4 summed indices,
// lane l = r + 16 k forms coef(j0 + r, i0 + k), the A operand's (r, k) entry
// (by the phase recurrence of PhaseWalk),
// and loads in[i0 + k][16 t + r], the B operand's (k, r) entry of tile t; the
// accumulator of tile t holds out[j0 + c][16 t + r'] at lane r' + 16 (c % 4),
// register c / 4.  blockIdx.y selects the range of i, as in k_dft.
typedef double dmfma4 __attribute__((ext_vector_type(4)));

template <class Coef, int NIN, int NE>
__global__ __launch_bounds__(256) void k_dft_mfma(int nj, int ni, int m2, int ichunk, const cd* __restrict__ in0,
                                                  const cd* __restrict__ in1, cd* __restrict__ out, Coef coef)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, k = lane >> 4;
    const int j0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    const int ib = blockIdx.y * ichunk;
    const int ie = min(ni, ib + ichunk);
    out += (size_t)blockIdx.y * nj * m2;
    dmfma4 accR[NE], accI[NE];
#pragma unroll
    for (int t = 0; t < NE; ++t) {
        accR[t] = dmfma4{0.0, 0.0, 0.0, 0.0};
        accI[t] = dmfma4{0.0, 0.0, 0.0, 0.0};
    }
    const int j = j0 + r;
    const bool jok = j < nj;
    PhaseWalk st{};
    for (int i0 = ib, step = 0; i0 < ie; i0 += 4, ++step) {
        const int i = i0 + k;
        if constexpr (Coef::WALK) {
            if ((step & 15) == 0) coef.walk_start(jok ? j : 0, i, st);
            else st.p = cmul(st.p, st.m4);
        }
        cd c[2] = {cd{0, 0}, cd{0, 0}};
        if (i < ie && jok) coef.walk_coef(st, j, i, c);
#pragma unroll
        for (int q = 0; q < NIN; ++q) {
            const cd* in = q == 0 ? in0 : in1;
            cd b[NE];
#pragma unroll
            for (int t = 0; t < NE; ++t) {
                const int e = 16 * t + r;
                b[t] = (i < ie && e < m2) ? in[(size_t)i * m2 + e] : cd{0, 0};
            }
            // four passes over the tiles: consecutive MFMAs update different accumulators
            const double ar = c[q].x, ai = c[q].y;
#pragma unroll
            for (int t = 0; t < NE; ++t) accR[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, b[t].x, accR[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NE; ++t) accI[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, b[t].y, accI[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NE; ++t) accR[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, b[t].y, accR[t], 0, 0, 0);
#pragma unroll
            for (int t = 0; t < NE; ++t) accI[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, b[t].x, accI[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < NE; ++t) {
        const int e = 16 * t + r;
        if (e >= m2) continue;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int jj = j0 + k + 4 * g;
            if (jj < nj) out[(size_t)jj * m2 + e] = cd{accR[t][g], accI[t][g]};
        }
    }
}

struct DftShape {
    int TJ, TI, ks, ichunk, jgroups, rj;  // jgroups > 0: the register-tiled kernel
};

// Tiles of the transform and the split of the summed index i over
// workgroups: enough workgroups to fill the device (about 1024) when the
// output alone gives fewer.
// The matrix-core form (k_dft_mfma) for m2 <= 128: rows in tiles of 64 per
// workgroup (16 per wavefront), the summed index split so that the launch has
// about 4096 wavefronts (4 per SIMD), in multiples of 4.
constexpr int kDftMfmaMaxM2 = 128;

inline DftShape dft_shape_mfma(int nj, int ni)
{
    DftShape d{};
    d.TJ = 64;
    d.TI = 4;
    const int gx = (nj + d.TJ - 1) / d.TJ;
    const int steps = (ni + 3) / 4;
    int ks = (1024 + gx - 1) / gx;
    if (ks > steps) ks = steps;
    if (ks > 32) ks = 32;
    if (ks < 1) ks = 1;
    d.ichunk = (steps + ks - 1) / ks * 4;
    d.ks = (ni + d.ichunk - 1) / d.ichunk;
    d.jgroups = -1;
    return d;
}

inline DftShape dft_shape(int nj, int ni, int m2, int nin, bool phase_row)
{
    if (m2 <= kDftMfmaMaxM2) return dft_shape_mfma(nj, ni);
    DftShape d;
    if (m2 <= 256 && nin == 2) {  // register-tiled (tau -> iw; measured slower for iw -> tau): TJ = jgroups * rj rows
        d.jgroups = 256 / m2;
        d.rj = 256 / d.jgroups;
        if (d.rj > 8) d.rj = 8;
        d.TJ = d.jgroups * d.rj;
        d.TI = 3072 / (nin * (d.TJ + m2) + (phase_row ? d.TJ + 1 : 0));
        if (d.TI > 32) d.TI = 32;
        if (d.TI < 1) d.TI = 1;
    } else {
        d.jgroups = 0;
        d.rj = 0;
        d.TJ = 2048 / m2;
        if (d.TJ > 32) d.TJ = 32;
        if (d.TJ < 1) d.TJ = 1;
        d.TI = 2048 / (m2 * nin);
        if (d.TI > 32) d.TI = 32;
        if (d.TI < 1) d.TI = 1;
    }
    const int gx = (nj + d.TJ - 1) / d.TJ;
    const int itiles = (ni + d.TI - 1) / d.TI;
    int ks = (1024 + gx - 1) / gx;
    if (ks > itiles) ks = itiles;
    if (ks > 16) ks = 16;
    if (ks < 1) ks = 1;
    const int per = (itiles + ks - 1) / ks;
    d.ichunk = per * d.TI;
    d.ks = (ni + d.ichunk - 1) / d.ichunk;
    return d;
}

// scratch (in cd) the split transform needs for its partial sums
inline size_t dft_part_count(int nj, int ni, int m2, int nin, bool phase_row)
{
    const DftShape d = dft_shape(nj, ni, m2, nin, phase_row);
    return d.ks > 1 ? (size_t)d.ks * nj * m2 : 0;
}

template <class Coef, int NIN>
hipError_t launch_dft(int nj, int ni, int m2, const cd* in0, const cd* in1, cd* out, Coef coef, cd* parts,
                      hipStream_t s)
{
    const DftShape d = dft_shape(nj, ni, m2, NIN, Coef::PHASE_ROW);
    if (d.jgroups < 0) {  // matrix cores
        if (d.ks > 1 && !parts) return hipErrorInvalidValue;
        cd* dstm = d.ks > 1 ? parts : out;
        dim3 gridm((nj + d.TJ - 1) / d.TJ, d.ks);
#define M_(NE)                                                                                                    \
    if (m2 <= 16 * NE) {                                                                                          \
        hipLaunchKernelGGL((k_dft_mfma<Coef, NIN, NE>), gridm, dim3(256), 0, s, nj, ni, m2, d.ichunk, in0, in1, dstm, \
                           coef);                                                                                 \
    } else
        M_(1) M_(2) M_(4) M_(6) M_(7) M_(8) return hipErrorInvalidValue;
#undef M_
        hipError_t e = hipGetLastError();
        if (e != hipSuccess || d.ks == 1) return e;
        const size_t cnt = (size_t)nj * m2;
        hipLaunchKernelGGL(k_dft_sum_parts, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, parts, d.ks, cnt, out);
        return hipGetLastError();
    }
    const int TJ = d.TJ, TI = d.TI;
    const int opt = (TJ * m2 + 255) / 256;
    size_t shm = sizeof(cd) * ((size_t)NIN * TJ * TI + (size_t)NIN * TI * m2 +
                               (Coef::PHASE_ROW ? (size_t)TJ * (TI + 1) : 0));
    dim3 grid((nj + TJ - 1) / TJ, d.ks);
    cd* dst = d.ks > 1 ? parts : out;
    if (d.ks > 1 && !parts) return hipErrorInvalidValue;
    if (d.jgroups > 0) {
        hipError_t e = hipFuncSetAttribute((const void*)k_dft_rt<Coef, NIN, 8>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_dft_rt<Coef, NIN, 8>), grid, dim3(256), shm, s, nj, ni, m2, TJ, TI, d.ichunk, d.jgroups,
                           d.rj, in0, in1, dst, coef);
        e = hipGetLastError();
        if (e != hipSuccess || d.ks == 1) return e;
        const size_t cnt = (size_t)nj * m2;
        hipLaunchKernelGGL(k_dft_sum_parts, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, parts, d.ks, cnt, out);
        return hipGetLastError();
    }
#define L_(O)                                                                                                     \
    {                                                                                                             \
        hipError_t e = hipFuncSetAttribute((const void*)k_dft<Coef, NIN, O>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                           (int)shm);                                                             \
        if (e != hipSuccess) return e;                                                                            \
        hipLaunchKernelGGL((k_dft<Coef, NIN, O>), grid, dim3(256), shm, s, nj, ni, m2, TJ, TI, d.ichunk, in0, in1,  \
                           dst, coef);                                                                            \
        e = hipGetLastError();                                                                                    \
        if (e != hipSuccess || d.ks == 1) return e;                                                               \
        const size_t cnt = (size_t)nj * m2;                                                                       \
        hipLaunchKernelGGL(k_dft_sum_parts, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, parts, d.ks, cnt, out); \
        return hipGetLastError();                                                                                 \
    }
    if (opt <= 1) L_(1)
    if (opt <= 2) L_(2)
    if (opt <= 4) L_(4)
    L_(8)
#undef L_
}

unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

#define HIPCHK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(h, DMFTGPU_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

static char* ft_ws(dmftgpu_handle h, size_t bytes, int* rc)
{
    hipError_t e;
    char* p = (char*)h->ft.get(bytes, &e);
    if (!p) *rc = fail(h, DMFTGPU_ERR_HIP, "ft workspace %zu bytes: %s", bytes, hipGetErrorString(e));
    return p;
}

struct Carve {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count)
    {
        T* p = (T*)(base + off);
        off += (count * sizeof(T) + 255) / 256 * 256;
        return p;
    }
};

extern "C" int dmftgpu_moments(dmftgpu_handle h, int d, int n_freq, double beta, const double* fw, int green_asym,
                               double* moments, int* windows)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (d < 1 || d > 32) return fail(h, DMFTGPU_ERR_UNSUPPORTED, "d=%d outside 1..32", d);
    if (!fw || !moments) return fail(h, DMFTGPU_ERR_ARG, "fw and moments are required");
    if (!(beta > 0.0)) return fail(h, DMFTGPU_ERR_ARG, "beta=%g", beta);
    const int nm = green_asym ? 2 : 4;
    const int range = (green_asym ? n_freq - 2 : n_freq - 4) + 1;
    int Nv = n_freq / 20;
    if (Nv == 0) Nv = 1;
    const int ncent = range - 2 * Nv;
    // the reference asserts on the window bounds (:321, :440)
    if (range < 1 || ncent < 1) return fail(h, DMFTGPU_ERR_ARG, "n_freq=%d too small for the window search", n_freq);
    const int m2 = d * d;
    // the basis table and the inverted normal matrices depend on (n_freq, beta,
    // green_asym) only: kept in the handle and formed again when those change
    dmftgpu_handle_s::MomEntry* ent = nullptr;
    for (auto& m : h->mom_cache)
        if (m.valid && m.nf == n_freq && m.beta == beta && m.green == green_asym) ent = &m;
    const bool hit = ent != nullptr;
    if (!hit) {  // the least recently used entry (an unused one first)
        ent = &h->mom_cache[0];
        for (auto& m : h->mom_cache)
            if (!m.valid || m.used < ent->used) {
                ent = &m;
                if (!m.valid) break;
            }
    }
    ent->used = ++h->mom_clock;
    hipError_t eg;
    char* kbase = (char*)ent->buf.get((size_t)n_freq * sizeof(double4) + 256 + (size_t)range * 16 * 8, &eg);
    if (!kbase) return fail(h, DMFTGPU_ERR_HIP, "moment cache: %s", hipGetErrorString(eg));
    double4* tab = (double4*)kbase;
    double* kinv = (double*)(kbase + ((size_t)n_freq * sizeof(double4) + 255) / 256 * 256);
    size_t need = (size_t)nm * range * m2 * 16 + (size_t)nm * range * 8 +
                  (size_t)nm * ncent * m2 * 16 + (size_t)ncent * m2 * 16 * 6 + (size_t)ncent * d * 8 * 2 + 64 * 256;
    int rc = 0;
    char* base = ft_ws(h, need, &rc);
    if (!base) return rc;
    Carve cv{base};
    cd* mt = cv.take<cd>((size_t)nm * range * m2);
    double* trace = cv.take<double>((size_t)nm * range);
    cd* avg = cv.take<cd>((size_t)nm * ncent * m2);
    double* w1 = cv.take<double>((size_t)ncent * d);
    cd* U1 = cv.take<cd>((size_t)ncent * m2);
    cd* T = cv.take<cd>((size_t)ncent * m2);
    cd* scr = cv.take<cd>((size_t)ncent * m2 * 3);
    double* wT = cv.take<double>((size_t)ncent * d);
    int* cen = cv.take<int>(4);
    hipStream_t s = h->stream;
    if (!hit) {
        ent->valid = false;
        std::vector<double4> htab(n_freq);
        for (int wn = 0; wn < n_freq; ++wn) {
            const double z = (2 * wn + 1) * kPi / beta;  // (:224, :364)
            htab[wn] = double4{-1. / std::pow(z, 1), -1. / std::pow(z, 2), 1. / std::pow(z, 3), 1.0 / z};
        }
        HIPCHK(hipMemcpyAsync(tab, htab.data(), sizeof(double4) * n_freq, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));  // htab is a host temporary
        const int npair = green_asym ? 3 : 10;
        hipLaunchKernelGGL(k_mom_ktk, dim3(nblk((size_t)range * npair)), dim3(256), 0, s, n_freq, range, green_asym, tab,
                           kinv);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_mom_kinv, dim3(nblk(range)), dim3(256), 0, s, range, green_asym, kinv);
        HIPCHK(hipGetLastError());
        ent->valid = true;
        ent->nf = n_freq;
        ent->beta = beta;
        ent->green = green_asym;
    }
    hipLaunchKernelGGL(k_mom_fit, dim3(nblk((size_t)range * m2)), dim3(256), 0, s, (const cd*)fw, d, n_freq, range, tab,
                       green_asym, kinv, mt);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_mom_trace, dim3(nblk((size_t)range * nm)), dim3(256), 0, s, mt, d, range, nm, trace);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_mom_avg, dim3(nblk((size_t)nm * ncent * m2)), dim3(256), 0, s, mt, m2, range, nm, Nv, avg);
    HIPCHK(hipGetLastError());
    if (!green_asym) HIPCHK(dg::launch_herm_eig(avg + (size_t)1 * ncent * m2, d, 0, d, nullptr, 0, w1, U1, ncent, s));
    for (int stage = green_asym ? -1 : 0; stage <= (green_asym ? -1 : 4); ++stage) {
        hipLaunchKernelGGL(k_mom_posmat, dim3(nblk((size_t)ncent * m2)), dim3(256), 0, s, avg, d, ncent, stage, w1, U1,
                           T, scr);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(dg::launch_herm_eig(T, d, 0, d, nullptr, 0, wT, nullptr, ncent, s));
    // window choice and final moments on the device: no host round trip
    cd* dm = (cd*)moments;
    hipLaunchKernelGGL(k_mom_init, dim3(nblk((size_t)4 * m2)), dim3(256), 0, s, dm, d, green_asym, cen);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_mom_select, dim3(nm), dim3(kSelThreads), 0, s, trace, w1, wT, d, range, ncent, Nv, green_asym, cen);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(k_mom_final, dim3(nblk(m2), nm), dim3(256), 0, s, mt, d, range, green_asym, cen, Nv, dm);
    HIPCHK(hipGetLastError());
    if (windows) {  // the only wait of the call
        int win[4];
        HIPCHK(hipMemcpyAsync(win, cen, sizeof(win), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::memcpy(windows, win, sizeof(win));
    }
    return DMFTGPU_OK;
}

// Two-pole tail set-up of FourierTransform_tau (:22-44): returns V^H U_h and eig values e_h.
static void tail_setup(int d, const std::vector<hl::cx>& m1, const std::vector<hl::cx>& m2, const std::vector<hl::cx>& m3,
                       std::vector<hl::cx>& A1, std::vector<double>& e1, std::vector<hl::cx>& A2, std::vector<double>& e2)
{
    using namespace hl;
    Mat V = zeros(d), H1 = zeros(d), H2 = zeros(d);
    if (min_eig(d, m1) > 0) {
        V = herm_func(d, m1, [](double x) { return std::sqrt(x > 0 ? x : 0.0); });
        Mat Vi, ViH;
        inverse(d, V, Vi);
        inverse(d, adj(d, V), ViH);
        Mat M2 = mul(d, mul(d, ViH, lin(m2, 2.0, m2, 0.0)), Vi);
        Mat M3 = mul(d, mul(d, ViH, lin(m3, 2.0, m3, 0.0)), Vi);
        Mat X = lin(M3, 0.5, mul(d, M2, M2), -0.25);
        if (min_eig(d, X) > 0) {
            Mat sq = herm_func(d, X, [](double x) { return std::sqrt(x > 0 ? x : 0.0); });
            H1 = lin(M2, 0.5, sq, 1.0);
        } else {
            H1 = lin(M2, 0.5, M2, 0.0);
        }
        H2 = lin(M2, 1.0, H1, -1.0);
    }
    Mat U;
    Mat VH = adj(d, V);
    eigh(d, H1, e1, U);
    A1 = mul(d, VH, U);
    eigh(d, H2, e2, U);
    A2 = mul(d, VH, U);
}

extern "C" int dmftgpu_ft_w2tau(dmftgpu_handle h, int d, int n_freq, int n_tau, double beta, const double* fw,
                                const double* moments, double* gtau, int* patched)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (d < 1 || d > 32) return fail(h, DMFTGPU_ERR_UNSUPPORTED, "d=%d outside 1..32", d);
    if (!fw || !moments || !gtau) return fail(h, DMFTGPU_ERR_ARG, "fw, moments and gtau are required");
    if (n_freq < 1 || n_tau < 1 || !(beta > 0.0)) return fail(h, DMFTGPU_ERR_ARG, "n_freq=%d n_tau=%d beta=%g", n_freq, n_tau, beta);
    const int m2 = d * d, ntau1 = n_tau + 1;
    hipStream_t s = h->stream;
    std::vector<hl::cx> hm(4 * (size_t)m2);
    HIPCHK(hipMemcpyAsync(hm.data(), moments, sizeof(cd) * 4 * m2, hipMemcpyDefault, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<hl::cx> m1(hm.begin() + m2, hm.begin() + 2 * m2), mm2(hm.begin() + 2 * m2, hm.begin() + 3 * m2),
        m3(hm.begin() + 3 * m2, hm.begin() + 4 * m2);
    std::vector<hl::cx> A1, A2;
    std::vector<double> e1, e2;
    tail_setup(d, m1, mm2, m3, A1, e1, A2, e2);
    const size_t npart = dft_part_count(ntau1, n_freq, m2, 1, CoefW2Tau::PHASE_ROW);
    size_t need = (size_t)n_freq * m2 * 16 + (size_t)ntau1 * m2 * 16 + 4 * (size_t)m2 * 16 + 4 * (size_t)d * 8 +
                  (size_t)ntau1 * d * 8 + npart * 16 + 16 * 256;
    int rc = 0;
    char* base = ft_ws(h, need, &rc);
    if (!base) return rc;
    Carve cv{base};
    cd* D = cv.take<cd>((size_t)n_freq * m2);
    cd* X = cv.take<cd>((size_t)ntau1 * m2);
    cd* parts = npart ? cv.take<cd>(npart) : nullptr;
    cd* dA1 = cv.take<cd>(m2);
    cd* dA2 = cv.take<cd>(m2);
    double* de1 = cv.take<double>(d);
    double* de2 = cv.take<double>(d);
    double* wmin = cv.take<double>((size_t)ntau1 * d);
    HIPCHK(hipMemcpyAsync(dA1, A1.data(), sizeof(cd) * m2, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dA2, A2.data(), sizeof(cd) * m2, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(de1, e1.data(), sizeof(double) * d, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(de2, e2.data(), sizeof(double) * d, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_ft_subtract_tail, dim3(nblk((size_t)n_freq * m2)), dim3(256), 0, s, (const cd*)fw, d, n_freq,
                       beta, dA1, de1, dA2, de2, D);
    HIPCHK(hipGetLastError());
    CoefW2Tau cw{beta, kPi / beta, (2 * kPi) / beta, beta / n_tau, nullptr};
    HIPCHK((launch_dft<CoefW2Tau, 1>(ntau1, n_freq, m2, D, nullptr, X, cw, parts, s)));
    cd* out = (cd*)gtau;
    hipLaunchKernelGGL(k_ft_finish, dim3(nblk((size_t)ntau1 * m2)), dim3(256), 0, s, X, d, ntau1, beta, beta / n_tau,
                       (const double*)nullptr, dA1, de1, dA2, de2, out);
    HIPCHK(hipGetLastError());
    // causality patch (:145-175): min eigenvalue of every G(tau_j)
    HIPCHK(dg::launch_herm_eig(out, d, 0, d, nullptr, 0, wmin, nullptr, ntau1, s));
    std::vector<double> hw((size_t)ntau1 * d);
    HIPCHK(hipMemcpyAsync(hw.data(), wmin, sizeof(double) * hw.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    bool any = false;
    for (int t = 0; t <= n_tau; ++t) any |= hw[(size_t)t * d] >= 0;
    int check = 0;
    if (any) {  // rare path: restate the reference loop on the host
        std::vector<hl::cx> g((size_t)ntau1 * m2);
        HIPCHK(hipMemcpyAsync(g.data(), out, sizeof(cd) * g.size(), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        double mine11 = 0.0;
        for (int t = 0; t < ntau1; ++t) {
            const double tau = t * beta / n_tau;
            const double mine = hw[(size_t)t * d];
            mine11 = mine;
            if (mine >= 0 && (t != 0 && t != n_tau)) {
                check = 1;
                for (int t1 = t + 1; t1 < ntau1; ++t1) {
                    const double tau1 = t1 * beta / n_tau, taup = (t - 1) * beta / n_tau;
                    for (int e = 0; e < m2; ++e)
                        g[(size_t)t * m2 + e] = (g[(size_t)t1 * m2 + e] - g[(size_t)(t - 1) * m2 + e]) / (tau1 - taup) *
                                                    (tau - taup) + g[(size_t)(t - 1) * m2 + e];
                    hl::Mat gm(g.begin() + (size_t)t * m2, g.begin() + (size_t)(t + 1) * m2);
                    mine11 = hl::min_eig(d, gm);
                    if (mine11 < 0) break;
                }
            }
            if (mine11 >= 0) {
                check = 1;
                for (int e = 0; e < m2; ++e) g[(size_t)t * m2 + e] = 0.0;
                for (int l = 0; l < d; ++l) g[(size_t)t * m2 + l + (size_t)l * d] = -1e-10;
            }
        }
        HIPCHK(hipMemcpyAsync(out, g.data(), sizeof(cd) * g.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    if (patched) *patched = check;
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_ft_tau2w(dmftgpu_handle h, int d, int n_freq, int n_tau, double beta, const double* gtau,
                                double* fw, int* violations)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (d < 1 || d > 32) return fail(h, DMFTGPU_ERR_UNSUPPORTED, "d=%d outside 1..32", d);
    if (!gtau || !fw) return fail(h, DMFTGPU_ERR_ARG, "gtau and fw are required");
    if (n_freq < 1 || n_tau < 1 || !(beta > 0.0)) return fail(h, DMFTGPU_ERR_ARG, "n_freq=%d n_tau=%d beta=%g", n_freq, n_tau, beta);
    const int m2 = d * d;
    const double dt = beta / n_tau;
    hipStream_t s = h->stream;
    const size_t npart = dft_part_count(n_freq, n_tau, m2, 2, CoefTau2W::PHASE_ROW);
    size_t need = 2 * (size_t)n_tau * m2 * 16 + (size_t)n_freq * m2 * 16 + (size_t)n_freq * d * 8 + npart * 16 + 8 * 256;
    int rc = 0;
    char* base = ft_ws(h, need, &rc);
    if (!base) return rc;
    Carve cv{base};
    cd* at = cv.take<cd>((size_t)n_tau * m2);
    cd* bt = cv.take<cd>((size_t)n_tau * m2);
    cd* ah = cv.take<cd>((size_t)n_freq * m2);
    double* wmax = cv.take<double>((size_t)n_freq * d);
    cd* parts = npart ? cv.take<cd>(npart) : nullptr;
    hipLaunchKernelGGL(k_ft_segments, dim3(nblk((size_t)n_tau * m2)), dim3(256), 0, s, (const cd*)gtau, m2, n_tau, dt,
                       at, bt);
    HIPCHK(hipGetLastError());
    CoefTau2W ct{beta, dt};
    HIPCHK((launch_dft<CoefTau2W, 2>(n_freq, n_tau, m2, at, bt, (cd*)fw, ct, parts, s)));
    hipLaunchKernelGGL(k_antiherm, dim3(nblk((size_t)n_freq * m2)), dim3(256), 0, s, (const cd*)fw, d, n_freq, ah);
    HIPCHK(hipGetLastError());
    HIPCHK(dg::launch_herm_eig(ah, d, 0, d, nullptr, 0, wmax, nullptr, n_freq, s));
    std::vector<double> hw((size_t)n_freq * d);
    HIPCHK(hipMemcpyAsync(hw.data(), wmax, sizeof(double) * hw.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int viol = 0;
    for (int n = 0; n < n_freq; ++n) viol += hw[(size_t)n * d + d - 1] > 0;
    if (violations) *violations = viol;
    return DMFTGPU_OK;
}


// FourierTransform_tau (FourierTransform.cpp:16-118) at arbitrary tau values
// (no causality patch: the single-tau overloads of FourierTransform have none).
extern "C" int dmftgpu_ft_w2tau_points(dmftgpu_handle h, int d, int n_freq, double beta, const double* fw,
                                       const double* moments, int n_points, const double* taus_host, double* out)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (d < 1 || d > 32) return fail(h, DMFTGPU_ERR_UNSUPPORTED, "d=%d outside 1..32", d);
    if (!fw || !moments || !out || !taus_host || n_points < 1 || n_freq < 1 || !(beta > 0.0))
        return fail(h, DMFTGPU_ERR_ARG, "bad arguments to ft_w2tau_points");
    const int m2 = d * d;
    hipStream_t s = h->stream;
    std::vector<hl::cx> hm(4 * (size_t)m2);
    HIPCHK(hipMemcpyAsync(hm.data(), moments, sizeof(cd) * 4 * m2, hipMemcpyDefault, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<hl::cx> m1(hm.begin() + m2, hm.begin() + 2 * m2), mm2(hm.begin() + 2 * m2, hm.begin() + 3 * m2),
        m3(hm.begin() + 3 * m2, hm.begin() + 4 * m2);
    std::vector<hl::cx> A1, A2;
    std::vector<double> e1, e2;
    tail_setup(d, m1, mm2, m3, A1, e1, A2, e2);
    const size_t npart = dft_part_count(n_points, n_freq, m2, 1, CoefW2Tau::PHASE_ROW);
    size_t need = (size_t)n_freq * m2 * 16 + (size_t)n_points * m2 * 16 + 4 * (size_t)m2 * 16 + 4 * (size_t)d * 8 +
                  (size_t)n_points * 8 + npart * 16 + 16 * 256;
    int rc = 0;
    char* base = ft_ws(h, need, &rc);
    if (!base) return rc;
    Carve cv{base};
    cd* D = cv.take<cd>((size_t)n_freq * m2);
    cd* X = cv.take<cd>((size_t)n_points * m2);
    cd* parts = npart ? cv.take<cd>(npart) : nullptr;
    cd* dA1 = cv.take<cd>(m2);
    cd* dA2 = cv.take<cd>(m2);
    double* de1 = cv.take<double>(d);
    double* de2 = cv.take<double>(d);
    double* dtau = cv.take<double>(n_points);
    HIPCHK(hipMemcpyAsync(dA1, A1.data(), sizeof(cd) * m2, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dA2, A2.data(), sizeof(cd) * m2, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(de1, e1.data(), sizeof(double) * d, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(de2, e2.data(), sizeof(double) * d, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dtau, taus_host, sizeof(double) * n_points, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_ft_subtract_tail, dim3(nblk((size_t)n_freq * m2)), dim3(256), 0, s, (const cd*)fw, d, n_freq,
                       beta, dA1, de1, dA2, de2, D);
    HIPCHK(hipGetLastError());
    CoefW2Tau cw{beta, kPi / beta, (2 * kPi) / beta, 0.0, dtau};
    HIPCHK((launch_dft<CoefW2Tau, 1>(n_points, n_freq, m2, D, nullptr, X, cw, parts, s)));
    hipLaunchKernelGGL(k_ft_finish, dim3(nblk((size_t)n_points * m2)), dim3(256), 0, s, X, d, n_points, beta, 0.0,
                       (const double*)dtau, dA1, de1, dA2, de2, (cd*)out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return DMFTGPU_OK;
}
