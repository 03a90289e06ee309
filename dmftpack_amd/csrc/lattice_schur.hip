// Schur-complement lattice path, host side: dispatch of the one-matrix-per-lane
// kernels (lattice_schur_lane.hpp, instantiated per n_c in lattice_lane{1..6}.hip)
// and of the column-per-lane kernels (lattice_schur_cols.hpp), the Sigma table,
// and the per-k preparation (W = H_cd V_d, the per-k blocks).
#include "lattice_kernels.hpp"

namespace dg {

// -------------------------------------------------------------------------
// n_c <= 6: one matrix per lane (this file) or, on request, column per lane;
// 7..16: column per lane (lattice_schur_cols.hpp, buckets 4/5/6/8/10/12/16)
bool schur_supported(int n, int nc) { return nc >= 1 && nc <= 16 && n >= nc && n <= 32; }
bool schur_multi_supported(int nc) { return nc >= 1 && nc <= 6; }
static int cols_bucket(int nc)
{
    return nc <= 4 ? 4 : nc <= 6 ? nc : nc <= 8 ? 8 : nc <= 10 ? 10 : nc <= 12 ? 12 : 16;
}
int schur_slots(int nc, int mode, bool fp32, bool cols)
{
    if (nc <= 6 && !cols) return 64;
    (void)mode;
    (void)fp32;
    return 64 / cols_bucket(nc);  // one wavefront of 64/N slots per workgroup
}

static hipError_t dispatch_nc(const LatArgs& a, int mode, bool fp32, const cd* sig_all, hipStream_t s, int wg,
                              const SchurMulti* m)
{
    switch (a.nc) {
        case 1: return launch_lane_1(a, mode, fp32, sig_all, s, wg, m);
        case 2: return launch_lane_2(a, mode, fp32, sig_all, s, wg, m);
        case 3: return launch_lane_3(a, mode, fp32, sig_all, s, wg, m);
        case 4: return launch_lane_4(a, mode, fp32, sig_all, s, wg, m);
        case 5: return launch_lane_5(a, mode, fp32, sig_all, s, wg, m);
        case 6: return launch_lane_6(a, mode, fp32, sig_all, s, wg, m);
        default: return hipErrorInvalidValue;
    }
}

// Sigma at every frequency of the density sum (moment tail beyond n_freq)
__global__ void k_sigma_all(const cd* __restrict__ sigma, const cd* __restrict__ mom, int nc, int n_freq, int nw_tot,
                            double beta, cd* __restrict__ out)
{
    const int m2 = nc * nc;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nw_tot * m2) return;
    const int nw = t / m2, e = t % m2;
    out[t] = sigma_elem(sigma, mom, nc, n_freq, beta, nw, e % nc, e / nc);
}

hipError_t launch_sigma_all(const LatArgs& a, hipStream_t s)
{
    const int m2 = a.nc * a.nc;
    hipLaunchKernelGGL(k_sigma_all, dim3((a.nw_tot * m2 + 255) / 256), dim3(256), 0, s, a.sigma, a.mom, a.nc, a.n_freq,
                       a.nw_tot, a.beta, a.pSigAll);
    return hipGetLastError();
}

bool schur_can_fuse(int nc, bool cols, bool real_axis) { return nc >= 1 && nc <= 6 && !cols && !real_axis; }

hipError_t launch_lattice_schur(const LatArgs& a, int mode, bool fp32, hipStream_t s, int wg, bool cols, bool fused)
{
    if (fused && !schur_can_fuse(a.nc, cols, a.real_axis)) return hipErrorInvalidValue;
    // wg = 0: the adjoint call at 128 frequencies per workgroup (its one wavefront
    // per SIMD gains from a second sharing each per-k scalar fill), the FP32R
    // forward over the first n_freq frequencies at n_c >= 5 as two wavefronts
    // splitting the k walk over 64 frequencies (its LDS accumulators, n_c^2 x 1 KB
    // per 64 frequencies, then allow the 8 wavefronts per CU that its registers
    // allow, against 6 at n_c = 5 and 4 at n_c = 6), the other launches at 64
    const int wg_call = wg ? wg : mode == MODE_VJP ? 128 : 64;
    const int wg_lo = wg ? wg : mode == MODE_FWD && fp32 && a.nc >= 5 ? SCHUR_WG_KPAIR : wg_call;
    hipError_t e = launch_sigma_all(a, s);
    if (e != hipSuccess) return e;
    if (a.real_axis && mode == MODE_FWD && a.nc <= 6 && !cols)
        return dispatch_nc(a, MODE_FWD_REAL, fp32, a.pSigAll, s, wg_call, nullptr);
    if (a.nc > 6 || cols || a.real_axis) {  // the per-lane kernels take the real axis in the forward only
        switch (cols_bucket(a.nc)) {
            case 4: return launch_cols_4(a, mode, fp32, a.pSigAll, s);
            case 5: return launch_cols_5(a, mode, fp32, a.pSigAll, s);
            case 6: return launch_cols_6(a, mode, fp32, a.pSigAll, s);
            case 8: return launch_cols_8(a, mode, fp32, a.pSigAll, s);
            case 10: return launch_cols_10(a, mode, fp32, a.pSigAll, s);
            case 12: return launch_cols_12(a, mode, fp32, a.pSigAll, s);
            default: return launch_cols_16(a, mode, fp32, a.pSigAll, s);
        }
    }
    if (mode == MODE_FWD && a.nw_tot > a.n_freq) {
        // Frequencies beyond n_freq enter NumMatrix (and N) only: the Hermitian
        // part of G_cc is accumulated there, half the accumulators.
        LatArgs lo = a, hi = a;
        lo.nw_end = a.n_freq;
        hi.nw_begin = a.n_freq;
        e = dispatch_nc(lo, MODE_FWD, fp32, a.pSigAll, s, wg_lo, nullptr);
        if (e != hipSuccess) return e;
        // fused: the adjoint that follows needs tr G^2 there too (a.pTr2)
        return dispatch_nc(hi, fused ? MODE_DENS_HERM : MODE_FWD_HERM, fp32, a.pSigAll, s, wg_call, nullptr);
    }
    if (mode == MODE_VJP && a.nw_tot > a.n_freq) {
        // Frequencies beyond n_freq contribute only tr G and tr G^2 (dN/dmu):
        // the density kernel does that at twice the occupancy of the adjoint.
        LatArgs lo = a, hi = a;
        lo.nw_end = a.n_freq;
        hi.nw_begin = a.n_freq;
        e = dispatch_nc(lo, MODE_VJP, fp32, a.pSigAll, s, wg_call, nullptr);
        if (e != hipSuccess || fused) return e;  // fused: the forward wrote tr G^2 beyond n_freq
        return dispatch_nc(hi, MODE_DENS, fp32, a.pSigAll, s, wg_call, nullptr);
    }
    return dispatch_nc(a, mode, fp32, a.pSigAll, s, wg_lo, nullptr);
}

hipError_t launch_lattice_schur_multi(const LatArgs& proto, int mode, bool fp32, hipStream_t s, int wg,
                                      const SchurMulti& m)
{
    return dispatch_nc(proto, mode, fp32, nullptr, s, wg, &m);
}

hipError_t launch_lattice_schur_cols_multi(const LatArgs& proto, int mode, bool fp32, hipStream_t s,
                                           const SchurMulti& m)
{
    switch (cols_bucket(proto.nc)) {
        case 8: return launch_cols_multi_8(proto, mode, fp32, s, m);
        case 10: return launch_cols_multi_10(proto, mode, fp32, s, m);
        case 12: return launch_cols_multi_12(proto, mode, fp32, s, m);
        case 16: return launch_cols_multi_16(proto, mode, fp32, s, m);
        default: return hipErrorInvalidValue;
    }
}

// W(k) = H_cd(k) V_d(k): W_ad = sum_e H(a, nc+e) V_d(e, d)
__global__ void k_schur_prep(const cd* __restrict__ Hk, const cd* __restrict__ Vd, int n, int nc, int nk,
                             cd* __restrict__ W)
{
    const int nd = n - nc;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nk * nc * nd) return;
    const int k = (int)(t / ((size_t)nc * nd));
    const int r = (int)(t % ((size_t)nc * nd));
    const int aa = r % nc, d = r / nc;
    const cd* H = Hk + (size_t)k * n * n;
    const cd* V = Vd + (size_t)k * nd * nd;
    cd s = cd{0.0, 0.0};
    for (int e = 0; e < nd; ++e) s = cfma(H[aa + (size_t)(nc + e) * n], V[e + (size_t)d * nd], s);
    W[t] = s;
}

// Per-k block of the one-matrix-per-lane kernels, laid out as the LDS
// staging buffer of stage_k: [H_cc nc*nc][P_d(a<=b) nd*np][e_d nd doubles],
// P_d(a, b) = W_ad conj(W_bd), pairs ordered b-major.
__global__ void k_schur_kdata(const cd* __restrict__ Hk, const cd* __restrict__ W, const double* __restrict__ ed, int n,
                              int nc, int nk, int bs, cd* __restrict__ out)
{
    const int nd = n - nc, np = nc * (nc + 1) / 2;
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (size_t)nk * bs) return;
    const int k = (int)(gid / bs), t = (int)(gid % bs);
    cd v = cd{0.0, 0.0};
    if (t < nc * nc) {
        v = Hk[(size_t)k * n * n + (t % nc) + (size_t)(t / nc) * n];
    } else if (t < nc * nc + nd * np) {
        const int q = t - nc * nc, d = q / np, p = q % np;
        int ib = 0;
        while ((ib + 1) * (ib + 2) / 2 <= p) ++ib;
        const int ia = p - ib * (ib + 1) / 2;
        const cd* Wk = W + (size_t)k * nc * nd;
        v = cmulc(Wk[ia + d * nc], Wk[ib + d * nc]);
    } else {
        const int q = 2 * (t - nc * nc - nd * np);  // two e_d per cd slot
        const double* e = ed + (size_t)k * nd;
        v.x = q < nd ? e[q] : 0.0;
        v.y = q + 1 < nd ? e[q + 1] : 0.0;
    }
    out[gid] = v;
}

int schur_kdata_size(int n, int nc) { return nc * nc + (n - nc) * (nc * (nc + 1) / 2) + (n - nc + 1) / 2; }

hipError_t launch_schur_kdata(const cd* Hk, const cd* W, const double* ed, int n, int nc, int nk, cd* out,
                              hipStream_t s)
{
    const int bs = schur_kdata_size(n, nc);
    const size_t total = (size_t)nk * bs;
    hipLaunchKernelGGL(k_schur_kdata, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, Hk, W, ed, n, nc, nk, bs,
                       out);
    return hipGetLastError();
}

hipError_t launch_schur_prep(const cd* Hk, const cd* Vd, int n, int nc, int nk, cd* W, hipStream_t s)
{
    const size_t total = (size_t)nk * nc * (n - nc);
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_schur_prep, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, Hk, Vd, n, nc, nk, W);
    return hipGetLastError();
}

}  // namespace dg
