// C ABI of the lattice Green's-function library (include/dmftgpu.h): host
// logic around the HIP kernels.  Each entry point restates the control flow of
// the reference routine it replaces; the per-(k, iw) work runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dmftgpu.h"
#include "api_internal.hpp"
#include "lattice_kernels.hpp"

using dg::cd;

const char* dmftgpu_version(void) { return "dmftgpu 0.1 (gfx950)"; }

namespace dgi {

int fail(dmftgpu_handle h, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    return code;
}

void* Workspace::get(size_t bytes, hipError_t* e)
{
    *e = hipSuccess;
    if (bytes <= cap) return ptr;
    if (ptr) {
        *e = hipFree(ptr);
        if (*e != hipSuccess) return nullptr;
        ptr = nullptr;
        cap = 0;
    }
    *e = hipMalloc(&ptr, bytes);
    if (*e != hipSuccess) {
        ptr = nullptr;
        return nullptr;
    }
    cap = bytes;
    return ptr;
}

void Workspace::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

// Carves consecutive 256-byte aligned regions out of one allocation.
struct Carver {
    size_t off = 0;
    template <class T>
    size_t take(size_t count)
    {
        size_t o = off;
        off += (count * sizeof(T) + 255) / 256 * 256;
        return o;
    }
};

}  // namespace dgi

using namespace dgi;

#define HIPCHK(expr)                                                                           \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(h, DMFTGPU_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

extern "C" int dmftgpu_create(dmftgpu_handle* out, void* stream)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!out) return DMFTGPU_ERR_ARG;
    auto* h = new dmftgpu_handle_s();
    h->stream = (hipStream_t)stream;
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return DMFTGPU_ERR_HIP;
    }
    // the no-alloc scratch of this device (dmftgpu_greenftn_w_noalloc); without
    // it only that entry point is unavailable
    if (dgi::scratch_address(h->device, &h->scratch) != hipSuccess) {
        (void)hipGetLastError();
        h->scratch = nullptr;
    }
    *out = h;
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_destroy(dmftgpu_handle h)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    (void)hipStreamSynchronize(h->stream);
    for (int m = 0; m < 3; ++m) {
        if (h->ev0[m]) (void)hipEventDestroy(h->ev0[m]);
        if (h->ev1[m]) (void)hipEventDestroy(h->ev1[m]);
    }
    h->ft.release();
    for (auto& m : h->mom_cache) m.buf.release();
    h->ws.release();
    h->prep.release();
    h->stat.release();
    h->static_cache.release();
    h->fused.release();
    h->cot.release();
    h->core.release();
    delete h;
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_set_stream(dmftgpu_handle h, void* stream)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    h->stream = (hipStream_t)stream;
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_enable_timing(dmftgpu_handle h, int enable)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (enable)
        for (int m = 0; m < 3; ++m) {
            if (!h->ev0[m]) HIPCHK(hipEventCreate(&h->ev0[m]));
            if (!h->ev1[m]) HIPCHK(hipEventCreate(&h->ev1[m]));
        }
    h->timing = enable != 0;
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_kernel_time(dmftgpu_handle h, int mode, double* ms)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h || !ms || mode < 0 || mode > 2) return DMFTGPU_ERR_ARG;
    if (!h->ev_used[mode]) return fail(h, DMFTGPU_ERR_STATE, "no timed kernel of mode %d", mode);
    HIPCHK(hipEventSynchronize(h->ev1[mode]));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, h->ev0[mode], h->ev1[mode]));
    *ms = t;
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_set_option(dmftgpu_handle h, int option, int value)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    switch (option) {
        case DMFTGPU_OPT_REFINE_STEPS:
            if (value < 1 || value > 2) return fail(h, DMFTGPU_ERR_ARG, "refine steps %d outside 1..2", value);
            h->refine_steps = value;
            return DMFTGPU_OK;
        case DMFTGPU_OPT_TARGET_WORKGROUPS:
            if (value < 0) return fail(h, DMFTGPU_ERR_ARG, "target workgroups %d < 0", value);
            h->target_wg = value;
            return DMFTGPU_OK;
        case DMFTGPU_OPT_SCHUR_WG:
            if (value != 0 && value != 64 && value != 128)
                return fail(h, DMFTGPU_ERR_ARG, "Schur workgroup width %d not 0/64/128", value);
            h->schur_wg = value;
            return DMFTGPU_OK;
        case DMFTGPU_OPT_SCHUR_LAYOUT:
            if (value != 0 && value != 1) return fail(h, DMFTGPU_ERR_ARG, "Schur layout %d not 0/1", value);
            h->schur_layout = value;
            return DMFTGPU_OK;
        case DMFTGPU_OPT_UPFOLD_CHUNK:
            if (value < 0) return fail(h, DMFTGPU_ERR_ARG, "upfolding chunk %d < 0", value);
            h->upfold_chunk = value;
            return DMFTGPU_OK;
        case DMFTGPU_OPT_DENSITY_KERNEL:
            if (value != 0 && value != 1) return fail(h, DMFTGPU_ERR_ARG, "density kernel %d not 0/1", value);
            h->density_kernel = value;
            return DMFTGPU_OK;
        default: return fail(h, DMFTGPU_ERR_ARG, "unknown option %d", option);
    }
}

extern "C" const char* dmftgpu_last_error(dmftgpu_handle h) { return h ? h->err.c_str() : "null handle"; }

static int check_lattice(dmftgpu_handle h, const dmftgpu_lattice* L)
{
    if (!L) return fail(h, DMFTGPU_ERR_ARG, "lattice descriptor is null");
    if (L->n < 1 || L->n > 32) return fail(h, DMFTGPU_ERR_UNSUPPORTED, "n=%d outside 1..32", L->n);
    if (L->n_corr < 1 || L->n_corr > L->n) return fail(h, DMFTGPU_ERR_ARG, "n_corr=%d outside 1..n", L->n_corr);
    if (L->n_cluster < 1 || L->n_corr % L->n_cluster != 0)
        return fail(h, DMFTGPU_ERR_ARG, "n_cluster=%d does not divide n_corr=%d", L->n_cluster, L->n_corr);
    if (L->n_k < 1 || L->n_k_global < 1) return fail(h, DMFTGPU_ERR_ARG, "n_k=%d n_k_global=%d", L->n_k, L->n_k_global);
    if (L->n_freq < 1 || L->high_freq < 1) return fail(h, DMFTGPU_ERR_ARG, "n_freq=%d high_freq=%d", L->n_freq, L->high_freq);
    if (!(L->beta > 0.0) || !std::isfinite(L->beta)) return fail(h, DMFTGPU_ERR_ARG, "beta=%g", L->beta);
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_prepare(dmftgpu_handle h, const dmftgpu_lattice* L, const double* Hk, const double* moments)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    int rc = check_lattice(h, L);
    if (rc) return rc;
    if (!Hk || !moments) return fail(h, DMFTGPU_ERR_ARG, "Hk or moments is null");
    h->lat = *L;
    h->Hk = (const cd*)Hk;
    const int n = L->n, nc = L->n_corr, nk = L->n_k;
    const int nd = n - nc;
    Carver cv;
    size_t o_mom = cv.take<cd>(4 * (size_t)nc * nc);
    size_t o_ev = cv.take<double>((size_t)nk * n);
    size_t o_U = cv.take<cd>((size_t)nk * n * n);
    size_t o_ed = cv.take<double>((size_t)nk * (nd > 0 ? nd : 1));
    size_t o_Vd = cv.take<cd>((size_t)nk * (nd > 0 ? nd * nd : 1));
    size_t o_W = cv.take<cd>((size_t)nk * nc * (nd > 0 ? nd : 1) + dg::kKdataPadBytes / sizeof(cd));
    const bool lane_kernels = dg::schur_supported(n, nc) && dg::schur_multi_supported(nc);
    size_t o_kd = cv.take<cd>(lane_kernels ? (size_t)nk * dg::schur_kdata_size(n, nc) + dg::kKdataPadBytes / sizeof(cd)
                                          : 1);
    hipError_t e;
    char* base = (char*)h->prep.get(cv.off, &e);
    if (!base) return fail(h, DMFTGPU_ERR_HIP, "prepare alloc %zu bytes: %s", cv.off, hipGetErrorString(e));
    h->mom = (cd*)(base + o_mom);
    h->ev = (double*)(base + o_ev);
    h->U = (cd*)(base + o_U);
    h->ed = (double*)(base + o_ed);
    h->Vd = (cd*)(base + o_Vd);
    h->W = (cd*)(base + o_W);
    h->kdata = lane_kernels ? (cd*)(base + o_kd) : nullptr;
    HIPCHK(hipMemcpyAsync(h->mom, moments, sizeof(cd) * 4 * nc * nc, hipMemcpyDefault, h->stream));
    // static part: eig(H(k) + P m0 P^H)   (GreenFtn.cpp:197-215)
    HIPCHK(dg::launch_herm_eig(h->Hk, n, 0, n, h->mom, nc, h->ev, h->U, nk, h->stream));
    // Schur data: eig(H_dd(k)) and W = H_cd V_d
    h->schur_ready = false;
    if (dg::schur_supported(n, nc)) {
        if (nd > 0) {
            HIPCHK(dg::launch_herm_eig(h->Hk, n, nc, nd, nullptr, 0, h->ed, h->Vd, nk, h->stream));
            HIPCHK(dg::launch_schur_prep(h->Hk, h->Vd, n, nc, nk, h->W, h->stream));
        }
        if (lane_kernels) HIPCHK(dg::launch_schur_kdata(h->Hk, h->W, h->ed, n, nc, nk, h->kdata, h->stream));
        h->schur_ready = true;
    }
    h->core_n_orb = 0;  // Kohn-Sham bands belong to the previous H(k)
    h->prepared = true;
    ++h->prep_gen;
    return DMFTGPU_OK;
}

// New Sigma tail moments for the prepared H(k): only the static part
// eig(H(k) + P m0 P^H) (GreenFtn.cpp:197-215) depends on them; the Schur data
// eig(H_dd(k)) and W = H_cd V_d depend on H(k) alone and are kept.
extern "C" int dmftgpu_update_moments(dmftgpu_handle h, const double* moments)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (!h->prepared) return fail(h, DMFTGPU_ERR_STATE, "dmftgpu_prepare has not been called");
    if (!moments) return fail(h, DMFTGPU_ERR_ARG, "moments is null");
    const int n = h->lat.n, nc = h->lat.n_corr, nk = h->lat.n_k;
    HIPCHK(hipMemcpyAsync(h->mom, moments, sizeof(cd) * 4 * nc * nc, hipMemcpyDefault, h->stream));
    HIPCHK(dg::launch_herm_eig(h->Hk, n, 0, n, h->mom, nc, h->ev, h->U, nk, h->stream));
    ++h->prep_gen;
    return DMFTGPU_OK;
}

// Kohn-Sham bands for the core-band term of GiwToNele (GreenFtn.cpp:622-631):
// ks_energy [n_k][n_orb] (host or device), val_to_ks [n_k][n] HOST, the KS
// band of each window band (FromValToKS), -1 for a padded band.  The bands of a
// k point that no window band names are the ones the term counts.
extern "C" int dmftgpu_set_core_bands(dmftgpu_handle h, int n_orb, const double* ks_energy, const int* val_to_ks)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (!h->prepared) return fail(h, DMFTGPU_ERR_STATE, "dmftgpu_prepare has not been called");
    ++h->prep_gen;  // the cached static totals include the core term
    if (n_orb == 0) {
        h->core_n_orb = 0;
        return DMFTGPU_OK;
    }
    const int n = h->lat.n, nk = h->lat.n_k;
    if (n_orb < 0 || n_orb > 4096 || !ks_energy || !val_to_ks)
        return fail(h, DMFTGPU_ERR_ARG, "n_orb=%d (1..4096) needs ks_energy and val_to_ks", n_orb);
    std::vector<unsigned char> out((size_t)nk * n_orb, 1);
    for (int k = 0; k < nk; ++k)
        for (int i = 0; i < n; ++i) {
            const int b = val_to_ks[(size_t)k * n + i];
            if (b == -1) continue;
            if (b < 0 || b >= n_orb)
                return fail(h, DMFTGPU_ERR_ARG, "val_to_ks[%d][%d] = %d outside 0..%d", k, i, b, n_orb - 1);
            if (!out[(size_t)k * n_orb + b])
                return fail(h, DMFTGPU_ERR_ARG, "k point %d names KS band %d twice", k, b);
            out[(size_t)k * n_orb + b] = 0;
        }
    const size_t eb = ((size_t)nk * n_orb * sizeof(double) + 255) / 256 * 256;
    hipError_t e;
    char* base = (char*)h->core.get(eb + out.size(), &e);
    if (!base) return fail(h, DMFTGPU_ERR_HIP, "core bands: %s", hipGetErrorString(e));
    h->core_eps = (double*)base;
    h->core_out = (unsigned char*)(base + eb);
    HIPCHK(hipMemcpyAsync(h->core_eps, ks_energy, sizeof(double) * (size_t)nk * n_orb, hipMemcpyDefault, h->stream));
    HIPCHK(hipMemcpyAsync(h->core_out, out.data(), out.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));  // out is a host temporary
    h->core_n_orb = n_orb;
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_static_eig(dmftgpu_handle h, double* eps, double* U)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (!h->prepared) return fail(h, DMFTGPU_ERR_STATE, "dmftgpu_prepare has not been called");
    const dmftgpu_lattice& L = h->lat;
    if (eps) HIPCHK(hipMemcpyAsync(eps, h->ev, sizeof(double) * L.n_k * L.n, hipMemcpyDefault, h->stream));
    if (U) HIPCHK(hipMemcpyAsync(U, h->U, sizeof(cd) * (size_t)L.n_k * L.n * L.n, hipMemcpyDefault, h->stream));
    return DMFTGPU_OK;
}

namespace {

struct Plan {
    int n_chunks, k_per_chunk;
};

// Work items per lattice launch that set the k chunking: the option's value,
// or by default 32768 for the one-matrix-per-lane Schur kernels (their
// forward is capped by LDS or registers at 6-8 wavefronts per CU, and shorter
// k walks over more workgroups keep the CUs fuller: 24.0 -> 21.4 ms for the
// config-2 forward call), but no fewer than 64 k per chunk (the per-chunk
// partial sums grow with the chunk count: with 1000 real-axis energies 16 k
// per chunk made the call 25% slower), and 8192 for the others.
int lattice_target(dmftgpu_handle h, int method, int nc, bool cols, int nk, int nw_tot, int slots)
{
    if (h->target_wg > 0) return h->target_wg;
    if (method == DMFTGPU_PLAIN_F64 || cols || nc > 6) return 8192;
    const long n_wt = (nw_tot + slots - 1) / slots;
    const long cap = n_wt * std::max(1, nk / 64);
    return (int)std::max(8192L, std::min(32768L, cap));
}

Plan make_plan(int nk, int nw_tot, int slots, int target_wg)
{
    const int n_wt = (nw_tot + slots - 1) / slots;
    int chunks = (target_wg + n_wt - 1) / n_wt;
    if (chunks < 1) chunks = 1;
    if (chunks > nk) chunks = nk;
    int kpc = (nk + chunks - 1) / chunks;
    chunks = (nk + kpc - 1) / kpc;
    return Plan{chunks, kpc};
}

// Static weights and their per-k reduction: fills st_out[m2] (complex) and
// st_scalars[2] = {sum w, sum dN terms} (device), all unnormalised.  With
// Kohn-Sham bands bound by dmftgpu_set_core_bands the scalars include the
// core-band term of GiwToNele (GreenFtn.cpp:622-631) and its mu-derivative.
// The result depends only on (mu, prepared lattice, core bands) and is kept
// per handle, in the handle's workspaces; with from_ws every buffer is carved
// from ws instead and nothing is kept.  wt_out (optional) receives the
// per-level weights w_kl [nk][n].
int run_static(dmftgpu_handle h, double mu, cd* st_out, double* st_scalars, char* ws, size_t& off,
               bool from_ws = false, const double** wt_out = nullptr, size_t cap = SIZE_MAX)
{
    const dmftgpu_lattice& L = h->lat;
    const int n = L.n, nc = L.n_corr, nk = L.n_k, m2 = nc * nc;
    const int nw_tot = L.n_freq * L.high_freq;
    const size_t wbytes = ((size_t)nk * n * sizeof(double) + 255) / 256 * 256;
    hipError_t e;
    char *cache, *wbuf;
    if (from_ws) {
        cache = ws + off;
        off += ((size_t)m2 * sizeof(cd) + 256 + 255) / 256 * 256;
        wbuf = ws + off;
        off += 2 * wbytes;
    } else {
        cache = (char*)h->static_cache.get((size_t)m2 * sizeof(cd) + 256, &e);
        if (!cache) return fail(h, DMFTGPU_ERR_HIP, "static cache: %s", hipGetErrorString(e));
        wbuf = (char*)h->stat.get(2 * wbytes, &e);
        if (!wbuf) return fail(h, DMFTGPU_ERR_HIP, "static weights: %s", hipGetErrorString(e));
    }
    double* wt = (double*)wbuf;
    double* dt = (double*)(wbuf + wbytes);
    if (wt_out) *wt_out = wt;
    if (from_ws || !(h->static_cache_on && h->static_gen == h->prep_gen && h->static_mu == mu)) {
        const int kpc = (nk + 255) / 256;
        const int chunks = (nk + kpc - 1) / kpc;
        const int cchunks = h->core_n_orb > 0 ? chunks : 0;  // core-band partials follow the static ones
        const int tot = chunks + cchunks;
        cd* pNum = (cd*)(ws + off);
        off += ((size_t)chunks * m2 * sizeof(cd) + 255) / 256 * 256;
        double* pN = (double*)(ws + off);
        off += ((size_t)tot * 2 * sizeof(double) + 255) / 256 * 256;
        double* pdN = pN + tot;
        if (off > cap) return fail(h, DMFTGPU_ERR_STATE, "static-part scratch needs %zu of %zu bytes", off, cap);
        HIPCHK(dg::launch_static_terms(h->ev, n, nk, nw_tot, L.beta, mu, wt, dt, h->stream));
        HIPCHK(dg::launch_static_reduce(wt, dt, h->U, n, nc, nk, kpc, chunks, pNum, pN, pdN, h->stream));
        if (cchunks)
            HIPCHK(dg::launch_core_bands(h->core_eps, h->core_out, h->core_n_orb, nk, kpc, cchunks, L.beta, mu,
                                         pN + chunks, pdN + chunks, h->stream));
        HIPCHK(dg::launch_reduce_chunks(pNum, chunks, m2, (cd*)cache, h->stream));
        HIPCHK(dg::launch_reduce_chunks_d(pN, tot, 1, (double*)(cache + (size_t)m2 * sizeof(cd)), h->stream));
        HIPCHK(dg::launch_reduce_chunks_d(pdN, tot, 1, (double*)(cache + (size_t)m2 * sizeof(cd)) + 1, h->stream));
        if (!from_ws) {
            h->static_gen = h->prep_gen;
            h->static_mu = mu;
        }
    }
    HIPCHK(hipMemcpyAsync(st_out, cache, sizeof(cd) * m2, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(st_scalars, cache + (size_t)m2 * sizeof(cd), 2 * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    return DMFTGPU_OK;
}

// workspace run_static takes from ws (from_ws: its cache and weights too)
size_t static_ws_bytes(const dmftgpu_lattice& L, bool from_ws = false)
{
    const int kpc = (L.n_k + 255) / 256;
    const int chunks = (L.n_k + kpc - 1) / kpc;
    size_t b = ((size_t)chunks * L.n_corr * L.n_corr * 16 + 255) / 256 * 256 + ((size_t)2 * chunks * 16 + 255) / 256 * 256 +
               4096;
    if (from_ws)
        b += 2 * (((size_t)L.n_k * L.n * 8 + 255) / 256 * 256) + ((size_t)L.n_corr * L.n_corr * 16 + 256 + 255) / 256 * 256;
    return b;
}

bool valid_method(int m) { return m == DMFTGPU_PLAIN_F64 || m == DMFTGPU_SCHUR_F64 || m == DMFTGPU_SCHUR_F32R; }

bool method_ok(dmftgpu_handle h, int method, int* rc)
{
    *rc = DMFTGPU_OK;
    if (!h->prepared) *rc = fail(h, DMFTGPU_ERR_STATE, "dmftgpu_prepare has not been called");
    else if (!valid_method(method)) *rc = fail(h, DMFTGPU_ERR_ARG, "unknown method %d", method);
    else if (method != DMFTGPU_PLAIN_F64 && !h->schur_ready)
        *rc = fail(h, DMFTGPU_ERR_UNSUPPORTED, "Schur method needs n_corr <= 16 (n=%d, n_corr=%d)", h->lat.n,
                   h->lat.n_corr);
    return *rc == DMFTGPU_OK;
}

// Runs the per-(k, iw) kernel of the chosen method in the given mode and
// reduces the per-chunk partials.  Outputs (device, in ws):
//   Gsum [nw_tot][m2] (FWD), trsum [nw_tot], tr2sum [nw_tot] (DENS/VJP),
//   Sbsum [n_freq][m2] and musum [n_freq] (VJP).
struct Sums {
    cd *Gsum = nullptr, *tr = nullptr, *tr2 = nullptr, *Sb = nullptr, *mu = nullptr;
};

// Real-axis frequency grid z_n = e0 + n de + i eta, n < n_e (retarded_GreenFtn2).
struct RealAxis {
    int n_e;
    double e0, de, eta;
};

// Per-k outputs: one k per chunk, the per-chunk partials are the results
// (g [n_k][nw][nc*nc], tr [n_k][nw]); no reduction over k.
struct PerK {
    cd* g;
    cd* tr;
};

// fused_tr2 (one-matrix-per-lane Schur only): the [chunks][nw_tot] tr G^2 partials
// shared by a forward that fills them beyond n_freq and the adjoint that follows
// (dmftgpu_gloc_loss_grad), which fills the rest and reduces them.
int run_lattice(dmftgpu_handle h, int method, int mode, const double* sigma, double mu, const double* gbar,
                double nbar, char* ws, size_t& off, Sums& S, bool want_trace = false, const RealAxis* ra = nullptr,
                const PerK* pk = nullptr, cd* fused_tr2 = nullptr, int fused_chunks = 0, int max_chunks = 0,
                size_t cap = SIZE_MAX)
{
    const dmftgpu_lattice& L = h->lat;
    const int n = L.n, nc = L.n_corr, nk = L.n_k, m2 = nc * nc;
    const int nw_tot = ra ? ra->n_e : L.n_freq * L.high_freq;
    const int n_freq = ra ? ra->n_e : L.n_freq;
    int slots;
    if (method == DMFTGPU_PLAIN_F64) slots = dg::plain_slots(n, mode);
    else slots = dg::schur_slots(nc, mode, method == DMFTGPU_SCHUR_F32R, h->schur_layout == 1 || (ra && mode != dg::MODE_FWD));
    const bool cols = h->schur_layout == 1 || (ra && mode != dg::MODE_FWD);
    Plan P = pk ? Plan{nk, 1} : make_plan(nk, nw_tot, slots, lattice_target(h, method, nc, cols, nk, nw_tot, slots));
    if (max_chunks > 0 && P.n_chunks > max_chunks) {  // partial sums held to a fixed scratch size
        const int kpc = (nk + max_chunks - 1) / max_chunks;
        P = Plan{(nk + kpc - 1) / kpc, kpc};
    }
    if (fused_tr2 && P.n_chunks != fused_chunks)  // both halves of a fused pair must chunk k alike
        return fail(h, DMFTGPU_ERR_STATE, "fused partials laid out for %d k chunks, this call uses %d", fused_chunks,
                    P.n_chunks);
    dg::LatArgs a{};
    a.Hk = h->Hk;
    a.sigma = (const cd*)sigma;
    a.mom = h->mom;
    a.n = n;
    a.nc = nc;
    a.nk = nk;
    a.n_freq = n_freq;
    a.nw_tot = nw_tot;
    if (ra) {
        a.real_axis = 1;
        a.e0 = ra->e0;
        a.de = ra->de;
        a.eta = ra->eta;
    }
    a.beta = L.beta;
    a.mu = mu;
    a.k_per_chunk = P.k_per_chunk;
    a.n_chunks = P.n_chunks;
    a.gbar = (const cd*)gbar;
    a.nbar = nbar;
    a.ed = h->ed;
    a.W = h->W;
    a.kdata = h->kdata;
    a.refine = h->refine_steps;
    a.want_trace = want_trace ? 1 : 0;
    auto take = [&](size_t bytes) {
        char* p = ws + off;
        off += (bytes + 255) / 256 * 256;
        return p;
    };
    const size_t C = P.n_chunks;
    if (pk) {
        a.pTr = pk->tr;
        a.pG = pk->g;
    } else {
        a.pTr = (cd*)take(C * nw_tot * sizeof(cd));
        if (mode == dg::MODE_FWD) a.pG = (cd*)take(C * nw_tot * m2 * sizeof(cd));
    }
    if (fused_tr2) a.pTr2 = fused_tr2;
    else if (mode != dg::MODE_FWD) a.pTr2 = (cd*)take(C * nw_tot * sizeof(cd));
    if (mode == dg::MODE_VJP) {
        a.pSb = (cd*)take(C * n_freq * m2 * sizeof(cd));
        a.pMu = (cd*)take(C * n_freq * sizeof(cd));
    }
    if (method != DMFTGPU_PLAIN_F64) a.pSigAll = (cd*)take((size_t)nw_tot * m2 * sizeof(cd));
    // the reduced sums: every region is carved before any kernel runs
    if (!pk) {
        S.tr = (cd*)take(nw_tot * sizeof(cd));
        if (mode == dg::MODE_FWD) S.Gsum = (cd*)take((size_t)nw_tot * m2 * sizeof(cd));
        else S.tr2 = (cd*)take(nw_tot * sizeof(cd));
        if (mode == dg::MODE_VJP) {
            S.Sb = (cd*)take((size_t)n_freq * m2 * sizeof(cd));
            S.mu = (cd*)take(n_freq * sizeof(cd));
        }
    }
    if (off > cap) return fail(h, DMFTGPU_ERR_STATE, "lattice scratch needs %zu of %zu bytes", off, cap);
    if (h->timing) HIPCHK(hipEventRecord(h->ev0[mode], h->stream));
    if (method == DMFTGPU_PLAIN_F64) {
        HIPCHK(dg::launch_lattice_plain(a, mode, h->stream, nullptr));
    } else {
        // schur_wg 0: each launch at its default width (launch_lattice_schur)
        HIPCHK(dg::launch_lattice_schur(a, mode, method == DMFTGPU_SCHUR_F32R, h->stream, h->schur_wg,
                                        h->schur_layout == 1, fused_tr2 != nullptr));
    }
    if (h->timing) {
        HIPCHK(hipEventRecord(h->ev1[mode], h->stream));
        h->ev_used[mode] = true;
    }
    if (pk) return DMFTGPU_OK;
    HIPCHK(dg::launch_reduce_chunks(a.pTr, P.n_chunks, nw_tot, S.tr, h->stream));
    if (mode == dg::MODE_FWD) {
        HIPCHK(dg::launch_reduce_chunks(a.pG, P.n_chunks, (size_t)nw_tot * m2, S.Gsum, h->stream));
    } else {
        HIPCHK(dg::launch_reduce_chunks(a.pTr2, P.n_chunks, nw_tot, S.tr2, h->stream));
    }
    if (mode == dg::MODE_VJP) {
        HIPCHK(dg::launch_reduce_chunks(a.pSb, P.n_chunks, (size_t)n_freq * m2, S.Sb, h->stream));
        HIPCHK(dg::launch_reduce_chunks(a.pMu, P.n_chunks, n_freq, S.mu, h->stream));
    }
    return DMFTGPU_OK;
}

size_t lattice_ws_bytes(dmftgpu_handle h, int method, int mode, const RealAxis* ra = nullptr)
{
    const dmftgpu_lattice& L = h->lat;
    const int nc = L.n_corr, m2 = nc * nc;
    const size_t nw_tot = ra ? (size_t)ra->n_e : (size_t)L.n_freq * L.high_freq;
    const size_t n_freq = ra ? (size_t)ra->n_e : (size_t)L.n_freq;
    int slots = method == DMFTGPU_PLAIN_F64 ? dg::plain_slots(L.n, mode)
                                            : dg::schur_slots(nc, mode, method == DMFTGPU_SCHUR_F32R,
                                                              h->schur_layout == 1 || (ra && mode != dg::MODE_FWD));
    Plan P = make_plan(L.n_k, (int)nw_tot, slots,
                       lattice_target(h, method, nc, h->schur_layout == 1 || (ra && mode != dg::MODE_FWD), L.n_k,
                                      (int)nw_tot, slots));
    size_t C = P.n_chunks + 1;  // +1 for the reduced copy
    size_t b = (C + 1) * nw_tot * 16 * (2 + (size_t)m2) + C * n_freq * 16 * (m2 + 1) + 64 * 256;
    return b;
}

}  // namespace

// GreenFtn_w (GreenFtn.cpp:142-422)
namespace {

// ext_ws (optional): caller-provided scratch of ext_cap bytes used instead of
// the handle's workspaces (the static part then keeps no cache); the k
// chunking is reduced until the partial sums fit.
int forward_impl(dmftgpu_handle h, int method, const double* sigma, double mu, double* gloc, double* gloc_hf,
                 double* num_matrix, double* n_total, cd* fused_tr2, int fused_chunks, char* ext_ws = nullptr,
                 size_t ext_cap = 0)
{
    const dmftgpu_lattice& L = h->lat;
    const int nc = L.n_corr, m2 = nc * nc;
    const int nw_tot = L.n_freq * L.high_freq;
    char* ws;
    int max_chunks = 0;
    if (ext_ws) {
        // bytes: fixed part (Sigma table, reduced sums, totals, static part) plus
        // per k chunk the partials of tr G and G_cc, each region 256-aligned
        const size_t fixed = 3 * (((size_t)nw_tot * m2 * 16 + 255) / 256 * 256) + ((size_t)nw_tot * 16 + 255) / 256 * 256 +
                             ((size_t)(m2 + 3) * 16 + 255) / 256 * 256 + ((size_t)m2 * 16 + 255) / 256 * 256 + 256 +
                             static_ws_bytes(L, true);
        const size_t per_chunk = (size_t)nw_tot * 16 * (m2 + 1);
        if (ext_cap < fixed + per_chunk + 512)
            return fail(h, DMFTGPU_ERR_UNSUPPORTED,
                        "lattice needs at least %zu bytes of scratch, the no-alloc maximum is %zu", fixed + per_chunk + 512,
                        ext_cap);
        max_chunks = (int)std::min<size_t>((ext_cap - fixed - 512) / per_chunk, 1 << 20);
        ws = ext_ws;
    } else {
        hipError_t e;
        size_t need = lattice_ws_bytes(h, method, dg::MODE_FWD) + static_ws_bytes(L) + (size_t)(m2 + 8) * 32 + 8192;
        ws = (char*)h->ws.get(need, &e);
        if (!ws) return fail(h, DMFTGPU_ERR_HIP, "workspace %zu bytes: %s", need, hipGetErrorString(e));
    }
    size_t off = 0;
    Sums S;
    // totals and the static part first, then the lattice: each checks the
    // capacity before it launches anything
    cd* tot = (cd*)(ws + off);
    off += ((size_t)(m2 + 3) * sizeof(cd) + 255) / 256 * 256;
    cd* st = (cd*)(ws + off);
    off += ((size_t)m2 * sizeof(cd) + 255) / 256 * 256;
    double* sts = (double*)(ws + off);
    off += 256;
    const size_t cap = ext_ws ? ext_cap : SIZE_MAX;
    int rc = run_static(h, mu, st, sts, ws, off, ext_ws != nullptr, nullptr, cap);
    if (rc) return rc;
    rc = run_lattice(h, method, dg::MODE_FWD, sigma, mu, nullptr, 0.0, ws, off, S, n_total != nullptr, nullptr,
                     nullptr, fused_tr2, fused_chunks, max_chunks, cap);
    if (rc) return rc;
    HIPCHK(dg::launch_freq_totals(S.Gsum, m2, nc, S.tr, nullptr, nw_tot, nullptr, 0, tot, h->stream));
    const double inv_nk = 1.0 / (double)L.n_k_global;
    HIPCHK(dg::launch_gloc_finalize(S.Gsum, L.n_freq, nc, L.n_cluster, inv_nk, (cd*)gloc, h->stream));
    if (gloc_hf)
        HIPCHK(dg::launch_gloc_hf(h->ev, h->U, L.n, nc, L.n_k, L.n_freq, L.beta, mu, inv_nk, L.n_cluster, (cd*)gloc_hf,
                                  h->stream));
    // NumMatrix = [ (1/beta) sum_n sum_k (G_cc + G_cc^H) + sum_k (U w U^H)_cc ] / knum_mpiGlobal
    // (host arrays on the stack: n_corr <= 32)
    cd htot[32 * 32 + 3], hst[32 * 32], num[32 * 32];
    double hsts[2];
    HIPCHK(hipMemcpyAsync(htot, tot, sizeof(cd) * (m2 + 3), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(hst, st, sizeof(cd) * m2, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(hsts, sts, sizeof(double) * 2, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (n_total) {
        *n_total = (htot[m2].x / L.beta + hsts[0]) / (double)L.n_k_global;
        if (!std::isfinite(*n_total)) return fail(h, DMFTGPU_ERR_NUMERIC, "N(mu) is not finite");
    }
    for (int q = 0; q < m2; ++q) {
        num[q].x = (htot[q].x / L.beta + hst[q].x) * inv_nk;
        num[q].y = (htot[q].y / L.beta + hst[q].y) * inv_nk;
        if (!std::isfinite(num[q].x) || !std::isfinite(num[q].y))
            return fail(h, DMFTGPU_ERR_NUMERIC, "NumMatrix is not finite");
    }
    HIPCHK(hipMemcpyAsync(num_matrix, num, sizeof(cd) * m2, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return DMFTGPU_OK;
}

}  // namespace

extern "C" int dmftgpu_greenftn_w(dmftgpu_handle h, int method, const double* sigma, double mu, double* gloc,
                                  double* gloc_hf, double* num_matrix, double* n_total)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (!h->prepared) return fail(h, DMFTGPU_ERR_STATE, "dmftgpu_prepare has not been called");
    if (!valid_method(method)) return fail(h, DMFTGPU_ERR_ARG, "unknown method %d", method);
    if (method != DMFTGPU_PLAIN_F64 && !h->schur_ready)
        return fail(h, DMFTGPU_ERR_UNSUPPORTED, "Schur method needs n_corr <= 16 (n=%d, n_corr=%d)", h->lat.n, h->lat.n_corr);
    if (!sigma || !gloc || !num_matrix) return fail(h, DMFTGPU_ERR_ARG, "sigma, gloc and num_matrix are required");
    if (!std::isfinite(mu)) return fail(h, DMFTGPU_ERR_ARG, "mu is not finite");
    return forward_impl(h, method, sigma, mu, gloc, gloc_hf, num_matrix, n_total, nullptr, 0);
}

// dmftgpu_greenftn_w without device allocation: every scratch buffer is carved
// from this device's fixed scratch array (noalloc.hip), the k chunking reduced
// until the partial sums fit; the static part is recomputed each call (no
// cached copy can live in a shared array).  No hipMalloc / hipFree, no host heap.
extern "C" int dmftgpu_greenftn_w_noalloc(dmftgpu_handle h, int method, const double* sigma, double mu, double* gloc,
                                          double* gloc_hf, double* num_matrix, double* n_total)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    int rc;
    if (!method_ok(h, method, &rc)) return rc;
    if (!sigma || !gloc || !num_matrix) return fail(h, DMFTGPU_ERR_ARG, "sigma, gloc and num_matrix are required");
    if (!std::isfinite(mu)) return fail(h, DMFTGPU_ERR_ARG, "mu is not finite");
    if (!h->scratch) return fail(h, DMFTGPU_ERR_HIP, "no-alloc scratch unavailable on device %d", h->device);
    return forward_impl(h, method, sigma, mu, gloc, gloc_hf, num_matrix, n_total, nullptr, 0, (char*)h->scratch,
                        dgi::kNoAllocScratchBytes);
}

// retarded_GreenFtn2 (GreenFtn.cpp:425-466) summed over k, as qs_spectra
// accumulates it (qs_spectra.cpp:118-131)
extern "C" int dmftgpu_greenftn_real(dmftgpu_handle h, int method, int n_e, double e0, double de, double eta,
                                     const double* sigma_r, double mu, double* gloc_r, double* trace)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (!h->prepared) return fail(h, DMFTGPU_ERR_STATE, "dmftgpu_prepare has not been called");
    if (!valid_method(method)) return fail(h, DMFTGPU_ERR_ARG, "unknown method %d", method);
    if (method != DMFTGPU_PLAIN_F64 && !h->schur_ready)
        return fail(h, DMFTGPU_ERR_UNSUPPORTED, "Schur method needs n_corr <= 16 (n=%d, n_corr=%d)", h->lat.n, h->lat.n_corr);
    if (n_e < 1 || !std::isfinite(e0) || !std::isfinite(de) || !(eta > 0.0) || !std::isfinite(eta) ||
        !std::isfinite(mu))
        return fail(h, DMFTGPU_ERR_ARG, "n_e=%d e0=%g de=%g eta=%g mu=%g (eta must be > 0)", n_e, e0, de, eta, mu);
    if (!sigma_r || !gloc_r) return fail(h, DMFTGPU_ERR_ARG, "sigma_r and gloc_r are required");
    const dmftgpu_lattice& L = h->lat;
    const int m2 = L.n_corr * L.n_corr;
    const RealAxis ra{n_e, e0, de, eta};
    hipError_t e;
    size_t need = lattice_ws_bytes(h, method, dg::MODE_FWD, &ra) + 8192;
    char* ws = (char*)h->ws.get(need, &e);
    if (!ws) return fail(h, DMFTGPU_ERR_HIP, "workspace %zu bytes: %s", need, hipGetErrorString(e));
    size_t off = 0;
    Sums S;
    int rc = run_lattice(h, method, dg::MODE_FWD, sigma_r, mu, nullptr, 0.0, ws, off, S, trace != nullptr, &ra);
    if (rc) return rc;
    const double inv_nk = 1.0 / (double)L.n_k_global;
    HIPCHK(dg::launch_scale(S.Gsum, (size_t)n_e * m2, inv_nk, (cd*)gloc_r, h->stream));
    if (trace) HIPCHK(dg::launch_scale(S.tr, (size_t)n_e, inv_nk, (cd*)trace, h->stream));
    return DMFTGPU_OK;
}

// Per-k real-axis Green's function (qs_spectra.cpp:197-221, the band spectra)
extern "C" int dmftgpu_greenftn_real_k(dmftgpu_handle h, int method, int n_e, double e0, double de, double eta,
                                       const double* sigma_r, double mu, double* gcc_k, double* trace_k)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (!h->prepared) return fail(h, DMFTGPU_ERR_STATE, "dmftgpu_prepare has not been called");
    if (!valid_method(method)) return fail(h, DMFTGPU_ERR_ARG, "unknown method %d", method);
    if (method != DMFTGPU_PLAIN_F64 && !h->schur_ready)
        return fail(h, DMFTGPU_ERR_UNSUPPORTED, "Schur method needs n_corr <= 16 (n=%d, n_corr=%d)", h->lat.n, h->lat.n_corr);
    if (n_e < 1 || !std::isfinite(e0) || !std::isfinite(de) || !(eta > 0.0) || !std::isfinite(eta) ||
        !std::isfinite(mu))
        return fail(h, DMFTGPU_ERR_ARG, "n_e=%d e0=%g de=%g eta=%g mu=%g (eta must be > 0)", n_e, e0, de, eta, mu);
    if (!sigma_r || (!gcc_k && !trace_k)) return fail(h, DMFTGPU_ERR_ARG, "sigma_r and gcc_k or trace_k are required");
    const dmftgpu_lattice& L = h->lat;
    const size_t m2 = (size_t)L.n_corr * L.n_corr;
    const RealAxis ra{n_e, e0, de, eta};
    // the kernels always write both partials: scratch for the one not asked for
    const size_t g_scratch = gcc_k ? 0 : (size_t)L.n_k * n_e * m2 * sizeof(cd);
    const size_t t_scratch = trace_k ? 0 : (size_t)L.n_k * n_e * sizeof(cd);
    hipError_t e;
    size_t need = lattice_ws_bytes(h, method, dg::MODE_FWD, &ra) + g_scratch + t_scratch + 8192;
    char* ws = (char*)h->ws.get(need, &e);
    if (!ws) return fail(h, DMFTGPU_ERR_HIP, "workspace %zu bytes: %s", need, hipGetErrorString(e));
    size_t off = 0;
    PerK pk{gcc_k ? (cd*)gcc_k : (cd*)(ws + off), nullptr};
    off += (g_scratch + 255) / 256 * 256;
    pk.tr = trace_k ? (cd*)trace_k : (cd*)(ws + off);
    off += (t_scratch + 255) / 256 * 256;
    Sums S;
    return run_lattice(h, method, dg::MODE_FWD, sigma_r, mu, nullptr, 0.0, ws, off, S, trace_k != nullptr, &ra, &pk);
}

// N(mu), dN/dmu
static int density_impl(dmftgpu_handle h, int method, const double* sigma, double mu, double* n_out, double* dn_out)
{
    const dmftgpu_lattice& L = h->lat;
    const int nc = L.n_corr, m2 = nc * nc;
    const int nw_tot = L.n_freq * L.high_freq;
    hipError_t e;
    size_t need = lattice_ws_bytes(h, method, dg::MODE_DENS) + static_ws_bytes(L) + (size_t)(m2 + 8) * 32 + 8192;
    char* ws = (char*)h->ws.get(need, &e);
    if (!ws) return fail(h, DMFTGPU_ERR_HIP, "workspace %zu bytes: %s", need, hipGetErrorString(e));
    size_t off = 0;
    Sums S;
    int rc = run_lattice(h, method, dg::MODE_DENS, sigma, mu, nullptr, 0.0, ws, off, S);
    if (rc) return rc;
    cd* tot = (cd*)(ws + off);
    off += ((size_t)(m2 + 3) * sizeof(cd) + 255) / 256 * 256;
    cd* st = (cd*)(ws + off);
    off += ((size_t)m2 * sizeof(cd) + 255) / 256 * 256;
    double* sts = (double*)(ws + off);
    off += 256;
    rc = run_static(h, mu, st, sts, ws, off);
    if (rc) return rc;
    HIPCHK(dg::launch_freq_totals(nullptr, m2, nc, S.tr, S.tr2, nw_tot, nullptr, 0, tot, h->stream));
    std::vector<cd> htot(m2 + 3);
    double hsts[2];
    HIPCHK(hipMemcpyAsync(htot.data(), tot, sizeof(cd) * (m2 + 3), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(hsts, sts, sizeof(double) * 2, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const double inv_nk = 1.0 / (double)L.n_k_global;
    const double N = (htot[m2].x / L.beta + hsts[0]) * inv_nk;
    const double dN = (-htot[m2 + 1].x / L.beta + hsts[1]) * inv_nk;
    if (!std::isfinite(N) || !std::isfinite(dN)) return fail(h, DMFTGPU_ERR_NUMERIC, "N(mu) is not finite");
    if (n_out) *n_out = N;
    if (dn_out) *dn_out = dN;
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_density(dmftgpu_handle h, int method, const double* sigma, double mu, double* n_out,
                               double* dn_dmu_out)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (!h->prepared) return fail(h, DMFTGPU_ERR_STATE, "dmftgpu_prepare has not been called");
    if (!valid_method(method)) return fail(h, DMFTGPU_ERR_ARG, "unknown method %d", method);
    if (method != DMFTGPU_PLAIN_F64 && !h->schur_ready)
        return fail(h, DMFTGPU_ERR_UNSUPPORTED, "Schur method needs n_corr <= 16 (n=%d, n_corr=%d)", h->lat.n, h->lat.n_corr);
    if (!sigma) return fail(h, DMFTGPU_ERR_ARG, "sigma is null");
    if (!std::isfinite(mu)) return fail(h, DMFTGPU_ERR_ARG, "mu is not finite");
    return density_impl(h, method, sigma, mu, n_out, dn_dmu_out);
}

// GreenFtn_w_adjust_mu (GreenFtn.cpp:66-139).  The reference brackets with
// steps of dmu and bisects until |dmu| <= 1e-5; here N(mu) is monotone
// increasing with the analytic slope dN/dmu, so Newton steps are taken inside
// a maintained bracket and replaced by bisection whenever they leave it.
extern "C" int dmftgpu_adjust_mu(dmftgpu_handle h, int method, const double* sigma, double mu0, double n_target,
                                 double tol, double* mu_out, int* n_eval)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    int rc0;
    if (!method_ok(h, method, &rc0)) return rc0;
    if (!sigma || !mu_out) return fail(h, DMFTGPU_ERR_ARG, "sigma and mu_out are required");
    if (!std::isfinite(mu0) || !std::isfinite(n_target)) return fail(h, DMFTGPU_ERR_ARG, "mu0 or n_target is not finite");
    if (!(tol > 0.0)) tol = 1e-5;
    double mu = mu0, N, dN;
    int evals = 0;
    double lo = -INFINITY, hi = INFINITY;  // N(lo) < target < N(hi)
    for (int it = 0; it < 200; ++it) {
        int rc = density_impl(h, method, sigma, mu, &N, &dN);
        ++evals;
        if (rc) return rc;
        const double r = N - n_target;
        if (r == 0.0) break;
        if (r < 0.0) lo = mu;
        else hi = mu;
        double step = (dN > 0.0) ? -r / dN : (r < 0.0 ? 1.0 : -1.0);
        double next = mu + step;
        if (std::isfinite(lo) && std::isfinite(hi)) {
            if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
        } else {
            // no bracket yet: limit the Newton step (the reference steps by dmu = 1)
            if (step > 1.0) next = mu + 1.0;
            if (step < -1.0) next = mu - 1.0;
        }
        if (std::fabs(next) > 600.0) return fail(h, DMFTGPU_ERR_NUMERIC, "mu search diverged (|mu| > 600)");
        const double dmu = next - mu;
        mu = next;
        if (std::fabs(dmu) <= tol) break;
    }
    *mu_out = mu;
    if (n_eval) *n_eval = evals;
    return DMFTGPU_OK;
}

namespace {

int vjp_impl(dmftgpu_handle h, int method, const double* sigma, double mu, const double* gloc_bar, double n_bar,
             double* sigma_bar, double* mu_bar, cd* fused_tr2, int fused_chunks)
{
    const dmftgpu_lattice& L = h->lat;
    const int nc = L.n_corr, m2 = nc * nc;
    const int nw_tot = L.n_freq * L.high_freq;
    hipError_t e;
    size_t need = lattice_ws_bytes(h, method, dg::MODE_VJP) + static_ws_bytes(L) + (size_t)(m2 + 8) * 32 + 8192;
    char* ws = (char*)h->ws.get(need, &e);
    if (!ws) return fail(h, DMFTGPU_ERR_HIP, "workspace %zu bytes: %s", need, hipGetErrorString(e));
    size_t off = 0;
    Sums S;
    int rc = run_lattice(h, method, dg::MODE_VJP, sigma, mu, gloc_bar, n_bar, ws, off, S, false, nullptr, nullptr,
                         fused_tr2, fused_chunks);
    if (rc) return rc;
    cd* tot = (cd*)(ws + off);
    off += ((size_t)(m2 + 3) * sizeof(cd) + 255) / 256 * 256;
    cd* st = (cd*)(ws + off);
    off += ((size_t)m2 * sizeof(cd) + 255) / 256 * 256;
    double* sts = (double*)(ws + off);
    off += 256;
    rc = run_static(h, mu, st, sts, ws, off);
    if (rc) return rc;
    const double inv_nk = 1.0 / (double)L.n_k_global;
    HIPCHK(dg::launch_scale(S.Sb, (size_t)L.n_freq * m2, inv_nk, (cd*)sigma_bar, h->stream));
    // (tr G itself is not needed here; after a fused forward only its first N_freq entries are formed)
    HIPCHK(dg::launch_freq_totals(nullptr, m2, nc, nullptr, S.tr2, nw_tot, S.mu, L.n_freq, tot, h->stream));
    std::vector<cd> htot(m2 + 3);
    double hsts[2];
    HIPCHK(hipMemcpyAsync(htot.data(), tot, sizeof(cd) * (m2 + 3), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(hsts, sts, sizeof(double) * 2, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const double dN = (-htot[m2 + 1].x / L.beta + hsts[1]) * inv_nk;
    const double mub = -htot[m2 + 2].x * inv_nk + n_bar * dN;
    if (!std::isfinite(mub)) return fail(h, DMFTGPU_ERR_NUMERIC, "mu_bar is not finite");
    if (mu_bar) *mu_bar = mub;
    return DMFTGPU_OK;
}


}  // namespace

extern "C" int dmftgpu_gloc_vjp(dmftgpu_handle h, int method, const double* sigma, double mu, const double* gloc_bar,
                                double n_bar, double* sigma_bar, double* mu_bar)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    int rc;
    if (!method_ok(h, method, &rc)) return rc;
    if (!sigma || !gloc_bar || !sigma_bar) return fail(h, DMFTGPU_ERR_ARG, "sigma, gloc_bar and sigma_bar are required");
    if (!std::isfinite(mu) || !std::isfinite(n_bar)) return fail(h, DMFTGPU_ERR_ARG, "mu or n_bar is not finite");
    return vjp_impl(h, method, sigma, mu, gloc_bar, n_bar, sigma_bar, mu_bar, nullptr, 0);
}

// Jacobian-vector product of (G_loc, N) in (Sigma(iw_n), n < N_freq; mu), moments
// held fixed.  Per (k, iw_n):  dG_cc = G (dSigma - dmu S') G  with (G^2)_cc = G S' G,
// and dN = (1/Nk) [ (2/beta) Re sum_{n<N_freq} tr(dSigma_n (G^2)_cc) ] + dmu dN/dmu.
// G X G is the conjugate transpose of G^H X^H G^H, the product the adjoint forms
// with V = gbar + c_n S'^H (c_n = 2 nbar / beta): with gbar = dSigma^H and
// c_n = -dmu the adjoint kernels accumulate exactly dG_cc^H, and their per-frequency
// mu terms tr(gbar^H (G^2)_cc) = tr(dSigma (G^2)_cc).  So the same kernels run with
// those inputs, and the results are transposed back.
static int jvp_impl(dmftgpu_handle h, int method, const double* sigma, double mu, const double* dsigma,
                    double dmu, double* dgloc, double* dn)
{
    const dmftgpu_lattice& L = h->lat;
    const int nc = L.n_corr, m2 = nc * nc;
    const int nw_tot = L.n_freq * L.high_freq;
    hipError_t e;
    const size_t gb_bytes = ((size_t)L.n_freq * m2 * sizeof(cd) + 255) / 256 * 256;
    size_t need = lattice_ws_bytes(h, method, dg::MODE_VJP) + static_ws_bytes(L) + (size_t)(m2 + 8) * 32 + 8192 +
                  gb_bytes;
    char* ws = (char*)h->ws.get(need, &e);
    if (!ws) return fail(h, DMFTGPU_ERR_HIP, "workspace %zu bytes: %s", need, hipGetErrorString(e));
    size_t off = 0;
    cd* gbarH = (cd*)ws;  // dSigma^H
    off += gb_bytes;
    HIPCHK(dg::launch_adjoint_stack((const cd*)dsigma, (size_t)L.n_freq, nc, 1.0, gbarH, h->stream));
    Sums S;
    int rc = run_lattice(h, method, dg::MODE_VJP, sigma, mu, (const double*)gbarH, -0.5 * dmu * L.beta, ws, off, S);
    if (rc) return rc;
    cd* tot = (cd*)(ws + off);
    off += ((size_t)(m2 + 3) * sizeof(cd) + 255) / 256 * 256;
    cd* st = (cd*)(ws + off);
    off += ((size_t)m2 * sizeof(cd) + 255) / 256 * 256;
    double* sts = (double*)(ws + off);
    off += 256;
    rc = run_static(h, mu, st, sts, ws, off);
    if (rc) return rc;
    const double inv_nk = 1.0 / (double)L.n_k_global;
    HIPCHK(dg::launch_adjoint_stack(S.Sb, (size_t)L.n_freq, nc, inv_nk, (cd*)dgloc, h->stream));
    HIPCHK(dg::launch_freq_totals(nullptr, m2, nc, nullptr, S.tr2, nw_tot, S.mu, L.n_freq, tot, h->stream));
    cd htot[32 * 32 + 3];
    double hsts[2];
    HIPCHK(hipMemcpyAsync(htot, tot, sizeof(cd) * (m2 + 3), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(hsts, sts, sizeof(double) * 2, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    const double dNdmu = (-htot[m2 + 1].x / L.beta + hsts[1]) * inv_nk;
    const double d = 2.0 / L.beta * htot[m2 + 2].x * inv_nk + dmu * dNdmu;
    if (!std::isfinite(d)) return fail(h, DMFTGPU_ERR_NUMERIC, "dN is not finite");
    if (dn) *dn = d;
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_gloc_jvp(dmftgpu_handle h, int method, const double* sigma, double mu, const double* dsigma,
                                double dmu, double* dgloc, double* dn)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    int rc;
    if (!method_ok(h, method, &rc)) return rc;
    if (!sigma || !dsigma || !dgloc) return fail(h, DMFTGPU_ERR_ARG, "sigma, dsigma and dgloc are required");
    if (!std::isfinite(mu) || !std::isfinite(dmu)) return fail(h, DMFTGPU_ERR_ARG, "mu or dmu is not finite");
    return jvp_impl(h, method, sigma, mu, dsigma, dmu, dgloc, dn);
}

// Loss L = sum_{n<N_freq} ||G_loc(iw_n) - G_tgt(iw_n)||_F^2 + (N - N_tgt)^2 and its
// gradient: forward, cotangents on the device, adjoint.  On the one-matrix-per-lane
// Schur path the forward also forms tr G^2 beyond N_freq (fused), so the adjoint
// only covers the first N_freq frequencies.
extern "C" int dmftgpu_gloc_loss_grad(dmftgpu_handle h, int method, const double* sigma, double mu,
                                      const double* gloc_target, double n_target, double* gloc, double* num_matrix,
                                      double* n_total, double* loss, double* sigma_bar, double* mu_bar)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    int rc;
    if (!method_ok(h, method, &rc)) return rc;
    if (!sigma || !gloc_target || !gloc || !num_matrix || !sigma_bar)
        return fail(h, DMFTGPU_ERR_ARG, "sigma, gloc_target, gloc, num_matrix and sigma_bar are required");
    if (!std::isfinite(mu) || !std::isfinite(n_target)) return fail(h, DMFTGPU_ERR_ARG, "mu or n_target is not finite");
    const dmftgpu_lattice& L = h->lat;
    const int nc = L.n_corr, m2 = nc * nc;
    const int nw_tot = L.n_freq * L.high_freq;
    hipError_t e;
    cd* ftr2 = nullptr;
    int fchunks = 0;
    if (method != DMFTGPU_PLAIN_F64 && dg::schur_can_fuse(nc, h->schur_layout == 1, false) && nw_tot > L.n_freq) {
        const int slots = dg::schur_slots(nc, dg::MODE_FWD, method == DMFTGPU_SCHUR_F32R, false);
        const Plan P = make_plan(L.n_k, nw_tot, slots, lattice_target(h, method, nc, false, L.n_k, nw_tot, slots));
        fchunks = P.n_chunks;
        ftr2 = (cd*)h->fused.get((size_t)P.n_chunks * nw_tot * sizeof(cd), &e);
        if (!ftr2) return fail(h, DMFTGPU_ERR_HIP, "fused partials: %s", hipGetErrorString(e));
    }
    double N = 0.0;
    rc = forward_impl(h, method, sigma, mu, gloc, nullptr, num_matrix, &N, ftr2, fchunks);
    if (rc) return rc;
    // cotangents: gbar = 2 (G_loc - G_tgt) on the returned (cluster-block-diagonal)
    // entries, 0 on the inter-cluster ones, which the output fixes at zero
    cd* gbar = (cd*)h->cot.get((size_t)L.n_freq * m2 * sizeof(cd) + 256, &e);
    if (!gbar) return fail(h, DMFTGPU_ERR_HIP, "cotangent buffer: %s", hipGetErrorString(e));
    double* dloss = (double*)(gbar + (size_t)L.n_freq * m2);
    HIPCHK(dg::launch_loss_cotangent((const cd*)gloc, (const cd*)gloc_target, L.n_freq, nc, nc / L.n_cluster, gbar, dloss,
                                     h->stream));
    const double nbar = 2.0 * (N - n_target);
    rc = vjp_impl(h, method, sigma, mu, (const double*)gbar, nbar, sigma_bar, mu_bar, ftr2, fchunks);
    if (rc) return rc;
    if (n_total) *n_total = N;
    if (loss) {
        double hl = 0.0;
        HIPCHK(hipMemcpyAsync(&hl, dloss, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        *loss = hl + (N - n_target) * (N - n_target);
    }
    return DMFTGPU_OK;
}

// densityMatDFT of GreenFtn_w (GreenFtn.cpp:291-319), full n x n per k.
extern "C" int dmftgpu_density_matrix(dmftgpu_handle h, const double* sigma, double mu, double* dens)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!h) return DMFTGPU_ERR_ARG;
    if (!h->prepared) return fail(h, DMFTGPU_ERR_STATE, "dmftgpu_prepare has not been called");
    if (!sigma || !dens) return fail(h, DMFTGPU_ERR_ARG, "sigma and dens are required");
    if (!std::isfinite(mu)) return fail(h, DMFTGPU_ERR_ARG, "mu is not finite");
    const dmftgpu_lattice& L = h->lat;
    const int m2 = L.n_corr * L.n_corr;
    hipError_t e;
    const size_t nw_tot = (size_t)L.n_freq * L.high_freq;
    const bool block = h->density_kernel == 0 && h->schur_ready && h->kdata && dg::density_lr_supported(L.n, L.n_corr);
    const size_t sig_bytes = block ? (nw_tot * m2 * sizeof(cd) + 255) / 256 * 256 : 0;
    size_t need = static_ws_bytes(L) + (size_t)(m2 + 8) * 32 + 8192 + sig_bytes;
    char* ws = (char*)h->ws.get(need, &e);
    if (!ws) return fail(h, DMFTGPU_ERR_HIP, "workspace %zu bytes: %s", need, hipGetErrorString(e));
    size_t off = 0;
    cd* st = (cd*)(ws + off);
    off += ((size_t)m2 * sizeof(cd) + 255) / 256 * 256;
    double* sts = (double*)(ws + off);
    off += 256;
    cd* sig_all = block ? (cd*)(ws + off) : nullptr;
    off += sig_bytes;
    const double* wt = nullptr;
    int rc = run_static(h, mu, st, sts, ws, off, false, &wt);  // the static weights for this mu
    if (rc) return rc;
    dg::LatArgs a{};
    a.Hk = h->Hk;
    a.sigma = (const cd*)sigma;
    a.mom = h->mom;
    a.n = L.n;
    a.nc = L.n_corr;
    a.nk = L.n_k;
    a.n_freq = L.n_freq;
    a.nw_tot = L.n_freq * L.high_freq;
    a.beta = L.beta;
    a.mu = mu;
    a.kdata = h->kdata;
    if (block) {
        a.pSigAll = sig_all;
        HIPCHK(dg::launch_sigma_all(a, h->stream));
        HIPCHK(dg::launch_density_matrix_lr(a, h->W, h->Vd, h->ed, h->U, wt, (cd*)dens, h->stream));
    } else {
        HIPCHK(dg::launch_density_matrix(a, h->U, wt, (cd*)dens, h->stream));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    return DMFTGPU_OK;
}
