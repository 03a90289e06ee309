// Batched lattice path: many independent lattices (uneven n, n_c, Nk) in one
// replayable step.
//
// Per system the work is the same as dmftgpu_greenftn_w + dmftgpu_gloc_vjp
// (GreenFtn.cpp:142-422 and its adjoint) on the loss
//     L = sum_w ||G_loc(iw) - G_tgt(iw)||_F^2 + (N - N_tgt)^2 ,
// but
//   * the per-(k, iw) kernels of all systems with the same n_c run as ONE
//     persistent launch over a static work list of (system, frequency tile,
//     k chunk) items, one launch per (n_c, mode): the one-matrix-per-lane
//     kernels for n_c <= 6 (k_lattice_schur_multi), the column-per-lane
//     kernels for 6 < n_c <= 16 (k_lattice_schur_cols_multi);
//   * everything else (static eigen-decompositions, Sigma tables, chunk
//     reductions, finalisation, loss cotangents) is device-side, so the whole
//     step is captured once into a hipGraph and replayed with hipGraphLaunch;
//   * an autotuner picks, per n_c group and mode, the workgroup width
//     (frequencies per block: 64 or 128; one matrix per lane) and the k chunk
//     size (work per item) by timing the candidates on the batch itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "../../include/dmftgpu.h"
#include "api_internal.hpp"
#include "lattice_kernels.hpp"

using dg::cd;
using dgi::fail;

namespace {

using namespace dg;

// num = (tot/beta + st) inv_nk ; N = (tr/beta + sum w) inv_nk     (GreenFtn.cpp:322-356)
__global__ void k_fin_fwd(const cd* __restrict__ tot, const cd* __restrict__ st, const double* __restrict__ sts,
                          int m2, double beta, double inv_nk, cd* __restrict__ num, double* __restrict__ n_total)
{
    const int q = threadIdx.x;
    if (q < m2) num[q] = cscale(cadd(cscale(tot[q], 1.0 / beta), st[q]), inv_nk);
    if (q == 0 && n_total) *n_total = (tot[m2].x / beta + sts[0]) * inv_nk;
}

// loss cotangents: gbar = 2 (G_loc - G_tgt) inside cluster blocks (npc orbitals
// each), 0 across them (G_loc is returned block-diagonal, so those entries do not
// depend on Sigma); nbar = 2 (N - N_tgt)
__global__ void k_loss_cot(const cd* __restrict__ gloc, const cd* __restrict__ gtgt, size_t len, int nc, int npc,
                           const double* __restrict__ n_total, double n_tgt, cd* __restrict__ gbar,
                           double* __restrict__ nbar)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) {
        const int e = (int)(i % ((size_t)nc * nc)), a = e % nc, b = e / nc;
        gbar[i] = (a / npc == b / npc) ? cscale(csub(gloc[i], gtgt[i]), 2.0) : cd{0.0, 0.0};
    }
    if (i == 0) *nbar = 2.0 * (*n_total - n_tgt);
}

// mu_bar = -Re sum tr(gbar^H (G^2)_cc)/Nk + nbar dN/dmu
__global__ void k_fin_vjp(const cd* __restrict__ tot, const double* __restrict__ sts, int m2, double beta,
                          double inv_nk, const double* __restrict__ nbar, double* __restrict__ mu_bar)
{
    const double dN = (-tot[m2 + 1].x / beta + sts[1]) * inv_nk;
    *mu_bar = -tot[m2 + 2].x * inv_nk + (*nbar) * dN;
}

unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

struct Sys {
    dmftgpu_system d;
    int n, nc, nd, nk, nf, nw_tot, m2;
    // device data
    cd *mom, *U, *Vd, *W, *kdata, *sig_all, *gbar;
    double *ev, *ed, *nbar;
    // static part
    double *wt, *dt, *pN;
    cd *pNum, *st, *tot_f, *tot_v;
    double* sts;
    int st_kpc, st_chunks;
    // lattice partials / sums (sized for the largest chunk count)
    int max_chunks;
    cd *pTrF, *pG, *pTrV, *pTr2, *pSb, *pMu;
    cd *trF, *Gsum, *trV, *tr2, *Sb, *musum;
    // chosen plans
    int chunks_f, kpc_f, chunks_v, kpc_v;
};

// per-mode slots of a group: 0 forward, 1 adjoint
inline int slot(int mode) { return mode == MODE_VJP ? 1 : 0; }

struct Group {
    int nc;
    bool cols = false;  // n_c > 6: column-per-lane kernels, one work list over every frequency
    std::vector<int> systems;
    int wg[2] = {64, 64};
    double target[2] = {0, 0};
    double ms[2] = {0, 0};
    LatArgs* dargs[2] = {nullptr, nullptr};
    int4* ditems[2] = {nullptr, nullptr};
    int n_items[2] = {0, 0};
    // frequencies n >= N_freq: the forward accumulates the Hermitian part of G_cc
    // (MODE_FWD_HERM), the adjoint only needs tr G, tr G^2 (density kernel); with
    // an adjoint in the step both happen in the forward (MODE_DENS_HERM, writing
    // the adjoint's tr G^2 partials, same k chunking) and the adjoint has no
    // upper work list
    int mode_hi[2] = {MODE_FWD_HERM, MODE_DENS};
    LatArgs* dargs_hi[2] = {nullptr, nullptr};
    int4* ditems_hi[2] = {nullptr, nullptr};
    int n_items_hi[2] = {0, 0};
    int max_nd = 0;
};

}  // namespace

struct dmftgpu_batch_s {
    dmftgpu_handle h = nullptr;   // internal handle on the batch's own stream
    dmftgpu_handle hu = nullptr;  // caller's handle (stream ordering, error text)
    hipStream_t bs = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    int method = DMFTGPU_SCHUR_F64;
    int flags = 0;
    bool vjp = false;
    std::vector<Sys> sys;
    std::vector<Group> groups;
    dgi::Workspace mem;
    std::vector<void*> extra;  // per-group args/items allocations
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

#define HIPCHK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) return fail(h, DMFTGPU_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

namespace {

// rough per-(k, iw) cost of the Schur work, for k chunking
double lane_cost(int nc, int nd) { return 8.0 * nd * nc * (nc + 1) / 2 + 10.0 * nc * nc * nc + 100.0; }

int plan_chunks(const Sys& s, double target, int* kpc)
{
    int k = (int)std::max(1.0, std::floor(target / lane_cost(s.nc, s.nd)));
    if (k > s.nk) k = s.nk;
    int chunks = (s.nk + k - 1) / k;
    if (chunks > s.max_chunks) {
        chunks = s.max_chunks;
        k = (s.nk + chunks - 1) / chunks;
        chunks = (s.nk + k - 1) / k;
    }
    *kpc = k;
    return chunks;
}

LatArgs make_args(const dmftgpu_batch_s* b, const Sys& s, int mode, int chunks, int kpc)
{
    LatArgs a{};
    a.Hk = (const cd*)s.d.Hk;
    a.sigma = (const cd*)s.d.sigma;
    a.mom = s.mom;
    a.n = s.n;
    a.nc = s.nc;
    a.nk = s.nk;
    a.n_freq = s.nf;
    a.nw_tot = s.nw_tot;
    a.beta = s.d.lat.beta;
    a.mu = s.d.mu;
    a.k_per_chunk = kpc;
    a.n_chunks = chunks;
    a.ed = s.ed;
    a.W = s.W;
    a.kdata = s.kdata;
    a.refine = b->h->refine_steps;
    a.pSigAll = s.sig_all;
    if (mode == MODE_FWD) {
        a.pTr = s.pTrF;
        a.pG = s.pG;
        a.want_trace = 1;
    } else {
        a.pTr = s.pTrV;
        a.pTr2 = s.pTr2;
        a.pSb = s.pSb;
        a.pMu = s.pMu;
        a.gbar = s.gbar;
        a.nbar_dev = s.nbar;
    }
    return a;
}

// Builds the device args / work list of one group for one mode with (wg, target).
int build_group(dmftgpu_batch_s* b, Group& g, int mode, int wg, double target)
{
    dmftgpu_handle h = b->h;
    std::vector<LatArgs> args, args_hi;
    std::vector<int4> items, items_hi;
    for (size_t q = 0; q < g.systems.size(); ++q) {
        Sys& s = b->sys[g.systems[q]];
        int kpc;
        const int chunks = plan_chunks(s, target, &kpc);
        if (mode == MODE_FWD) {
            s.chunks_f = chunks;
            s.kpc_f = kpc;
        } else {
            s.chunks_v = chunks;
            s.kpc_v = kpc;
        }
        LatArgs a = make_args(b, s, mode, chunks, kpc);
        const bool split = s.nw_tot > s.nf && !g.cols;
        if (split) {  // [0, nf) forward / adjoint, [nf, nw_tot) see Group::mode_hi
            LatArgs hi = a;
            a.nw_end = s.nf;
            hi.nw_begin = s.nf;
            if (mode == MODE_FWD && b->vjp) hi.pTr2 = s.pTr2;  // fused: the adjoint's partials
            args_hi.push_back(hi);
        } else {
            args_hi.push_back(a);  // unused placeholder keeps the system index
        }
        args.push_back(a);
        // (the k-pair form covers 64 frequencies per item; the upper launch runs at 64)
        const int wf = wg == dg::SCHUR_WG_KPAIR ? 64 : wg;
        const int per_tile = g.cols ? dg::schur_slots(s.nc, mode, b->method == DMFTGPU_SCHUR_F32R, true) : wf;
        const int nwt = ((split ? s.nf : s.nw_tot) + per_tile - 1) / per_tile;
        for (int c = 0; c < chunks; ++c)
            for (int t = 0; t < nwt; ++t) items.push_back(int4{(int)q, t, c, 0});
        if (split && !(mode == MODE_VJP && b->vjp)) {  // fused adjoint: the forward did these
            const int nwh = (s.nw_tot - s.nf + wf - 1) / wf;
            for (int c = 0; c < chunks; ++c)
                for (int t = 0; t < nwh; ++t) items_hi.push_back(int4{(int)q, t, c, 0});
        }
    }
    g.mode_hi[0] = b->vjp ? MODE_DENS_HERM : MODE_FWD_HERM;
    // interleave systems so neighbouring workgroups mix large and small items
    std::stable_sort(items.begin(), items.end(), [](const int4& x, const int4& y) { return x.z < y.z; });
    std::stable_sort(items_hi.begin(), items_hi.end(), [](const int4& x, const int4& y) { return x.z < y.z; });
    if (!g.dargs[slot(mode)]) {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, sizeof(LatArgs) * args.size()));
        b->extra.push_back(p);
        g.dargs[slot(mode)] = (LatArgs*)p;
    }
    void* pi = nullptr;
    HIPCHK(hipMalloc(&pi, sizeof(int4) * items.size()));
    b->extra.push_back(pi);
    g.ditems[slot(mode)] = (int4*)pi;
    g.n_items[slot(mode)] = (int)items.size();
    g.wg[slot(mode)] = wg;
    g.target[slot(mode)] = target;
    HIPCHK(hipMemcpy(g.dargs[slot(mode)], args.data(), sizeof(LatArgs) * args.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(g.ditems[slot(mode)], items.data(), sizeof(int4) * items.size(), hipMemcpyHostToDevice));
    const int sl = slot(mode);
    g.n_items_hi[sl] = (int)items_hi.size();
    if (!g.dargs_hi[sl]) {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, sizeof(LatArgs) * args_hi.size()));
        b->extra.push_back(p);
        g.dargs_hi[sl] = (LatArgs*)p;
    }
    HIPCHK(hipMemcpy(g.dargs_hi[sl], args_hi.data(), sizeof(LatArgs) * args_hi.size(), hipMemcpyHostToDevice));
    if (!items_hi.empty()) {
        void* p = nullptr;
        HIPCHK(hipMalloc(&p, sizeof(int4) * items_hi.size()));
        b->extra.push_back(p);
        g.ditems_hi[sl] = (int4*)p;
        HIPCHK(hipMemcpy(g.ditems_hi[sl], items_hi.data(), sizeof(int4) * items_hi.size(), hipMemcpyHostToDevice));
    }
    return DMFTGPU_OK;
}

int launch_group(dmftgpu_batch_s* b, Group& g, int mode)
{
    dmftgpu_handle h = b->h;
    const bool fp32 = b->method == DMFTGPU_SCHUR_F32R;
    LatArgs proto{};
    proto.nc = g.nc;
    proto.refine = h->refine_steps;
    proto.want_trace = 1;
    SchurMulti m;
    m.args = g.dargs[slot(mode)];
    m.items = g.ditems[slot(mode)];
    m.n_items = g.n_items[slot(mode)];
    m.grid = std::min(g.n_items[slot(mode)], 256 * 16);
    m.max_nd = g.max_nd;
    if (g.cols) {  // every frequency in the one list (no upper-half work list)
        HIPCHK(dg::launch_lattice_schur_cols_multi(proto, mode, fp32, h->stream, m));
        return DMFTGPU_OK;
    }
    HIPCHK(dg::launch_lattice_schur_multi(proto, mode, fp32, h->stream, g.wg[slot(mode)], m));
    if (g.n_items_hi[slot(mode)] > 0) {
        SchurMulti mh = m;
        mh.args = g.dargs_hi[slot(mode)];
        mh.items = g.ditems_hi[slot(mode)];
        mh.n_items = g.n_items_hi[slot(mode)];
        mh.grid = std::min(mh.n_items, 256 * 16);
        const int wg_hi = g.wg[slot(mode)] == dg::SCHUR_WG_KPAIR ? 64 : g.wg[slot(mode)];
        HIPCHK(dg::launch_lattice_schur_multi(proto, g.mode_hi[slot(mode)], fp32, h->stream, wg_hi, mh));
    }
    return DMFTGPU_OK;
}

// prepare: moments copy, static eig, Schur data (GreenFtn.cpp:197-215)
int issue_prepare(dmftgpu_batch_s* b)
{
    dmftgpu_handle h = b->h;
    for (Sys& s : b->sys) {
        HIPCHK(hipMemcpyAsync(s.mom, s.d.moments, sizeof(cd) * 4 * s.m2, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(dg::launch_herm_eig((const cd*)s.d.Hk, s.n, 0, s.n, s.mom, s.nc, s.ev, s.U, s.nk, h->stream));
        if (s.nd > 0) {
            HIPCHK(dg::launch_herm_eig((const cd*)s.d.Hk, s.n, s.nc, s.nd, nullptr, 0, s.ed, s.Vd, s.nk, h->stream));
            HIPCHK(dg::launch_schur_prep((const cd*)s.d.Hk, s.Vd, s.n, s.nc, s.nk, s.W, h->stream));
        }
        if (s.kdata) HIPCHK(dg::launch_schur_kdata((const cd*)s.d.Hk, s.W, s.ed, s.n, s.nc, s.nk, s.kdata, h->stream));
    }
    return DMFTGPU_OK;
}

int issue_sigma_tables(dmftgpu_batch_s* b)
{
    dmftgpu_handle h = b->h;
    for (Sys& s : b->sys) {
        LatArgs a = make_args(b, s, MODE_FWD, 1, s.nk);
        HIPCHK(dg::launch_sigma_all(a, h->stream));
    }
    return DMFTGPU_OK;
}

// the whole step (stream-ordered, capturable: no host synchronisation)
int issue_step(dmftgpu_batch_s* b)
{
    dmftgpu_handle h = b->h;
    if (b->flags & DMFTGPU_BATCH_PREPARE) {
        int rc = issue_prepare(b);
        if (rc) return rc;
    }
    for (Sys& s : b->sys) {  // Sigma tables and static weights
        LatArgs a = make_args(b, s, MODE_FWD, 1, s.nk);
        HIPCHK(dg::launch_sigma_all(a, h->stream));
        const double beta = s.d.lat.beta;
        HIPCHK(dg::launch_static_terms(s.ev, s.n, s.nk, s.nw_tot, beta, s.d.mu, s.wt, s.dt, h->stream));
        HIPCHK(dg::launch_static_reduce(s.wt, s.dt, s.U, s.n, s.nc, s.nk, s.st_kpc, s.st_chunks, s.pNum, s.pN,
                                        s.pN + s.st_chunks, h->stream));
        HIPCHK(dg::launch_reduce_chunks(s.pNum, s.st_chunks, s.m2, s.st, h->stream));
        HIPCHK(dg::launch_reduce_chunks_d(s.pN, s.st_chunks, 1, s.sts, h->stream));
        HIPCHK(dg::launch_reduce_chunks_d(s.pN + s.st_chunks, s.st_chunks, 1, s.sts + 1, h->stream));
    }
    for (Group& g : b->groups) {  // forward: one persistent launch per n_c
        int rc = launch_group(b, g, MODE_FWD);
        if (rc) return rc;
    }
    for (Sys& s : b->sys) {
        const double inv_nk = 1.0 / (double)s.d.lat.n_k_global;
        HIPCHK(dg::launch_reduce_chunks(s.pTrF, s.chunks_f, s.nw_tot, s.trF, h->stream));
        HIPCHK(dg::launch_reduce_chunks(s.pG, s.chunks_f, (size_t)s.nw_tot * s.m2, s.Gsum, h->stream));
        HIPCHK(dg::launch_freq_totals(s.Gsum, s.m2, s.nc, s.trF, nullptr, s.nw_tot, nullptr, 0, s.tot_f, h->stream));
        HIPCHK(dg::launch_gloc_finalize(s.Gsum, s.nf, s.nc, s.d.lat.n_cluster, inv_nk, (cd*)s.d.gloc, h->stream));
        hipLaunchKernelGGL(k_fin_fwd, dim3(1), dim3(1024), 0, h->stream, s.tot_f, s.st, s.sts, s.m2, s.d.lat.beta,
                           inv_nk, (cd*)s.d.num_matrix, s.d.n_total);
        HIPCHK(hipGetLastError());
    }
    if (!b->vjp) return DMFTGPU_OK;
    for (Sys& s : b->sys) {
        const size_t len = (size_t)s.nf * s.m2;
        hipLaunchKernelGGL(k_loss_cot, dim3(nblk(len)), dim3(256), 0, h->stream, (const cd*)s.d.gloc,
                           (const cd*)s.d.gloc_target, len, s.nc, s.nc / s.d.lat.n_cluster, s.d.n_total, s.d.n_target, s.gbar,
                           s.nbar);
        HIPCHK(hipGetLastError());
    }
    for (Group& g : b->groups) {  // adjoint: one persistent launch per n_c
        int rc = launch_group(b, g, MODE_VJP);
        if (rc) return rc;
    }
    for (Sys& s : b->sys) {
        const double inv_nk = 1.0 / (double)s.d.lat.n_k_global;
        HIPCHK(dg::launch_reduce_chunks(s.pTrV, s.chunks_v, s.nw_tot, s.trV, h->stream));
        HIPCHK(dg::launch_reduce_chunks(s.pTr2, s.chunks_v, s.nw_tot, s.tr2, h->stream));
        HIPCHK(dg::launch_reduce_chunks(s.pSb, s.chunks_v, (size_t)s.nf * s.m2, s.Sb, h->stream));
        HIPCHK(dg::launch_reduce_chunks(s.pMu, s.chunks_v, s.nf, s.musum, h->stream));
        HIPCHK(dg::launch_scale(s.Sb, (size_t)s.nf * s.m2, inv_nk, (cd*)s.d.sigma_bar, h->stream));
        HIPCHK(dg::launch_freq_totals(nullptr, s.m2, s.nc, s.trV, s.tr2, s.nw_tot, s.musum, s.nf, s.tot_v, h->stream));
        hipLaunchKernelGGL(k_fin_vjp, dim3(1), dim3(1), 0, h->stream, s.tot_v, s.sts, s.m2, s.d.lat.beta, inv_nk,
                           s.nbar, s.d.mu_bar);
        HIPCHK(hipGetLastError());
    }
    return DMFTGPU_OK;
}

// the k-pair form exists for the FP32R forward of the one-matrix-per-lane kernels at n_c >= 5
bool kpair_ok(const dmftgpu_batch_s* b, const Group& g, int mode)
{
    return mode == MODE_FWD && b->method == DMFTGPU_SCHUR_F32R && !g.cols && g.nc >= 5;
}

// times one mode's group launch for each candidate and keeps the fastest
int tune_group(dmftgpu_batch_s* b, Group& g, int mode)
{
    dmftgpu_handle h = b->h;
    const int wgs[3] = {64, 128, dg::SCHUR_WG_KPAIR};
    const double targets[4] = {6.25e4, 2.5e5, 1.0e6, 4.0e6};
    struct Events {  // destroyed on every return path
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } ev;
    HIPCHK(hipEventCreate(&ev.e0));
    HIPCHK(hipEventCreate(&ev.e1));
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    double best = 1e300;
    int bw = 64;
    double bt = targets[2];
    for (int wg : wgs)
        for (double t : targets) {
            if (g.cols && wg != 64) continue;  // one width: a wavefront per workgroup
            if (wg == dg::SCHUR_WG_KPAIR && !kpair_ok(b, g, mode)) continue;
            int rc = build_group(b, g, mode, wg, t);
            if (rc) return rc;
            rc = launch_group(b, g, mode);  // warm-up
            if (rc) return rc;
            HIPCHK(hipEventRecord(e0, h->stream));
            rc = launch_group(b, g, mode);
            if (rc) return rc;
            HIPCHK(hipEventRecord(e1, h->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                bw = wg;
                bt = t;
            }
        }
    g.ms[slot(mode)] = best;
    return build_group(b, g, mode, bw, bt);
}

// Forward and adjoint of a fused step share the k chunking (the forward writes the
// adjoint's tr G^2 partials beyond N_freq), so one target is chosen for both: per
// target the best workgroup width of each, the target with the least sum.
int tune_group_pair(dmftgpu_batch_s* b, Group& g)
{
    dmftgpu_handle h = b->h;
    const int wgs[3] = {64, 128, dg::SCHUR_WG_KPAIR};
    const double targets[4] = {6.25e4, 2.5e5, 1.0e6, 4.0e6};
    struct Events {  // destroyed on every return path
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } ev;
    HIPCHK(hipEventCreate(&ev.e0));
    HIPCHK(hipEventCreate(&ev.e1));
    double best_sum = 1e300, best_ms[2] = {0, 0};
    int best_wg[2] = {64, 64};
    double best_t = targets[2];
    for (double t : targets) {
        double ms_mode[2];
        int wg_mode[2];
        for (int mode : {MODE_FWD, MODE_VJP}) {
            double bm = 1e300;
            int bw = 64;
            for (int wg : wgs) {
                if (g.cols && wg != 64) continue;  // one width: a wavefront per workgroup
                if (wg == dg::SCHUR_WG_KPAIR && !kpair_ok(b, g, mode)) continue;
                int rc = build_group(b, g, mode, wg, t);
                if (rc) return rc;
                rc = launch_group(b, g, mode);  // warm-up
                if (rc) return rc;
                HIPCHK(hipEventRecord(ev.e0, h->stream));
                rc = launch_group(b, g, mode);
                if (rc) return rc;
                HIPCHK(hipEventRecord(ev.e1, h->stream));
                HIPCHK(hipEventSynchronize(ev.e1));
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, ev.e0, ev.e1));
                if (ms < bm) {
                    bm = ms;
                    bw = wg;
                }
            }
            ms_mode[slot(mode)] = bm;
            wg_mode[slot(mode)] = bw;
        }
        if (ms_mode[0] + ms_mode[1] < best_sum) {
            best_sum = ms_mode[0] + ms_mode[1];
            best_t = t;
            for (int q = 0; q < 2; ++q) {
                best_ms[q] = ms_mode[q];
                best_wg[q] = wg_mode[q];
            }
        }
    }
    g.ms[0] = best_ms[0];
    g.ms[1] = best_ms[1];
    int rc = build_group(b, g, MODE_FWD, best_wg[0], best_t);
    if (rc) return rc;
    return build_group(b, g, MODE_VJP, best_wg[1], best_t);
}

}  // namespace

static int batch_create_impl(dmftgpu_batch_s* b, int count, const dmftgpu_system* systems, int method, int flags);

extern "C" int dmftgpu_batch_create(dmftgpu_handle hu, int count, const dmftgpu_system* systems, int method, int flags,
                                    dmftgpu_batch* out)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!hu || !out || !systems || count < 1) return hu ? fail(hu, DMFTGPU_ERR_ARG, "bad batch arguments") : DMFTGPU_ERR_ARG;
    if (method != DMFTGPU_SCHUR_F64 && method != DMFTGPU_SCHUR_F32R)
        return fail(hu, DMFTGPU_ERR_UNSUPPORTED, "batched path runs the Schur methods");
    auto* b = new dmftgpu_batch_s();
    b->hu = hu;
    hipError_t e = hipStreamCreateWithFlags(&b->bs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming);
    if (e != hipSuccess || dmftgpu_create(&b->h, b->bs) != DMFTGPU_OK) {
        if (b->bs) (void)hipStreamDestroy(b->bs);
        delete b;
        return fail(hu, DMFTGPU_ERR_HIP, "batch stream: %s", hipGetErrorString(e));
    }
    b->h->refine_steps = hu->refine_steps;
    // inputs were produced on the caller's stream
    (void)hipEventRecord(b->ev_in, hu->stream);
    (void)hipStreamWaitEvent(b->bs, b->ev_in, 0);
    int rc = batch_create_impl(b, count, systems, method, flags);
    if (rc) {
        hu->err = b->h->err;
        dmftgpu_batch_destroy(b);
        return rc;
    }
    *out = b;
    return DMFTGPU_OK;
}

static int batch_create_impl(dmftgpu_batch_s* b, int count, const dmftgpu_system* systems, int method, int flags)
{
    dmftgpu_handle h = b->h;
    b->method = method;
    b->flags = flags;
    b->sys.resize(count);
    std::map<int, int> gidx;
    size_t total = 0;
    for (int i = 0; i < count; ++i) {
        Sys& s = b->sys[i];
        s.d = systems[i];
        const dmftgpu_lattice& L = s.d.lat;
        if (L.n < 1 || L.n > 32 || L.n_corr < 1 || !dg::schur_supported(L.n, L.n_corr) || L.n_k < 1 ||
            L.n_freq < 1 || L.n_cluster < 1 || L.n_corr % L.n_cluster || !(L.beta > 0) || !s.d.Hk || !s.d.sigma ||
            !s.d.moments || !s.d.gloc || !s.d.num_matrix || !s.d.n_total) {
            return fail(h, DMFTGPU_ERR_ARG,
                        "system %d: invalid descriptor or unsupported size (n_corr <= 16, n <= 32)", i);
        }
        const bool v = s.d.gloc_target && s.d.sigma_bar && s.d.mu_bar;
        if (i == 0) b->vjp = v;
        if (v != b->vjp) {
            return fail(h, DMFTGPU_ERR_ARG, "either every system carries a loss target or none does");
        }
        s.n = L.n;
        s.nc = L.n_corr;
        s.nd = L.n - L.n_corr;
        s.nk = L.n_k;
        s.nf = L.n_freq;
        s.nw_tot = L.n_freq * L.high_freq;
        s.m2 = s.nc * s.nc;
        s.max_chunks = std::max(1, std::min(s.nk, 128));
        s.st_kpc = (s.nk + 255) / 256;
        s.st_chunks = (s.nk + s.st_kpc - 1) / s.st_kpc;
        auto it = gidx.find(s.nc);
        if (it == gidx.end()) {
            gidx[s.nc] = (int)b->groups.size();
            Group g;
            g.nc = s.nc;
            g.cols = !dg::schur_multi_supported(s.nc);
            b->groups.push_back(g);
            it = gidx.find(s.nc);
        }
        Group& g = b->groups[it->second];
        g.systems.push_back(i);
        g.max_nd = std::max(g.max_nd, s.nd);
    }
    // one allocation for every per-system buffer
    auto rnd = [](size_t x) { return (x + 255) / 256 * 256; };
    std::vector<size_t> need(count);
    for (int i = 0; i < count; ++i) {
        const Sys& s = b->sys[i];
        const size_t C = s.max_chunks, ndd = std::max(1, s.nd);
        size_t t = 0;
        t += rnd(16 * 4 * s.m2) + rnd(8 * (size_t)s.nk * s.n) + rnd(16 * (size_t)s.nk * s.n * s.n);
        t += rnd(8 * (size_t)s.nk * ndd) + rnd(16 * (size_t)s.nk * ndd * ndd) +
             rnd(16 * (size_t)s.nk * s.nc * ndd + dg::kKdataPadBytes);
        if (dg::schur_multi_supported(s.nc))
            t += rnd(16 * (size_t)s.nk * dg::schur_kdata_size(s.n, s.nc) + dg::kKdataPadBytes);
        t += rnd(16 * (size_t)s.nw_tot * s.m2) + rnd(16 * (size_t)s.nf * s.m2) + rnd(8);
        t += 2 * rnd(8 * (size_t)s.nk * s.n) + rnd(8 * 2 * (size_t)s.st_chunks) + rnd(16 * (size_t)s.st_chunks * s.m2);
        t += rnd(16 * s.m2) + 2 * rnd(16 * (s.m2 + 3)) + rnd(16);
        t += 2 * rnd(16 * C * s.nw_tot) + rnd(16 * C * s.nw_tot * s.m2) + rnd(16 * C * s.nw_tot);
        t += rnd(16 * C * s.nf * s.m2) + rnd(16 * C * s.nf);
        t += 3 * rnd(16 * (size_t)s.nw_tot) + rnd(16 * (size_t)s.nw_tot * s.m2) + rnd(16 * (size_t)s.nf * s.m2) +
             rnd(16 * (size_t)s.nf);
        need[i] = t;
        total += t;
    }
    hipError_t e;
    char* base = (char*)b->mem.get(total, &e);
    if (!base) return fail(h, DMFTGPU_ERR_HIP, "batch allocation %zu bytes: %s", total, hipGetErrorString(e));
    size_t off = 0;
    for (int i = 0; i < count; ++i) {
        Sys& s = b->sys[i];
        const size_t C = s.max_chunks, ndd = std::max(1, s.nd);
        auto take = [&](size_t bytes) {
            char* p = base + off;
            off += rnd(bytes);
            return (void*)p;
        };
        s.mom = (cd*)take(16 * 4 * s.m2);
        s.ev = (double*)take(8 * (size_t)s.nk * s.n);
        s.U = (cd*)take(16 * (size_t)s.nk * s.n * s.n);
        s.ed = (double*)take(8 * (size_t)s.nk * ndd);
        s.Vd = (cd*)take(16 * (size_t)s.nk * ndd * ndd);
        s.W = (cd*)take(16 * (size_t)s.nk * s.nc * ndd + dg::kKdataPadBytes);
        s.kdata = dg::schur_multi_supported(s.nc)
                      ? (cd*)take(16 * (size_t)s.nk * dg::schur_kdata_size(s.n, s.nc) + dg::kKdataPadBytes)
                      : nullptr;
        s.sig_all = (cd*)take(16 * (size_t)s.nw_tot * s.m2);
        s.gbar = (cd*)take(16 * (size_t)s.nf * s.m2);
        s.nbar = (double*)take(8);
        s.wt = (double*)take(8 * (size_t)s.nk * s.n);
        s.dt = (double*)take(8 * (size_t)s.nk * s.n);
        s.pN = (double*)take(8 * 2 * (size_t)s.st_chunks);
        s.pNum = (cd*)take(16 * (size_t)s.st_chunks * s.m2);
        s.st = (cd*)take(16 * s.m2);
        s.tot_f = (cd*)take(16 * (s.m2 + 3));
        s.tot_v = (cd*)take(16 * (s.m2 + 3));
        s.sts = (double*)take(16);
        s.pTrF = (cd*)take(16 * C * s.nw_tot);
        s.pTrV = (cd*)take(16 * C * s.nw_tot);
        s.pG = (cd*)take(16 * C * s.nw_tot * s.m2);
        s.pTr2 = (cd*)take(16 * C * s.nw_tot);
        s.pSb = (cd*)take(16 * C * s.nf * s.m2);
        s.pMu = (cd*)take(16 * C * s.nf);
        s.trF = (cd*)take(16 * (size_t)s.nw_tot);
        s.trV = (cd*)take(16 * (size_t)s.nw_tot);
        s.tr2 = (cd*)take(16 * (size_t)s.nw_tot);
        s.Gsum = (cd*)take(16 * (size_t)s.nw_tot * s.m2);
        s.Sb = (cd*)take(16 * (size_t)s.nf * s.m2);
        s.musum = (cd*)take(16 * (size_t)s.nf);
    }
    int rc = issue_prepare(b);
    if (!rc) rc = issue_sigma_tables(b);
    if (rc) return rc;
    for (Group& g : b->groups) {
        for (int mode : {MODE_FWD, MODE_VJP}) {
            if (mode == MODE_VJP && !b->vjp) continue;
            const int wg0 = kpair_ok(b, g, mode) ? dg::SCHUR_WG_KPAIR : 64;  // as the single-lattice default
            rc = (flags & DMFTGPU_BATCH_TUNE) ? 0 : build_group(b, g, mode, wg0, 1.0e6);
            if (rc) return rc;
        }
    }
    if (flags & DMFTGPU_BATCH_TUNE) {
        if (!b->vjp) {
            for (Group& g : b->groups)
                if ((rc = tune_group(b, g, MODE_FWD))) break;
        } else {
            // the adjoint needs the forward outputs (loss cotangents) in place while it is timed
            for (Group& g : b->groups)
                if (!rc) rc = build_group(b, g, MODE_FWD, 64, 1.0e6);
            for (Group& g : b->groups)
                if (!rc) rc = build_group(b, g, MODE_VJP, 64, 1.0e6);
            if (!rc) rc = issue_step(b);
            if (!rc && hipStreamSynchronize(h->stream) != hipSuccess) rc = fail(h, DMFTGPU_ERR_HIP, "tuning step");
            for (Group& g : b->groups)
                if (!rc) rc = tune_group_pair(b, g);
        }
        if (rc) return rc;
    }
    if (flags & DMFTGPU_BATCH_GRAPH) {
        hipError_t ce = hipStreamBeginCapture(h->stream, hipStreamCaptureModeGlobal);
        if (ce != hipSuccess) return fail(h, DMFTGPU_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(ce));
        rc = issue_step(b);
        hipGraph_t graph = nullptr;
        ce = hipStreamEndCapture(h->stream, &graph);
        if (rc || ce != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc ? rc : fail(h, DMFTGPU_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
        }
        b->graph = graph;
        ce = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
        if (ce != hipSuccess) return fail(h, DMFTGPU_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ce));
    }
    return DMFTGPU_OK;
}

// One step: ordered after the caller's stream and before its later work.
extern "C" int dmftgpu_batch_run(dmftgpu_batch b)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!b) return DMFTGPU_ERR_ARG;
    dmftgpu_handle h = b->hu;
    HIPCHK(hipEventRecord(b->ev_in, b->hu->stream));
    HIPCHK(hipStreamWaitEvent(b->bs, b->ev_in, 0));
    if (b->exec) {
        HIPCHK(hipGraphLaunch(b->exec, b->bs));
    } else {
        int rc = issue_step(b);
        if (rc) {
            b->hu->err = b->h->err;
            return rc;
        }
    }
    HIPCHK(hipEventRecord(b->ev_out, b->bs));
    HIPCHK(hipStreamWaitEvent(b->hu->stream, b->ev_out, 0));
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_batch_tuning(dmftgpu_batch b, int group, int* nc, int* systems, int* wg, double* target,
                                    double* ms)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!b) return DMFTGPU_ERR_ARG;
    if (group < 0 || group >= (int)b->groups.size()) return (int)b->groups.size();
    const Group& g = b->groups[group];
    if (nc) *nc = g.nc;
    if (systems) *systems = (int)g.systems.size();
    for (int m = 0; m < 2; ++m) {
        if (wg) wg[m] = g.wg[m];
        if (target) target[m] = g.target[m];
        if (ms) ms[m] = g.ms[m];
    }
    return DMFTGPU_OK;
}

extern "C" int dmftgpu_batch_destroy(dmftgpu_batch b)
{
    (void)hipGetLastError();  // drop an error left by an earlier failed call
    if (!b) return DMFTGPU_ERR_ARG;
    if (b->bs) (void)hipStreamSynchronize(b->bs);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void* p : b->extra) (void)hipFree(p);
    b->mem.release();
    if (b->h) dmftgpu_destroy(b->h);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_out) (void)hipEventDestroy(b->ev_out);
    if (b->bs) (void)hipStreamDestroy(b->bs);
    delete b;
    return DMFTGPU_OK;
}
