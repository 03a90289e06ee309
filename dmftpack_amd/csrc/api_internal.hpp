// Handle state shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/dmftgpu.h"
#include "dev_common.hpp"

namespace dgi {

// Grow-only device allocation owned by a handle.
struct Workspace {
    void* ptr = nullptr;
    size_t cap = 0;
    void* get(size_t bytes, hipError_t* e);
    void release();
};

int fail(dmftgpu_handle h, int code, const char* fmt, ...);

// Scratch of dmftgpu_greenftn_w_noalloc: a file-scope device array of this
// size per device (noalloc.hip); scratch_address returns its address on
// `device`, which must be current (the lookup runs once per device).
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
hipError_t scratch_address(int device, void** out);

}  // namespace dgi

struct dmftgpu_handle_s {
    hipStream_t stream = nullptr;
    int device = 0;
    void* scratch = nullptr;  // this device's no-alloc scratch (dgi::scratch_address), or null
    std::string err;
    // lattice bound by dmftgpu_prepare
    bool prepared = false;
    bool schur_ready = false;
    dmftgpu_lattice lat{};
    const dg::cd* Hk = nullptr;
    dg::cd* mom = nullptr;   // [4][nc][nc]
    double* ev = nullptr;    // [nk][n]    static eigenvalues
    dg::cd* U = nullptr;     // [nk][n][n] static eigenvectors
    double* ed = nullptr;    // [nk][nd]   eig(H_dd)
    dg::cd* Vd = nullptr;    // [nk][nd][nd]
    dg::cd* W = nullptr;     // [nk][nc][nd] H_cd V_d
    dg::cd* kdata = nullptr; // [nk][H_cc | P_d pairs | e_d] of the one-matrix-per-lane Schur kernels
    int target_wg = 0;       // work items per lattice launch (k chunking); 0: lattice_target's default
    // static-part totals for the last mu (valid until the next prepare)
    unsigned prep_gen = 0, static_gen = ~0u;
    double static_mu = 0.0;
    dgi::Workspace static_cache;  // [m2 cd][2 doubles]
    int refine_steps = 2;    // FP64-residual refinement steps of the FP32 Schur solve
    int schur_wg = 0;        // workgroup width of the Schur kernels (64 or 128; 0 = 64 forward, 128 adjoint)
    int schur_layout = 0;    // n_c <= 6: 0 one matrix per lane, 1 column per lane
    int upfold_chunk = 0;    // k per chunk of dmftgpu_upfold_density (0: scratch-bound)
    int density_kernel = 0;  // dmftgpu_density_matrix: 0 block form where supported, 1 full inverse
    bool static_cache_on = true;  // reuse static weights for a repeated mu
    dgi::Workspace prep;     // prepare-time data
    dgi::Workspace ws;       // per-call scratch
    dgi::Workspace ft;       // Fourier-transform scratch
    // moment fit: basis table + inverted normal matrices per (n_freq, beta, branch);
    // kMomCache grids are kept (a DMFT iteration alternates the Sigma-like and
    // G-like fits), the least recently used one is formed again
    static constexpr int kMomCache = 4;
    struct MomEntry {
        dgi::Workspace buf;
        bool valid = false;
        int nf = 0, green = 0;
        double beta = 0.0;
        unsigned long long used = 0;
    } mom_cache[kMomCache];
    unsigned long long mom_clock = 0;
    dgi::Workspace stat;     // static weights wt/dt [nk][n] of the cached mu
    // Kohn-Sham bands outside the model window (core-band term of GiwToNele,
    // GreenFtn.cpp:622-631), bound by dmftgpu_set_core_bands: energies
    // [nk][n_orb] and a flag per band, 1 = not one of the k point's window bands
    int core_n_orb = 0;
    double* core_eps = nullptr;
    unsigned char* core_out = nullptr;
    dgi::Workspace core;
    dgi::Workspace fused;    // tr G^2 partials shared by a fused forward and its adjoint
    dgi::Workspace cot;      // loss cotangents gbar [n_freq][nc][nc] + the loss sum
    // optional kernel timing (dmftgpu_enable_timing): events around the main
    // lattice kernel of each mode (FWD, DENS, VJP)
    bool timing = false;
    hipEvent_t ev0[3] = {nullptr, nullptr, nullptr}, ev1[3] = {nullptr, nullptr, nullptr};
    bool ev_used[3] = {false, false, false};
};
