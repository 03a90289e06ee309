// dmftgpu_shim.hpp -- drop-in C++ adapter keeping DMFTpack's call surface.
//
// Header-only, templated on the matrix type so it compiles against Eigen's
// MatrixXcd inside DMFTpack and against any type with the same small API
// (rows(), cols(), data() column-major complex<double>, setZero(r, c),
// operator()(i, j)).  The self-energy container needs getMatrix(n) (the
// ImgFreqFtn accessor, ImgTimeFtn.h:56).
//
// The reference reads its run parameters from globals (tight_common.h:61-79,
// TB.h:24, model.h:4); here they travel in dmftgpu_shim::Globals, which the
// integrator fills from those globals once per call site.  Every function
// gathers host data into device buffers, calls the C ABI of dmftgpu.h, and
// copies the results back in the reference's layout.  A non-zero status is
// turned into the reference's failure behaviour: message + abort (the
// reference asserts, GreenFtn.cpp:85, FourierTransform.cpp:321,440,557).
//
// MPI: with k split over ranks (knum < knum_mpiGlobal) the outputs are this
// rank's partial sums, exactly the operands of the reference's MPI_Allreduce
// calls (GreenFtn.cpp:356,368-369,633); Globals::allreduce, when set, is
// applied to them (in the same places) before they are returned.
#pragma once

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "dmftgpu.h"

namespace dmftgpu_shim {

using cplx = std::complex<double>;

struct Globals {
    int N_freq = 0;                   // tight_common.h:67
    int N_tau = 0;                    // tight_common.h:67
    double beta = 0.0;                // tight_common.h:64
    int knum = 0;                     // TB.h:24 (this rank's k points)
    int knum_mpiGlobal = 0;           // TB.h:24
    int highFreq2 = 2;                // GreenFtn.cpp:8
    double NumberOfElectron = 0.0;    // tight_common.h:64
    int method = DMFTGPU_PLAIN_F64;   // per-(k, iw) solver
    double mu_tolerance = 1e-5;       // |dmu| stop of GreenFtn_w_adjust_mu (GreenFtn.cpp:96)
    int mpi_rank = 0;                 // dmft_common.h:44; rank 0 prints and writes files (ifroot)
    bool write_files = false;         // GreenFtn_w writes Gw_loc*.dat / GwHF_loc*.dat (GreenFtn.cpp:387-421)
    // optional cross-rank SUM of a complex128 buffer (identity when unset)
    std::function<void(cplx*, size_t)> allreduce;
};

inline void die(dmftgpu_handle h, int rc, const char* where)
{
    if (rc == DMFTGPU_OK) return;
    std::fprintf(stderr, "[dmftgpu] %s failed (%d): %s\n", where, rc, dmftgpu_last_error(h));
    std::abort();
}

inline void hip_ok(hipError_t e, const char* where)
{
    if (e == hipSuccess) return;
    std::fprintf(stderr, "[dmftgpu] %s: %s\n", where, hipGetErrorString(e));
    std::abort();
}

// A device buffer of complex128 that owns its memory.
struct DevBuf {
    cplx* p = nullptr;
    size_t n = 0;
    explicit DevBuf(size_t count) : n(count) { hip_ok(hipMalloc((void**)&p, sizeof(cplx) * (count ? count : 1)), "hipMalloc"); }
    ~DevBuf() { if (p) (void)hipFree(p); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    void upload(const cplx* h) { hip_ok(hipMemcpy(p, h, sizeof(cplx) * n, hipMemcpyHostToDevice), "upload"); }
    void download(cplx* h) const { hip_ok(hipMemcpy(h, p, sizeof(cplx) * n, hipMemcpyDeviceToHost), "download"); }
};

template <class Matrix>
std::vector<cplx> pack(const std::vector<Matrix>& mats, size_t count, int rows)
{
    std::vector<cplx> out(count * rows * rows);
    for (size_t m = 0; m < count; ++m) std::memcpy(&out[m * rows * rows], mats[m].data(), sizeof(cplx) * rows * rows);
    return out;
}

// NBAND[k] may differ between k points (the energy window, downfolding.cpp:162-175).
// H(k) goes to the library at the largest NBAND: each smaller H(k) is padded with
// decoupled bands at kPadEnergy on the diagonal.  A decoupled band has G = G0
// at every frequency and a Fermi factor of exactly 0 (beta (E - mu) > 100), so
// G_loc, NumMatrix, N and the model block of densityMatDFT are those of the
// unpadded H(k).
constexpr double kPadEnergy = 1.0e4;

template <class Matrix>
int max_rows(const std::vector<Matrix>& mats, size_t count)
{
    int n = 0;
    for (size_t m = 0; m < count; ++m) n = std::max(n, (int)mats[m].rows());
    return n;
}

template <class Matrix>
std::vector<cplx> pack_hk(const std::vector<Matrix>& H_k, size_t count, int n)
{
    std::vector<cplx> out(count * n * n, cplx(0.0, 0.0));
    for (size_t k = 0; k < count; ++k) {
        const int nb = (int)H_k[k].rows();
        cplx* dst = &out[k * n * n];
        for (int j = 0; j < nb; ++j) std::memcpy(dst + (size_t)j * n, H_k[k].data() + (size_t)j * nb, sizeof(cplx) * nb);
        for (int i = nb; i < n; ++i) dst[i + (size_t)i * n] = kPadEnergy;
    }
    return out;
}

template <class SelfEnergy>
std::vector<cplx> pack_sigma(SelfEnergy& S, int n_freq, int nc)
{
    std::vector<cplx> out((size_t)n_freq * nc * nc);
    for (int n = 0; n < n_freq; ++n) {
        auto M = S.getMatrix(n);
        std::memcpy(&out[(size_t)n * nc * nc], M.data(), sizeof(cplx) * nc * nc);
    }
    return out;
}

// Keeps only the intra-cluster blocks (valenceHAMILTONIAN, GreenFtn.cpp:32-39).
inline void cluster_mask(cplx* m, int nc, int ncl, size_t count)
{
    const int npc = nc / ncl;
    for (size_t c = 0; c < count; ++c)
        for (int j = 0; j < nc; ++j)
            for (int i = 0; i < nc; ++i)
                if (i / npc != j / npc) m[c * nc * nc + i + (size_t)j * nc] = 0.0;
}

inline dmftgpu_lattice lattice_of(const Globals& g, int n, int nc, int ncl)
{
    dmftgpu_lattice L;
    L.n = n;
    L.n_corr = nc;
    L.n_cluster = ncl;
    L.n_k = g.knum;
    L.n_k_global = g.knum_mpiGlobal;
    L.n_freq = g.N_freq;
    L.high_freq = g.highFreq2;
    L.beta = g.beta;
    return L;
}

// getAsymto_moments (FourierTransform.cpp:184-465) on a device array.
inline std::vector<cplx> moments_dev(dmftgpu_handle h, int d, int n_freq, double beta, const DevBuf& fw, bool green)
{
    DevBuf mom(4 * (size_t)d * d);
    die(h, dmftgpu_moments(h, d, n_freq, beta, (const double*)fw.p, green ? 1 : 0, (double*)mom.p, nullptr),
        "getAsymto_moments");
    std::vector<cplx> out(mom.n);
    mom.download(out.data());
    return out;
}

namespace detail {
// Eigenvalues of a small Hermitian matrix (column-major): the real symmetric
// embedding [[Re A, -Im A], [Im A, Re A]] has each eigenvalue of A twice; it is
// diagonalised by cyclic Jacobi rotations (tan of the angle from the stable
// root of t^2 + 2 theta t - 1 = 0) until the off-diagonal mass is below 1e-30
// of the total.  Host only; used for the NumMatrix check of GreenFtn_w.
inline std::vector<double> herm_eigvals(const cplx* A, int n)
{
    const int m = 2 * n;
    std::vector<double> B((size_t)m * m);
    auto b = [&](int i, int j) -> double& { return B[i + (size_t)j * m]; };
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const cplx v = A[i + (size_t)j * n];
            b(i, j) = b(i + n, j + n) = v.real();
            b(i + n, j) = v.imag();
            b(i, j + n) = -v.imag();
        }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < m; ++i) (i == j ? tot : off) += b(i, j) * b(i, j);
        if (off <= 1e-30 * (tot + off)) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                if (b(p, q) == 0.0) continue;
                const double theta = (b(q, q) - b(p, p)) / (2.0 * b(p, q));
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
                for (int k = 0; k < m; ++k) {
                    const double bkp = b(k, p), bkq = b(k, q);
                    b(k, p) = c * bkp - sn * bkq;
                    b(k, q) = sn * bkp + c * bkq;
                }
                for (int k = 0; k < m; ++k) {
                    const double bpk = b(p, k), bqk = b(q, k);
                    b(p, k) = c * bpk - sn * bqk;
                    b(q, k) = sn * bpk + c * bqk;
                }
            }
    }
    std::vector<double> d(m);
    for (int i = 0; i < m; ++i) d[i] = b(i, i);
    std::sort(d.begin(), d.end());
    std::vector<double> w(n);
    for (int i = 0; i < n; ++i) w[i] = 0.5 * (d[2 * i] + d[2 * i + 1]);
    return w;
}

// The per-cluster files of GreenFtn_w (GreenFtn.cpp:387-421):
//   Gw_loc.dat<cl>, GwHF_loc.dat<cl>: "%0.5f" of w_n, then "    %e    %e" per diagonal entry
//   Gw_loc.full.dat<cl>: "%d %d %d %e    %e" per entry with |G| > 1e-8
//   GwHF_loc.full.dat<cl>: the same line for every entry
// (the solver loop reads the .full files back, dmft_main.cpp:394,398)
template <class Matrix>
void write_gw_loc_files(const Globals& g, int NumCluster, int npc, const std::vector<Matrix>& Gw,
                        const std::vector<Matrix>& GwHF)
{
    const double pi = 3.1415926535897932384626;  // dmft_common.h:43
    for (int at = 0; at < NumCluster; ++at) {
        const std::string sfx = std::to_string(at);
        std::FILE* f2 = std::fopen(("Gw_loc.dat" + sfx).c_str(), "w");
        std::FILE* f3 = std::fopen(("Gw_loc.full.dat" + sfx).c_str(), "w");
        std::FILE* h2 = std::fopen(("GwHF_loc.dat" + sfx).c_str(), "w");
        std::FILE* h3 = std::fopen(("GwHF_loc.full.dat" + sfx).c_str(), "w");
        if (!f2 || !f3 || !h2 || !h3) {
            std::fprintf(stderr, "[dmftgpu] GreenFtn_w: cannot write Gw_loc files\n");
            std::exit(1);
        }
        for (int n = 0; n < g.N_freq; ++n) {
            const double wn = pi * (2 * n + 1) / g.beta;
            std::fprintf(f2, "%0.5f", wn);
            std::fprintf(h2, "%0.5f", wn);
            for (int a = 0; a < npc; ++a) {
                const int aF = at * npc + a;
                const cplx ga = Gw[n](aF, aF), ha = GwHF[n](aF, aF);
                std::fprintf(f2, "    %e    %e", ga.real(), ga.imag());
                std::fprintf(h2, "    %e    %e", ha.real(), ha.imag());
                for (int b = 0; b < npc; ++b) {
                    const int bF = at * npc + b;
                    const cplx gab = Gw[n](aF, bF), hab = GwHF[n](aF, bF);
                    if (std::abs(gab) > 1e-8) std::fprintf(f3, "%d %d %d %e    %e\n", n, a, b, gab.real(), gab.imag());
                    std::fprintf(h3, "%d %d %d %e    %e\n", n, a, b, hab.real(), hab.imag());
                }
            }
            std::fprintf(f2, "\n");
            std::fprintf(h2, "\n");
        }
        std::fclose(f2);
        std::fclose(f3);
        std::fclose(h2);
        std::fclose(h3);
        std::printf("FILEOUT:Gw_loc.dat%d\nFILEOUT:Gw_loc.full.dat%d\n", at, at);
    }
}
}  // namespace detail

// void GreenFtn_w(int NumCluster, int NumHartrOrbit_per_cluster, vector<MatrixXcd>& H_k,
//                 ImgFreqFtn& SelfE_w, vector<MatrixXcd>& Gw, double mu, vector<MatrixXcd>& densityMatDFT)
// (GreenFtn.cpp:142).  NumMatrix (the global written at :356) is returned through num_matrix,
// with the reference's warning when an eigenvalue leaves [0, 1 + 1e-10] (:358-365).
// GwHF (the static-part Green's function of :277, which the reference only writes to
// GwHF_loc*.dat) is returned when GwHF is given or Globals::write_files is set; with
// write_files rank 0 writes the four per-cluster files of :387-421.
// densityMatDFT is filled when it is non-empty on entry (the caller pre-sizes it, TB.cpp:296-297).
template <class Matrix, class SelfEnergy>
void GreenFtn_w(dmftgpu_handle h, const Globals& g, int NumCluster, int NumHartrOrbit_per_cluster,
                std::vector<Matrix>& H_k, SelfEnergy& SelfE_w, std::vector<Matrix>& Gw, double mu,
                std::vector<Matrix>& densityMatDFT, Matrix& num_matrix, std::vector<Matrix>* GwHF = nullptr)
{
    const int nc = NumCluster * NumHartrOrbit_per_cluster;
    const int n = max_rows(H_k, g.knum);  // NBAND[k] may vary: padded to the largest
    const dmftgpu_lattice L = lattice_of(g, n, nc, NumCluster);
    const bool want_hf = GwHF != nullptr || g.write_files;
    DevBuf hk((size_t)g.knum * n * n), sig((size_t)g.N_freq * nc * nc), gl((size_t)g.N_freq * nc * nc), nm((size_t)nc * nc),
        ghf(want_hf ? (size_t)g.N_freq * nc * nc : 1);
    hk.upload(pack_hk(H_k, g.knum, n).data());
    sig.upload(pack_sigma(SelfE_w, g.N_freq, nc).data());
    std::vector<cplx> mom = moments_dev(h, nc, g.N_freq, g.beta, sig, false);  // :189-195
    DevBuf dmom(mom.size());
    dmom.upload(mom.data());
    die(h, dmftgpu_prepare(h, &L, (const double*)hk.p, (const double*)dmom.p), "prepare");
    die(h, dmftgpu_greenftn_w(h, g.method, (const double*)sig.p, mu, (double*)gl.p, want_hf ? (double*)ghf.p : nullptr,
                              (double*)nm.p, nullptr),
        "GreenFtn_w");
    std::vector<cplx> hgl(gl.n), hnm(nm.n), hhf(want_hf ? ghf.n : 0);
    gl.download(hgl.data());
    nm.download(hnm.data());
    if (want_hf) ghf.download(hhf.data());
    if (g.allreduce) {  // :356, :368-369
        g.allreduce(hnm.data(), hnm.size());
        g.allreduce(hgl.data(), hgl.size());
        if (want_hf) g.allreduce(hhf.data(), hhf.size());
    }
    for (double ev : detail::herm_eigvals(hnm.data(), nc))  // :358-365 (a warning, the run goes on)
        if ((ev < 0 || ev > 1 + 1e-10) && g.mpi_rank == 0) std::printf("[GreenFtn]: NumMatrix is unphysical, %g\n", ev);
    auto unpack = [&](const std::vector<cplx>& src, std::vector<Matrix>& dst) {
        dst.resize(g.N_freq);
        for (int w = 0; w < g.N_freq; ++w) {
            dst[w].setZero(nc, nc);
            std::memcpy(dst[w].data(), &src[(size_t)w * nc * nc], sizeof(cplx) * nc * nc);
        }
    };
    unpack(hgl, Gw);
    std::vector<Matrix> hf_local;
    std::vector<Matrix>& hf = GwHF ? *GwHF : hf_local;
    if (want_hf) unpack(hhf, hf);
    num_matrix.setZero(nc, nc);
    std::memcpy(num_matrix.data(), hnm.data(), sizeof(cplx) * nc * nc);
    if (!densityMatDFT.empty()) {  // :291-319
        DevBuf dm((size_t)g.knum * n * n);
        die(h, dmftgpu_density_matrix(h, (const double*)sig.p, mu, (double*)dm.p), "densityMatDFT");
        std::vector<cplx> hdm(dm.n);
        dm.download(hdm.data());
        for (int k = 0; k < g.knum; ++k) {  // the NBAND[k] x NBAND[k] block of each k
            const int nb = (int)H_k[k].rows();
            densityMatDFT[k].setZero(nb, nb);
            for (int j = 0; j < nb; ++j)
                std::memcpy(densityMatDFT[k].data() + (size_t)j * nb, &hdm[(size_t)k * n * n + (size_t)j * n],
                            sizeof(cplx) * nb);
        }
    }
    if (g.write_files && g.mpi_rank == 0) detail::write_gw_loc_files(g, NumCluster, NumHartrOrbit_per_cluster, Gw, hf);
}

// double GreenFtn_w_adjust_mu(vector<MatrixXcd>& H_k, ImgFreqFtn& SelfE_w, double mu, double dmu, int mu_adjust)
// (GreenFtn.cpp:66).  Sigma and m0 enter through their intra-cluster blocks, as in
// valenceHAMILTONIAN.  Newton on the analytic dN/dmu replaces the bracketing of :93-121.
// The target is Globals::NumberOfElectron, the total electron count: GiwToNele adds the
// Fermi weight of the Kohn-Sham bands outside the window (:622-631), from the globals
// KS_eigenEnergy[k][l] (l < NumOrbit) and FromValToKS[k][l0] (l0 < NBAND[k]) that arrive
// here as arguments (TB.h:7,19,31).
template <class Matrix, class SelfEnergy, class Energies>
double GreenFtn_w_adjust_mu(dmftgpu_handle h, const Globals& g, int NumCluster, int NumHartrOrbit_per_cluster,
                            std::vector<Matrix>& H_k, SelfEnergy& SelfE_w, double mu, double /*dmu*/,
                            int /*mu_adjust*/, const Energies& KS_eigenEnergy,
                            const std::vector<std::vector<int>>& FromValToKS, int NumOrbit)
{
    const int nc = NumCluster * NumHartrOrbit_per_cluster;
    const int n = max_rows(H_k, g.knum);  // NBAND[k] may vary: padded to the largest
    dmftgpu_lattice L = lattice_of(g, n, nc, NumCluster);
    DevBuf hk((size_t)g.knum * n * n), sig((size_t)g.N_freq * nc * nc);
    hk.upload(pack_hk(H_k, g.knum, n).data());
    std::vector<cplx> hs = pack_sigma(SelfE_w, g.N_freq, nc);
    sig.upload(hs.data());
    std::vector<cplx> mom = moments_dev(h, nc, g.N_freq, g.beta, sig, false);  // GreenFtn.cpp:523
    cluster_mask(hs.data(), nc, NumCluster, g.N_freq);
    cluster_mask(mom.data(), nc, NumCluster, 4);
    sig.upload(hs.data());
    DevBuf dmom(mom.size());
    dmom.upload(mom.data());
    die(h, dmftgpu_prepare(h, &L, (const double*)hk.p, (const double*)dmom.p), "prepare");
    if (NumOrbit > 0) {  // :622-631
        std::vector<double> eps((size_t)g.knum * NumOrbit);
        std::vector<int> v2k((size_t)g.knum * n, -1);  // -1: a padded band (l0 >= NBAND[k])
        for (int k = 0; k < g.knum; ++k) {
            for (int l = 0; l < NumOrbit; ++l) eps[(size_t)k * NumOrbit + l] = KS_eigenEnergy[k][l];
            const int nb = (int)H_k[k].rows();
            for (int l0 = 0; l0 < nb; ++l0) v2k[(size_t)k * n + l0] = FromValToKS[k][l0];
        }
        die(h, dmftgpu_set_core_bands(h, NumOrbit, eps.data(), v2k.data()), "GreenFtn_w_adjust_mu (core bands)");
    }
    double mu_out = mu;
    int ne = 0;
    die(h, dmftgpu_adjust_mu(h, g.method, (const double*)sig.p, mu, g.NumberOfElectron, g.mu_tolerance, &mu_out, &ne),
        "GreenFtn_w_adjust_mu");
    if (g.mpi_rank == 0) std::printf("Adjusted Chemical potential : From %2.5f to %2.5f  (RD= %+2.5f )\n", mu, mu_out, mu_out - mu);
    return mu_out;
}

// Model-space-only form (no Kohn-Sham bands outside the window: NumOrbit = NBAND and
// FromValToKS the identity, whose core-band term is zero).
template <class Matrix, class SelfEnergy>
double GreenFtn_w_adjust_mu(dmftgpu_handle h, const Globals& g, int NumCluster, int NumHartrOrbit_per_cluster,
                            std::vector<Matrix>& H_k, SelfEnergy& SelfE_w, double mu, double dmu, int mu_adjust)
{
    const std::vector<std::vector<double>> none;
    return GreenFtn_w_adjust_mu(h, g, NumCluster, NumHartrOrbit_per_cluster, H_k, SelfE_w, mu, dmu, mu_adjust, none,
                                std::vector<std::vector<int>>(), 0);
}

// void getAsymto_moments(vector<MatrixXcd>& moments, MatrixXcd* delta_w, bool greenFtnAsym) (tight_common.h:141)
template <class Matrix>
void getAsymto_moments(dmftgpu_handle h, const Globals& g, std::vector<Matrix>& moments, Matrix* delta_w,
                       bool greenFtnAsym = false)
{
    const int d = (int)delta_w[0].rows();
    DevBuf fw((size_t)g.N_freq * d * d);
    std::vector<cplx> hf((size_t)g.N_freq * d * d);
    for (int w = 0; w < g.N_freq; ++w) std::memcpy(&hf[(size_t)w * d * d], delta_w[w].data(), sizeof(cplx) * d * d);
    fw.upload(hf.data());
    std::vector<cplx> m = moments_dev(h, d, g.N_freq, g.beta, fw, greenFtnAsym);
    moments.resize(4);
    for (int q = 0; q < 4; ++q) {
        moments[q].setZero(d, d);
        std::memcpy(moments[q].data(), &m[(size_t)q * d * d], sizeof(cplx) * d * d);
    }
}

namespace detail {
template <class Matrix>
void ft_grid(dmftgpu_handle h, const Globals& g, Matrix* delta_w, Matrix* delta_t, const Matrix* moments1, bool green)
{
    const int d = (int)delta_w[0].rows();
    DevBuf fw((size_t)g.N_freq * d * d), gt((size_t)(g.N_tau + 1) * d * d);
    std::vector<cplx> hf((size_t)g.N_freq * d * d);
    for (int w = 0; w < g.N_freq; ++w) std::memcpy(&hf[(size_t)w * d * d], delta_w[w].data(), sizeof(cplx) * d * d);
    fw.upload(hf.data());
    std::vector<cplx> m = moments_dev(h, d, g.N_freq, g.beta, fw, green);
    if (moments1) {  // FourierTransform.cpp:485-486
        std::fill(m.begin(), m.begin() + (size_t)d * d, cplx(0.0));
        std::memcpy(&m[(size_t)d * d], moments1->data(), sizeof(cplx) * d * d);
    }
    DevBuf dm(m.size());
    dm.upload(m.data());
    int patched = 0;
    die(h, dmftgpu_ft_w2tau(h, d, g.N_freq, g.N_tau, g.beta, (const double*)fw.p, (const double*)dm.p,
                            (double*)gt.p, &patched), "FourierTransform");
    std::vector<cplx> ht(gt.n);
    gt.download(ht.data());
    for (int t = 0; t <= g.N_tau; ++t) {
        delta_t[t].setZero(d, d);
        std::memcpy(delta_t[t].data(), &ht[(size_t)t * d * d], sizeof(cplx) * d * d);
    }
    if (patched) std::fprintf(stderr, " Warning : positive Gt\n");  // :175
}

template <class Matrix>
void ft_point(dmftgpu_handle h, const Globals& g, Matrix* delta_w, Matrix& delta_t, double tau, const Matrix* moments1,
              bool green)
{
    const int d = (int)delta_w[0].rows();
    DevBuf fw((size_t)g.N_freq * d * d), out((size_t)d * d);
    std::vector<cplx> hf((size_t)g.N_freq * d * d);
    for (int w = 0; w < g.N_freq; ++w) std::memcpy(&hf[(size_t)w * d * d], delta_w[w].data(), sizeof(cplx) * d * d);
    fw.upload(hf.data());
    std::vector<cplx> m = moments_dev(h, d, g.N_freq, g.beta, fw, green);
    if (moments1) {
        std::fill(m.begin(), m.begin() + (size_t)d * d, cplx(0.0));
        std::memcpy(&m[(size_t)d * d], moments1->data(), sizeof(cplx) * d * d);
    }
    DevBuf dm(m.size());
    dm.upload(m.data());
    die(h, dmftgpu_ft_w2tau_points(h, d, g.N_freq, g.beta, (const double*)fw.p, (const double*)dm.p, 1, &tau,
                                   (double*)out.p), "FourierTransform(tau)");
    delta_t.setZero(d, d);
    out.download((cplx*)delta_t.data());
}
}  // namespace detail

// The four FourierTransform overloads (tight_common.h:142-145, FourierTransform.cpp:470-512).
template <class Matrix>
void FourierTransform(dmftgpu_handle h, const Globals& g, Matrix* delta_w, Matrix* delta_t, bool greenFtnAsym = false)
{
    detail::ft_grid(h, g, delta_w, delta_t, (const Matrix*)nullptr, greenFtnAsym);
}
template <class Matrix>
void FourierTransform(dmftgpu_handle h, const Globals& g, Matrix* delta_w, Matrix* delta_t, const Matrix& moments1,
                      bool greenFtnAsym = false)
{
    detail::ft_grid(h, g, delta_w, delta_t, &moments1, greenFtnAsym);
}
template <class Matrix>
void FourierTransform(dmftgpu_handle h, const Globals& g, Matrix* delta_w, Matrix& delta_t, double tau,
                      bool greenFtnAsym = false)
{
    detail::ft_point(h, g, delta_w, delta_t, tau, (const Matrix*)nullptr, greenFtnAsym);
}
template <class Matrix>
void FourierTransform(dmftgpu_handle h, const Globals& g, Matrix* delta_w, Matrix& delta_t, double tau,
                      const Matrix& moments1, bool greenFtnAsym = false)
{
    detail::ft_point(h, g, delta_w, delta_t, tau, &moments1, greenFtnAsym);
}

// void FT_t_to_w(MatrixXcd* delta_w, MatrixXcd* delta_t, int N_freq) (tight_common.h:146)
template <class Matrix>
void FT_t_to_w(dmftgpu_handle h, const Globals& g, Matrix* delta_w, Matrix* delta_t, int N_freq)
{
    const int d = (int)delta_t[0].rows();
    DevBuf gt((size_t)(g.N_tau + 1) * d * d), fw((size_t)N_freq * d * d);
    std::vector<cplx> ht(gt.n);
    for (int t = 0; t <= g.N_tau; ++t) std::memcpy(&ht[(size_t)t * d * d], delta_t[t].data(), sizeof(cplx) * d * d);
    gt.upload(ht.data());
    int viol = 0;
    die(h, dmftgpu_ft_tau2w(h, d, N_freq, g.N_tau, g.beta, (const double*)gt.p, (double*)fw.p, &viol), "FT_t_to_w");
    if (viol) {  // the reference asserts here (:557)
        std::fprintf(stderr, "[dmftgpu] FT_t_to_w: %d frequencies violate causality\n", viol);
        std::abort();
    }
    std::vector<cplx> hf(fw.n);
    fw.download(hf.data());
    for (int n = 0; n < N_freq; ++n) {
        delta_w[n].setZero(d, d);
        std::memcpy(delta_w[n].data(), &hf[(size_t)n * d * d], sizeof(cplx) * d * d);
    }
}

// void Construct_hyb_delta(int impurityDim, std::vector<int> impurityOrbit, ImgFreqFtn& SelfE_w,
//                          std::vector<MatrixXcd> Gw, double mu, ImgFreqFtn& weiss_fieldTB, int atom,
//                          MatrixXcd& SolverBasis)  (Construct_hyb_delta.cpp:47)
// The global impurity_site_Hamiltonian the reference reads becomes an argument;
// weiss_fieldTB needs setValueSubMat(n, atom, h1, h2, value) (ImgTimeFtn.h).
template <class Matrix, class SelfEnergy, class Weiss>
void Construct_hyb_delta(dmftgpu_handle h, const Globals& g, int impurityDim, const std::vector<int>& impurityOrbit,
                         SelfEnergy& SelfE_w, const std::vector<Matrix>& Gw, double mu, Weiss& weiss_fieldTB, int atom,
                         const Matrix& impurity_site_Hamiltonian)
{
    if (impurityDim == 0) return;
    const int nc = (int)Gw[0].rows();
    DevBuf gl((size_t)g.N_freq * nc * nc), sig((size_t)g.N_freq * nc * nc), hi((size_t)nc * nc),
        dl((size_t)g.N_freq * impurityDim * impurityDim);
    gl.upload(pack(Gw, g.N_freq, nc).data());
    sig.upload(pack_sigma(SelfE_w, g.N_freq, nc).data());
    std::vector<cplx> hh((size_t)nc * nc);
    for (int j = 0; j < nc; ++j)
        for (int i = 0; i < nc; ++i) hh[i + (size_t)j * nc] = impurity_site_Hamiltonian(i, j);
    hi.upload(hh.data());
    die(h, dmftgpu_hyb_delta(h, nc, g.N_freq, g.beta, mu, (const double*)gl.p, (const double*)sig.p,
                             (const double*)hi.p, impurityDim, impurityOrbit.data(), (double*)dl.p),
        "Construct_hyb_delta");
    std::vector<cplx> hd(dl.n);
    dl.download(hd.data());
    const size_t d2 = (size_t)impurityDim * impurityDim;
    for (int n = 0; n < g.N_freq; ++n)
        for (int h2 = 0; h2 < impurityDim; ++h2)
            for (int h1 = 0; h1 < impurityDim; ++h1)
                weiss_fieldTB.setValueSubMat(n, atom, h1, h2, hd[n * d2 + h1 + (size_t)h2 * impurityDim]);
}

// void write_hyb_t(ImgFreqFtn& weiss_field, MatrixXcd* delta_t, int atom)  (Construct_hyb_delta.cpp:126)
// Delta(iw) of the atom's NSpinOrbit_per_atom orbitals (getValue(n, h1, h2)) -> Delta(tau)
// through the moment fit and transform (FourierTransform(delta_w, delta_t, false), :136).
template <class Matrix, class Weiss>
void write_hyb_t(dmftgpu_handle h, const Globals& g, Weiss& weiss_field, Matrix* delta_t, int NSpinOrbit_per_atom)
{
    std::vector<Matrix> delta_w(g.N_freq);
    for (int n = 0; n < g.N_freq; ++n) {
        delta_w[n].setZero(NSpinOrbit_per_atom, NSpinOrbit_per_atom);
        for (int h2 = 0; h2 < NSpinOrbit_per_atom; ++h2)
            for (int h1 = 0; h1 < NSpinOrbit_per_atom; ++h1) delta_w[n](h1, h2) = weiss_field.getValue(n, h1, h2);
    }
    FourierTransform(h, g, delta_w.data(), delta_t, false);
}

// Hopping list of an Hk.HWR file (read_HR, read_TB_parm.cpp:6-76): the header
// lines "Norbital of Atom = N", then "R1 R2 R3 atom1 atom2 i j Re Im" lines
// with 1-based atoms and orbitals.  Orbital indices are returned 0-based and
// global: (orbitals of the atoms before atom1) + i - 1.
struct HoppingList {
    std::vector<int> accumulated;  // accumulated_Num_SpinOrbital, size atoms + 1
    std::vector<int> idx;          // [n_hop][5]: R1 R2 R3 i j
    std::vector<cplx> t;           // [n_hop]
    int n_orbital() const { return accumulated.empty() ? 0 : accumulated.back(); }
};

inline HoppingList read_HR(const char* path)
{
    HoppingList out;
    std::FILE* f = std::fopen(path, "r");
    if (!f) {
        std::fprintf(stderr, "Cannot open Hopping data file...\n");  // the reference exits here (:9-11)
        std::exit(1);
    }
    out.accumulated.push_back(0);
    char line[512];
    while (std::fgets(line, sizeof(line), f)) {
        int no = 0;
        if (std::sscanf(line, " Norbital of Atom = %d", &no) == 1) {
            out.accumulated.push_back(out.accumulated.back() + no);
            continue;
        }
        int r[3], a1, a2, i, j;
        double re, im;
        if (std::sscanf(line, "%d %d %d %d %d %d %d %lf %lf", &r[0], &r[1], &r[2], &a1, &a2, &i, &j, &re, &im) != 9)
            continue;
        const int i0 = out.accumulated[a1 - 1] + i - 1, j0 = out.accumulated[a2 - 1] + j - 1;
        out.idx.insert(out.idx.end(), {r[0], r[1], r[2], i0, j0});
        out.t.emplace_back(re, im);
    }
    std::fclose(f);
    return out;
}

// H_k_inModelSpace[k] = sum_R H(R) exp(-i k.a.R)  (ConstructModelHamiltonian.cpp:58-77)
// kphase [knum][3] holds kmesh[k][axis] * a_axis.
template <class Matrix>
void Construct_Hk(dmftgpu_handle h, const HoppingList& hr, const std::vector<double>& kphase,
                  std::vector<Matrix>& H_k)
{
    const int n = hr.n_orbital();
    const int nk = (int)(kphase.size() / 3);
    DevBuf hk((size_t)nk * n * n);
    double* dk = nullptr;
    hip_ok(hipMalloc((void**)&dk, sizeof(double) * kphase.size()), "hipMalloc");
    hip_ok(hipMemcpy(dk, kphase.data(), sizeof(double) * kphase.size(), hipMemcpyHostToDevice), "upload");
    const int rc = dmftgpu_hk_from_hr(h, n, (int)hr.t.size(), hr.idx.data(), (const double*)hr.t.data(), nk, dk,
                                      (double*)hk.p);
    (void)hipFree(dk);
    die(h, rc, "Construct_Hk");
    std::vector<cplx> hh(hk.n);
    hk.download(hh.data());
    H_k.resize(nk);
    for (int k = 0; k < nk; ++k) {
        H_k[k].setZero(n, n);
        std::memcpy(H_k[k].data(), &hh[(size_t)k * n * n], sizeof(cplx) * n * n);
    }
}

// The band loop of qs_spectra (qs_spectra.cpp:197-221) on a k path whose H(k)
// the handle holds (dmftgpu_prepare with n_k = the path's length and zero
// moments).  retarded_GreenFtn2 (GreenFtn.cpp:425-466) upfolds G^R to the full
// DFT basis, diag 1/(z - eps_b) over all NumOrbit KS bands plus
// C^H (G - G0) C on the window (:450-458), so its trace is
//   tr G^R_full = sum_b 1/(z - eps_kb) + tr G(k, z) - tr G0(k, z),  z = E_n + mu + i eta,
// G0 the model inverse without Sigma (DF_CorrBase drops out of the trace).
// tr G0 comes from a second call with Sigma = 0.
//   spectralWeight[k][n]            = -Im tr G^R_full / pi                    (:209-213)
//   partial_spectralWeight[k][n][i] = -Im G^R(i, i) / pi, i < n_corr          (:215-220)
// sigma_r: Sigma^R(E_n) [n_e] n_corr x n_corr; KS_eigenEnergy[k][b], b < NumOrbit.
template <class Matrix, class Energies>
void qs_band_spectra(dmftgpu_handle h, int method, int n_corr, int knum, int n_e, double E0, double dE,
                     double infinitesimal, double mu, const std::vector<Matrix>& sigma_r,
                     const Energies& KS_eigenEnergy, int NumOrbit, std::vector<std::vector<double>>& spectralWeight,
                     std::vector<std::vector<std::vector<double>>>& partial_spectralWeight)
{
    const int nc = n_corr;
    DevBuf sig((size_t)n_e * nc * nc), zero((size_t)n_e * nc * nc), gk((size_t)knum * n_e * nc * nc),
        tk((size_t)knum * n_e), t0((size_t)knum * n_e);
    sig.upload(pack(sigma_r, n_e, nc).data());
    std::vector<cplx> z0(zero.n, cplx(0.0, 0.0));
    zero.upload(z0.data());
    die(h, dmftgpu_greenftn_real_k(h, method, n_e, E0, dE, infinitesimal, (const double*)sig.p, mu, (double*)gk.p,
                                   (double*)tk.p), "qs_spectra (G)");
    die(h, dmftgpu_greenftn_real_k(h, method, n_e, E0, dE, infinitesimal, (const double*)zero.p, mu, nullptr,
                                   (double*)t0.p), "qs_spectra (G0)");
    std::vector<cplx> hg(gk.n), ht(tk.n), h0(t0.n);
    gk.download(hg.data());
    tk.download(ht.data());
    t0.download(h0.data());
    const double pi = 3.1415926535897932384626;  // dmft_common.h:43
    spectralWeight.assign(knum, std::vector<double>(n_e, 0.0));
    partial_spectralWeight.assign(knum, std::vector<std::vector<double>>(n_e, std::vector<double>(nc, 0.0)));
    for (int k = 0; k < knum; ++k)
        for (int n = 0; n < n_e; ++n) {
            const cplx z(E0 + n * dE + mu, infinitesimal);
            cplx tr = ht[(size_t)k * n_e + n] - h0[(size_t)k * n_e + n];
            for (int b = 0; b < NumOrbit; ++b) tr += 1.0 / (z - KS_eigenEnergy[k][b]);
            spectralWeight[k][n] = -tr.imag() / pi;
            for (int i = 0; i < nc; ++i)
                partial_spectralWeight[k][n][i] = -hg[((size_t)k * n_e + n) * nc * nc + i + (size_t)i * nc].imag() / pi;
        }
}

// upfolding_density (upfolding.cpp:38-139).  densityMatDFT[k] is taken to the
// DFT basis in place (upfolding_k, :10-35), as the reference leaves it; the
// three R sums come back in H_Rindex row order (DM_DFT_DMFT.HWR,
// dual_DM_direct.HWR, direct_DM_direct.HWR values, :89-137), after
// Globals::allreduce (:80-82).  Printing them needs the caller's
// orbital-to-atom tables and stays with the caller.  The reference's globals
// arrive as arguments: DF_CorrBase, KS_eigenEnergy (anything indexable as
// [k][band]: Eigen::VectorXd*), FromValToKS, and kphase [knum][3] =
// kmesh[k][axis] * a_axis.  An empty S_overlap means S = I.
template <class Matrix, class IndexMatrix, class Energies>
void upfolding_density(dmftgpu_handle h, const Globals& g, std::vector<Matrix>& densityMatDFT,
                       std::vector<Matrix>& KS_eigenVectors_orthoBasis, const IndexMatrix& H_Rindex, double mu,
                       std::vector<Matrix>& S_overlap, std::vector<Matrix>& transformMatrix_k,
                       const std::vector<Matrix>& DF_CorrBase, const Energies& KS_eigenEnergy,
                       const std::vector<std::vector<int>>& FromValToKS, const std::vector<double>& kphase,
                       std::vector<cplx> (&dmr)[3])
{
    const int nk = g.knum;
    // NBAND[k] may vary: the model blocks are padded to the largest, the padded
    // bands marked -1 in the band map (zero density, identity basis block)
    const int n = max_rows(densityMatDFT, nk);
    const int no = (int)KS_eigenVectors_orthoBasis[0].rows();
    const int nr = (int)H_Rindex.rows();
    DevBuf X((size_t)nk * n * n), Cb((size_t)nk * n * n), T((size_t)nk * no * no), Ko((size_t)nk * no * no),
        S(S_overlap.empty() ? 1 : (size_t)nk * no * no), D((size_t)nk * no * no), R(3 * (size_t)(nr > 0 ? nr : 1));
    {
        std::vector<cplx> hx((size_t)nk * n * n, cplx(0.0, 0.0)), hc((size_t)nk * n * n, cplx(0.0, 0.0));
        for (int k = 0; k < nk; ++k) {
            const int nb = (int)densityMatDFT[k].rows();
            for (int j = 0; j < nb; ++j) {
                std::memcpy(&hx[(size_t)k * n * n + (size_t)j * n], densityMatDFT[k].data() + (size_t)j * nb,
                            sizeof(cplx) * nb);
                std::memcpy(&hc[(size_t)k * n * n + (size_t)j * n], DF_CorrBase[k].data() + (size_t)j * nb,
                            sizeof(cplx) * nb);
            }
            for (int i = nb; i < n; ++i) hc[(size_t)k * n * n + i + (size_t)i * n] = 1.0;
        }
        X.upload(hx.data());
        Cb.upload(hc.data());
    }
    T.upload(pack(transformMatrix_k, nk, no).data());
    Ko.upload(pack(KS_eigenVectors_orthoBasis, nk, no).data());
    if (!S_overlap.empty()) S.upload(pack(S_overlap, nk, no).data());
    std::vector<double> eps((size_t)nk * no), kp(kphase.begin(), kphase.end());
    std::vector<int> v((size_t)nk * n), ridx(5 * (size_t)(nr > 0 ? nr : 1));
    for (int k = 0; k < nk; ++k) {
        for (int b = 0; b < no; ++b) eps[(size_t)k * no + b] = KS_eigenEnergy[k][b];
        const int nb = (int)densityMatDFT[k].rows();
        for (int i = 0; i < n; ++i) v[(size_t)k * n + i] = i < nb ? FromValToKS[k][i] : -1;
    }
    for (int r = 0; r < nr; ++r)
        for (int c = 0; c < 5; ++c) ridx[5 * (size_t)r + c] = H_Rindex(r, c);
    double *de = nullptr, *dk = nullptr;
    hip_ok(hipMalloc((void**)&de, sizeof(double) * eps.size()), "hipMalloc");
    hip_ok(hipMalloc((void**)&dk, sizeof(double) * kp.size()), "hipMalloc");
    hip_ok(hipMemcpy(de, eps.data(), sizeof(double) * eps.size(), hipMemcpyHostToDevice), "upload");
    hip_ok(hipMemcpy(dk, kp.data(), sizeof(double) * kp.size(), hipMemcpyHostToDevice), "upload");
    const int rc = dmftgpu_upfold_density(h, n, no, nk, g.knum_mpiGlobal, g.beta, mu, (const double*)X.p,
                                          (const double*)Cb.p, de, v.data(), (const double*)T.p, (const double*)Ko.p,
                                          S_overlap.empty() ? nullptr : (const double*)S.p, nr, ridx.data(), dk,
                                          (double*)D.p, nr > 0 ? (double*)R.p : nullptr);
    (void)hipFree(de);
    (void)hipFree(dk);
    die(h, rc, "upfolding_density");
    std::vector<cplx> hd(D.n), hr(R.n);
    D.download(hd.data());
    R.download(hr.data());
    for (int k = 0; k < nk; ++k) {
        densityMatDFT[k].setZero(no, no);
        std::memcpy(densityMatDFT[k].data(), &hd[(size_t)k * no * no], sizeof(cplx) * no * no);
    }
    for (int x = 0; x < 3; ++x) {
        dmr[x].assign(hr.begin() + (size_t)x * nr, hr.begin() + (size_t)(x + 1) * nr);
        if (g.allreduce && nr > 0) g.allreduce(dmr[x].data(), dmr[x].size());
    }
}

}  // namespace dmftgpu_shim
