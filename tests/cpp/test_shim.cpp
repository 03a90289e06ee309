// Drives include/dmftgpu_shim.hpp the way DMFTpack would (reference call
// surface, a column-major matrix type standing in for Eigen::MatrixXcd) and
// checks the results against the CPU oracle linked in as the checker.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <unistd.h>
#include <vector>

#include "dmftgpu.h"
#include "dmftgpu_shim.hpp"

using cplx = std::complex<double>;

struct Mat {  // the subset of Eigen::MatrixXcd the adapter uses
    int r = 0, c = 0;
    std::vector<cplx> v;
    int rows() const { return r; }
    int cols() const { return c; }
    cplx* data() { return v.data(); }
    const cplx* data() const { return v.data(); }
    void setZero(int rr, int cc) { r = rr; c = cc; v.assign((size_t)rr * cc, cplx(0.0)); }
    cplx& operator()(int i, int j) { return v[i + (size_t)j * r]; }
    const cplx& operator()(int i, int j) const { return v[i + (size_t)j * r]; }
};

struct Sigma {  // the subset of ImgFreqFtn the adapter uses
    std::vector<Mat> w;
    Mat getMatrix(int n) { return w[n]; }
    cplx getValue(int n, int i, int j) { return w[n](i, j); }
    void setValueSubMat(int n, int /*atom*/, int i, int j, cplx v) { w[n](i, j) = v; }
};

extern "C" {
struct orc_lattice { int n, n_corr, n_cluster, n_k, n_k_global, n_freq, high_freq; double beta; };
int orc_greenftn_w(const orc_lattice*, const cplx*, const cplx*, const cplx*, double, cplx*, cplx*, cplx*, cplx*, int);
int orc_asymto_moments(int, int, double, const cplx*, int, cplx*, int*);
int orc_ft_w2tau(int, int, int, double, const cplx*, const cplx*, cplx*, int*, int);
int orc_ft_tau2w(int, int, int, double, const cplx*, cplx*, int*);
int orc_ft_tau(int, int, double, const cplx*, const cplx*, double, cplx*);
int orc_hyb_delta(int, int, double, double, const cplx*, const cplx*, const cplx*, int, const int*, cplx*);
void orc_hk_from_hr(int, int, const int*, const cplx*, int, const double*, cplx*);
int orc_upfold_density(int, int, int, int, double, double, const cplx*, const cplx*, const double*, const int*,
                       const cplx*, const cplx*, const cplx*, int, const int*, const double*, cplx*, cplx*);
int orc_greenftn_real(const orc_lattice*, const cplx*, int, double, double, double, const cplx*, double, cplx*, cplx*,
                      int);
struct orc_core { int n_orb; const double* eps; const int* val_to_ks; };
double orc_giw_to_nele(const orc_lattice*, const cplx*, const cplx*, const cplx*, double, const orc_core*, int);
int orc_density(const orc_lattice*, const cplx*, const cplx*, const cplx*, double, double*, double*, const orc_core*,
                int);
}

struct IndexMat {  // the subset of Eigen::MatrixXi (H_Rindex) the adapter uses
    std::vector<int> v;  // row-major [rows][5]
    int rows() const { return (int)(v.size() / 5); }
    int operator()(int r, int c) const { return v[5 * (size_t)r + c]; }
};

static unsigned long long rng_state = 12345;
static double urand()
{
    rng_state = rng_state * 6364136223846793005ULL + 1442695040888963407ULL;
    return ((rng_state >> 11) * (1.0 / 9007199254740992.0)) - 0.5;
}

static int failures = 0;
static void check(const char* what, double err, double tol)
{
    std::printf("%-28s err %.3e (tol %.0e) %s\n", what, err, tol, err <= tol ? "ok" : "FAIL");
    if (!(err <= tol)) ++failures;
}

static double relerr(const cplx* a, const cplx* b, size_t n)
{
    double num = 0, den = 0;
    for (size_t i = 0; i < n; ++i) {
        num = std::fmax(num, std::abs(a[i] - b[i]));
        den = std::fmax(den, std::abs(b[i]));
    }
    return num / (den > 0 ? den : 1.0);
}

int main()
{
    const int n = 6, ncl = 2, npc = 2, nc = ncl * npc, nk = 8, nf = 48, ntau = 64;
    const double beta = 10.0, mu = 0.15, PI = 3.1415926535897932384626;
    dmftgpu_shim::Globals g;
    g.N_freq = nf;
    g.N_tau = ntau;
    g.beta = beta;
    g.knum = nk;
    g.knum_mpiGlobal = nk;
    g.NumberOfElectron = 2.5;
    // H(k): random Hermitian; Sigma: causal poles, block-diagonal per cluster
    std::vector<Mat> Hk(nk);
    for (auto& H : Hk) {
        H.setZero(n, n);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i <= j; ++i) {
                cplx v(urand(), i == j ? 0.0 : urand());
                H(i, j) = v;
                H(j, i) = std::conj(v);
            }
    }
    Sigma S;
    S.w.resize(nf);
    std::vector<double> amp(nc), pole(nc);
    for (int a = 0; a < nc; ++a) { amp[a] = 0.5 + urand(); pole[a] = 2 * urand(); }
    for (int w = 0; w < nf; ++w) {
        S.w[w].setZero(nc, nc);
        cplx z(0.0, PI * (2 * w + 1) / beta);
        for (int a = 0; a < nc; ++a) S.w[w](a, a) = 0.2 + amp[a] / (z - pole[a]);
        S.w[w](0, 1) = S.w[w](1, 0) = 0.1 / (z - 0.3);
    }
    dmftgpu_handle h;
    if (dmftgpu_create(&h, nullptr) != DMFTGPU_OK) { std::printf("no device\n"); return 2; }

    std::vector<Mat> Gw, dens(nk);
    Mat num;
    dmftgpu_shim::GreenFtn_w(h, g, ncl, npc, Hk, S, Gw, mu, dens, num);

    // oracle with the moments the adapter computed
    std::vector<Mat> mom;
    dmftgpu_shim::getAsymto_moments(h, g, mom, S.w.data(), false);
    std::vector<cplx> hk((size_t)nk * n * n), sg((size_t)nf * nc * nc), mm(4 * (size_t)nc * nc);
    for (int k = 0; k < nk; ++k) std::copy(Hk[k].v.begin(), Hk[k].v.end(), hk.begin() + (size_t)k * n * n);
    for (int w = 0; w < nf; ++w) std::copy(S.w[w].v.begin(), S.w[w].v.end(), sg.begin() + (size_t)w * nc * nc);
    for (int q = 0; q < 4; ++q) std::copy(mom[q].v.begin(), mom[q].v.end(), mm.begin() + (size_t)q * nc * nc);
    orc_lattice L{n, nc, ncl, nk, nk, nf, 2, beta};
    std::vector<cplx> gl((size_t)nf * nc * nc), ghf(gl.size()), nm((size_t)nc * nc), dm((size_t)nk * n * n);
    orc_greenftn_w(&L, hk.data(), sg.data(), mm.data(), mu, gl.data(), ghf.data(), nm.data(), dm.data(), 1);
    std::vector<cplx> got((size_t)nf * nc * nc), gotd((size_t)nk * n * n);
    for (int w = 0; w < nf; ++w) std::copy(Gw[w].v.begin(), Gw[w].v.end(), got.begin() + (size_t)w * nc * nc);
    for (int k = 0; k < nk; ++k) std::copy(dens[k].v.begin(), dens[k].v.end(), gotd.begin() + (size_t)k * n * n);
    check("GreenFtn_w Gw", relerr(got.data(), gl.data(), got.size()), 1e-12);
    check("GreenFtn_w NumMatrix", relerr(num.v.data(), nm.data(), nm.size()), 1e-12);
    check("GreenFtn_w densityMatDFT", relerr(gotd.data(), dm.data(), dm.size()), 1e-12);
    // the moments against the oracle's own fit
    std::vector<cplx> om(mm.size());
    orc_asymto_moments(nc, nf, beta, sg.data(), 0, om.data(), nullptr);
    check("getAsymto_moments", relerr(mm.data(), om.data(), om.size()), 1e-12);

    // NBAND[k] varying with k: the first half of the k points keep 5 of the 6 bands
    // (the last band and its couplings dropped); the adapter pads them, and the
    // oracle runs each band count on its own k points with the global k count
    {
        std::vector<Mat> Hr(nk);
        for (int k = 0; k < nk; ++k) {
            const int nb = k < nk / 2 ? n - 1 : n;
            Hr[k].setZero(nb, nb);
            for (int j = 0; j < nb; ++j)
                for (int i = 0; i < nb; ++i) Hr[k](i, j) = Hk[k](i, j);
        }
        std::vector<Mat> Gr, densr(nk);
        Mat numr;
        dmftgpu_shim::GreenFtn_w(h, g, ncl, npc, Hr, S, Gr, mu, densr, numr);
        std::vector<cplx> glr((size_t)nf * nc * nc, 0.0), nmr((size_t)nc * nc, 0.0);
        double dens_err = 0.0, dens_scale = 0.0;
        for (int part = 0; part < 2; ++part) {
            const int nb = part == 0 ? n - 1 : n, k0 = part == 0 ? 0 : nk / 2, kn = nk / 2;
            std::vector<cplx> hh((size_t)kn * nb * nb), g1(glr.size()), h1(glr.size()), n1(nmr.size()),
                d1((size_t)kn * nb * nb);
            for (int k = 0; k < kn; ++k) std::copy(Hr[k0 + k].v.begin(), Hr[k0 + k].v.end(), hh.begin() + (size_t)k * nb * nb);
            orc_lattice Lp{nb, nc, ncl, kn, nk, nf, 2, beta};
            orc_greenftn_w(&Lp, hh.data(), sg.data(), mm.data(), mu, g1.data(), h1.data(), n1.data(), d1.data(), 1);
            for (size_t e = 0; e < glr.size(); ++e) glr[e] += g1[e];
            for (size_t e = 0; e < nmr.size(); ++e) nmr[e] += n1[e];
            for (int k = 0; k < kn; ++k) {
                if (densr[k0 + k].rows() != nb) dens_err = 1.0;
                for (int e = 0; e < nb * nb; ++e) {
                    dens_err = std::fmax(dens_err, std::abs(densr[k0 + k].v[e] - d1[(size_t)k * nb * nb + e]));
                    dens_scale = std::fmax(dens_scale, std::abs(d1[(size_t)k * nb * nb + e]));
                }
            }
        }
        std::vector<cplx> gotr((size_t)nf * nc * nc);
        for (int w = 0; w < nf; ++w) std::copy(Gr[w].v.begin(), Gr[w].v.end(), gotr.begin() + (size_t)w * nc * nc);
        check("ragged NBAND Gw", relerr(gotr.data(), glr.data(), glr.size()), 1e-12);
        check("ragged NBAND NumMatrix", relerr(numr.v.data(), nmr.data(), nmr.size()), 1e-12);
        check("ragged NBAND densityMatDFT", dens_err / (dens_scale > 0 ? dens_scale : 1.0), 1e-12);
    }

    // GwHF (GreenFtn.cpp:277) and the per-cluster files of :387-421
    {
        std::vector<Mat> Gw2, hf, none;
        Mat num2;
        char tmpl[] = "/tmp/dmftgpu_shim_XXXXXX";
        const char* dir = mkdtemp(tmpl);
        char cwd[4096];
        if (!dir || !getcwd(cwd, sizeof(cwd)) || chdir(dir) != 0) { std::printf("no temp dir\n"); return 2; }
        dmftgpu_shim::Globals gf = g;
        gf.write_files = true;
        dmftgpu_shim::GreenFtn_w(h, gf, ncl, npc, Hk, S, Gw2, mu, none, num2, &hf);
        std::vector<cplx> ghf_got((size_t)nf * nc * nc);
        for (int w = 0; w < nf; ++w) std::copy(hf[w].v.begin(), hf[w].v.end(), ghf_got.begin() + (size_t)w * nc * nc);
        check("GreenFtn_w GwHF", relerr(ghf_got.data(), ghf.data(), ghf.size()), 1e-12);
        // GwHF_loc.full.dat<cl>: "n h1 h2 re im" for every entry of the cluster block, %e (7 digits)
        double ferr = 0.0;
        int lines = 0;
        for (int at = 0; at < ncl; ++at) {
            std::FILE* f = std::fopen(("GwHF_loc.full.dat" + std::to_string(at)).c_str(), "r");
            int w, a, b;
            double re, im;
            while (f && std::fscanf(f, "%d %d %d %lf %lf", &w, &a, &b, &re, &im) == 5) {
                const cplx ref = ghf[(size_t)w * nc * nc + (at * npc + a) + (size_t)(at * npc + b) * nc];
                ferr = std::fmax(ferr, std::abs(cplx(re, im) - ref) / std::fmax(std::abs(ref), 1e-300));
                ++lines;
            }
            if (f) std::fclose(f);
            std::FILE* q = std::fopen(("Gw_loc.dat" + std::to_string(at)).c_str(), "r");
            if (!q) ferr = 1.0;
            else std::fclose(q);
        }
        if (lines != ncl * nf * npc * npc) ferr = 1.0;
        check("GwHF_loc.full.dat values", ferr, 1e-6);
        if (chdir(cwd) != 0) ++failures;
    }

    // mu search: N(mu) = NumberOfElectron, checked with the oracle's GiwToNele at mu_out
    double mu_new = dmftgpu_shim::GreenFtn_w_adjust_mu(h, g, ncl, npc, Hk, S, 0.0, 1.0, 1);
    {
        const double Nn = orc_giw_to_nele(&L, hk.data(), sg.data(), mm.data(), mu_new, nullptr, 1);
        double Nd = 0, dN = 0;
        orc_density(&L, hk.data(), sg.data(), mm.data(), mu_new, &Nd, &dN, nullptr, 1);
        std::printf("adjusted mu %.10f, N(mu) %.12f, dN/dmu %.6f\n", mu_new, Nn, dN);
        check("adjust_mu N(mu_out) - target", std::fabs(Nn - g.NumberOfElectron), 1e-5 * dN);
    }
    // with Kohn-Sham bands outside the window (GiwToNele's core-band term,
    // GreenFtn.cpp:622-631): NumOrbit = n + 3, two bands below the window
    // (KS bands 0, 1) and one above; the target is the total count
    {
        const int no = n + 3;
        std::vector<std::vector<double>> ks(nk, std::vector<double>(no));
        std::vector<std::vector<int>> v2k(nk, std::vector<int>(n));
        std::vector<double> eps_flat((size_t)nk * no);
        std::vector<int> v_flat((size_t)nk * n);
        for (int k = 0; k < nk; ++k) {
            ks[k][0] = -1.0 + 0.3 * urand();
            ks[k][1] = -0.6 + 0.3 * urand();
            for (int l = 2; l < no - 1; ++l) ks[k][l] = urand();
            ks[k][no - 1] = 0.9 + 0.3 * urand();
            for (int l0 = 0; l0 < n; ++l0) v2k[k][l0] = 2 + (l0 + k) % n;
            for (int l = 0; l < no; ++l) eps_flat[(size_t)k * no + l] = ks[k][l];
            for (int l0 = 0; l0 < n; ++l0) v_flat[(size_t)k * n + l0] = v2k[k][l0];
        }
        dmftgpu_shim::Globals gc = g;
        gc.NumberOfElectron = 2.0 + 2.5;
        const double mu_c = dmftgpu_shim::GreenFtn_w_adjust_mu(h, gc, ncl, npc, Hk, S, 0.0, 1.0, 1, ks, v2k, no);
        orc_core core{no, eps_flat.data(), v_flat.data()};
        const double Nn = orc_giw_to_nele(&L, hk.data(), sg.data(), mm.data(), mu_c, &core, 1);
        double Nd = 0, dN = 0;
        orc_density(&L, hk.data(), sg.data(), mm.data(), mu_c, &Nd, &dN, &core, 1);
        std::printf("adjusted mu with core bands %.10f, N(mu) %.12f, dN/dmu %.6f\n", mu_c, Nn, dN);
        check("core bands N(mu_out) - target", std::fabs(Nn - gc.NumberOfElectron), 1e-5 * dN);
    }

    // FourierTransform (G-like input) and FT_t_to_w round trip
    std::vector<Mat> gt(ntau + 1), gw2(nf);
    dmftgpu_shim::FourierTransform(h, g, Gw.data(), gt.data(), true);
    std::vector<cplx> ogt((size_t)(ntau + 1) * nc * nc), gm(4 * (size_t)nc * nc);
    orc_asymto_moments(nc, nf, beta, got.data(), 1, gm.data(), nullptr);
    int patched = 0;
    orc_ft_w2tau(nc, nf, ntau, beta, got.data(), gm.data(), ogt.data(), &patched, 1);
    std::vector<cplx> ggt(ogt.size());
    for (int t = 0; t <= ntau; ++t) std::copy(gt[t].v.begin(), gt[t].v.end(), ggt.begin() + (size_t)t * nc * nc);
    check("FourierTransform", relerr(ggt.data(), ogt.data(), ogt.size()), 1e-12);
    Mat one;
    dmftgpu_shim::FourierTransform(h, g, Gw.data(), one, 3 * beta / ntau, true);
    check("FourierTransform(tau)", relerr(one.v.data(), &ggt[(size_t)3 * nc * nc], (size_t)nc * nc), 1e-12);
    // the single-tau form at a tau off the grid, against the oracle's FourierTransform_tau
    {
        const double tau = 0.3719 * beta;
        dmftgpu_shim::FourierTransform(h, g, Gw.data(), one, tau, true);
        std::vector<cplx> o1((size_t)nc * nc);
        orc_ft_tau(nc, nf, beta, got.data(), gm.data(), tau, o1.data());
        check("FourierTransform(tau) off grid", relerr(one.v.data(), o1.data(), o1.size()), 1e-12);
    }
    // the moments1 overloads (FourierTransform.cpp:475-485, 501-511): the fitted moments
    // with m0 = 0 and m1 = moments1, for the Sigma-like input (greenFtnAsym = false)
    {
        Mat m1;
        m1.setZero(nc, nc);
        for (int j = 0; j < nc; ++j)
            for (int i = 0; i < nc; ++i) m1(i, j) = (i == j ? 1.3 + 0.1 * i : 0.05 * (i + 1) * (j + 1)) + cplx(0, 0.02 * (i - j));
        std::vector<Mat> st(ntau + 1);
        dmftgpu_shim::FourierTransform(h, g, S.w.data(), st.data(), m1, false);
        std::vector<cplx> sm(4 * (size_t)nc * nc), ost((size_t)(ntau + 1) * nc * nc), gst(ost.size());
        orc_asymto_moments(nc, nf, beta, sg.data(), 0, sm.data(), nullptr);
        for (int e = 0; e < nc * nc; ++e) {
            sm[e] = 0.0;
            sm[(size_t)nc * nc + e] = m1.v[e];
        }
        int pt = 0;
        orc_ft_w2tau(nc, nf, ntau, beta, sg.data(), sm.data(), ost.data(), &pt, 1);
        for (int t = 0; t <= ntau; ++t) std::copy(st[t].v.begin(), st[t].v.end(), gst.begin() + (size_t)t * nc * nc);
        check("FourierTransform(moments1)", relerr(gst.data(), ost.data(), ost.size()), 1e-12);
        const double tau = 0.6113 * beta;
        Mat s1;
        dmftgpu_shim::FourierTransform(h, g, S.w.data(), s1, tau, m1, false);
        std::vector<cplx> o1((size_t)nc * nc);
        orc_ft_tau(nc, nf, beta, sg.data(), sm.data(), tau, o1.data());
        check("FourierTransform(tau, moments1)", relerr(s1.v.data(), o1.data(), o1.size()), 1e-12);
    }
    dmftgpu_shim::FT_t_to_w(h, g, gw2.data(), gt.data(), nf);
    std::vector<cplx> ow((size_t)nf * nc * nc), gw((size_t)nf * nc * nc);
    int viol = 0;
    orc_ft_tau2w(nc, nf, ntau, beta, ggt.data(), ow.data(), &viol);
    for (int w = 0; w < nf; ++w) std::copy(gw2[w].v.begin(), gw2[w].v.end(), gw.begin() + (size_t)w * nc * nc);
    check("FT_t_to_w", relerr(gw.data(), ow.data(), ow.size()), 1e-12);

    // Construct_hyb_delta for the impurity of cluster 1 (orbitals 2, 3), then write_hyb_t
    {
        std::vector<int> orbit = {2, 3};
        Mat himp;
        himp.setZero(nc, nc);
        for (int a = 0; a < nc; ++a) himp(a, a) = 0.1 * a;
        himp(2, 3) = himp(3, 2) = 0.05;
        Sigma weiss;
        weiss.w.resize(nf);
        for (auto& m : weiss.w) m.setZero(2, 2);
        dmftgpu_shim::Construct_hyb_delta(h, g, 2, orbit, S, Gw, mu, weiss, 1, himp);
        std::vector<cplx> od((size_t)nf * 4), gd((size_t)nf * 4);
        orc_hyb_delta(nc, nf, beta, mu, got.data(), sg.data(), himp.v.data(), 2, orbit.data(), od.data());
        for (int w = 0; w < nf; ++w) std::copy(weiss.w[w].v.begin(), weiss.w[w].v.end(), gd.begin() + (size_t)w * 4);
        check("Construct_hyb_delta", relerr(gd.data(), od.data(), od.size()), 1e-12);
        std::vector<Mat> dt(ntau + 1);
        dmftgpu_shim::write_hyb_t(h, g, weiss, dt.data(), 2);
        std::vector<cplx> dm4(16), odt((size_t)(ntau + 1) * 4), gdt(odt.size());
        orc_asymto_moments(2, nf, beta, od.data(), 0, dm4.data(), nullptr);
        int pt = 0;
        orc_ft_w2tau(2, nf, ntau, beta, od.data(), dm4.data(), odt.data(), &pt, 1);
        for (int t = 0; t <= ntau; ++t) std::copy(dt[t].v.begin(), dt[t].v.end(), gdt.begin() + (size_t)t * 4);
        check("write_hyb_t", relerr(gdt.data(), odt.data(), odt.size()), 1e-12);
    }

    // read_HR of a two-atom Hk.HWR file, then Construct_Hk on a small mesh
    {
        const char* path = "/tmp/dmftgpu_shim_test.HWR";
        std::FILE* f = std::fopen(path, "w");
        std::fprintf(f, "Norbital of Atom = 2\nNorbital of Atom = 1\n");
        std::fprintf(f, "+0 +0 +0 1 1 1 1  0.50 0.0\n+0 +0 +0 2 2 1 1 -0.70 0.0\n");
        std::fprintf(f, "+1 +0 +0 1 2 2 1 -0.25 0.10\n-1 +0 +0 2 1 1 2 -0.25 -0.10\n");
        std::fprintf(f, "+0 +1 +0 1 1 1 2  0.05 0.02\n+0 -1 +0 1 1 2 1  0.05 -0.02\n");
        std::fclose(f);
        dmftgpu_shim::HoppingList hr = dmftgpu_shim::read_HR(path);
        const int nkk = 12;
        std::vector<double> kp(3 * nkk);
        for (int k = 0; k < nkk; ++k)
            for (int a = 0; a < 3; ++a) kp[3 * k + a] = -PI + PI / 4 + 2 * PI * ((k >> a) % 4) / 4.0;
        std::vector<Mat> Hb;
        dmftgpu_shim::Construct_Hk(h, hr, kp, Hb);
        const int no = hr.n_orbital();
        std::vector<cplx> oh((size_t)nkk * no * no), gh(oh.size());
        orc_hk_from_hr(no, (int)hr.t.size(), hr.idx.data(), hr.t.data(), nkk, kp.data(), oh.data());
        for (int k = 0; k < nkk; ++k) std::copy(Hb[k].v.begin(), Hb[k].v.end(), gh.begin() + (size_t)k * no * no);
        check("read_HR + Construct_Hk", relerr(gh.data(), oh.data(), oh.size()) + (no == 3 && hr.t.size() == 6 ? 0 : 1),
              1e-14);
        double herm = 0;
        for (int k = 0; k < nkk; ++k)
            for (int i = 0; i < no; ++i)
                for (int j = 0; j < no; ++j) herm = std::fmax(herm, std::abs(Hb[k](i, j) - std::conj(Hb[k](j, i))));
        check("Construct_Hk Hermitian", herm, 1e-15);
    }

    // upfolding_density: 3 model bands inside 5 KS bands, 4 k points, a Hermitian
    // positive definite overlap, every (R, i0, m0) with R in {-1,0,1} x {0} x {0}
    {
        const int n = 3, no = 5, nk = 4;
        dmftgpu_shim::Globals gu;
        gu.knum = nk;
        gu.knum_mpiGlobal = nk;
        gu.beta = 30.0;
        const double mu = 0.1;
        std::vector<Mat> dm(nk), cb(nk), ko(nk), tr(nk), so(nk);
        std::vector<std::vector<double>> eps(nk, std::vector<double>(no));
        std::vector<std::vector<int>> v2k(nk);
        for (int k = 0; k < nk; ++k) {
            dm[k].setZero(n, n);
            cb[k].setZero(n, n);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j <= i; ++j) {
                    dm[k](i, j) = cplx(urand(), i == j ? 0.0 : urand());
                    dm[k](j, i) = std::conj(dm[k](i, j));
                }
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) cb[k](i, j) = cplx(i == j ? 1.0 : 0.0) + 0.2 * cplx(urand(), urand());
            ko[k].setZero(no, no);
            tr[k].setZero(no, no);
            so[k].setZero(no, no);
            for (int i = 0; i < no; ++i)
                for (int j = 0; j < no; ++j) {
                    ko[k](i, j) = cplx(i == j ? 1.0 : 0.0) + 0.3 * cplx(urand(), urand());
                    tr[k](i, j) = cplx(i == j ? 1.0 : 0.0) + 0.2 * cplx(urand(), urand());
                }
            for (int i = 0; i < no; ++i)
                for (int j = 0; j < no; ++j) {
                    cplx acc = i == j ? 1.0 : 0.0;
                    for (int l = 0; l < no; ++l) acc += 0.1 * tr[k](i, l) * std::conj(tr[k](j, l));
                    so[k](i, j) = acc;
                }
            for (int b = 0; b < no; ++b) eps[k][b] = 4.0 * urand() + (b == 0 ? -10.0 : 0.0);
            v2k[k] = {(k + 1) % no, (k + 3) % no, (k + 4) % no};
        }
        std::vector<double> kp(3 * nk);
        for (int k = 0; k < nk; ++k) kp[3 * k] = -PI + PI / 4 + 2 * PI * k / 4.0;
        IndexMat hri;
        for (int r = -1; r <= 1; ++r)
            for (int i = 0; i < no; ++i)
                for (int j = 0; j < no; ++j) hri.v.insert(hri.v.end(), {r, 0, 0, i, j});
        const int nr = hri.rows();
        // oracle on the same inputs
        std::vector<cplx> X((size_t)nk * n * n), C((size_t)nk * n * n), T((size_t)nk * no * no),
            K((size_t)nk * no * no), S((size_t)nk * no * no), od((size_t)nk * no * no), odmr(3 * (size_t)nr);
        std::vector<double> e((size_t)nk * no);
        std::vector<int> vv((size_t)nk * n);
        for (int k = 0; k < nk; ++k) {
            std::copy(dm[k].v.begin(), dm[k].v.end(), X.begin() + (size_t)k * n * n);
            std::copy(cb[k].v.begin(), cb[k].v.end(), C.begin() + (size_t)k * n * n);
            std::copy(tr[k].v.begin(), tr[k].v.end(), T.begin() + (size_t)k * no * no);
            std::copy(ko[k].v.begin(), ko[k].v.end(), K.begin() + (size_t)k * no * no);
            std::copy(so[k].v.begin(), so[k].v.end(), S.begin() + (size_t)k * no * no);
            for (int b = 0; b < no; ++b) e[(size_t)k * no + b] = eps[k][b];
            for (int i = 0; i < n; ++i) vv[(size_t)k * n + i] = v2k[k][i];
        }
        orc_upfold_density(n, no, nk, nk, gu.beta, mu, X.data(), C.data(), e.data(), vv.data(), T.data(), K.data(),
                           S.data(), nr, hri.v.data(), kp.data(), od.data(), odmr.data());
        std::vector<cplx> dmr[3];
        dmftgpu_shim::upfolding_density(h, gu, dm, ko, hri, mu, so, tr, cb, eps, v2k, kp, dmr);
        std::vector<cplx> gd((size_t)nk * no * no), gr(3 * (size_t)nr);
        for (int k = 0; k < nk; ++k) std::copy(dm[k].v.begin(), dm[k].v.end(), gd.begin() + (size_t)k * no * no);
        for (int x = 0; x < 3; ++x) std::copy(dmr[x].begin(), dmr[x].end(), gr.begin() + (size_t)x * nr);
        check("upfolding_density D_k", relerr(gd.data(), od.data(), od.size()) + (dm[0].rows() == no ? 0 : 1), 1e-13);
        check("upfolding_density DM(R)", relerr(gr.data(), odmr.data(), odmr.size()), 1e-13);

        // NBAND[k] varying: k = 0, 1 keep 2 of the 3 model bands; the oracle runs
        // each band count on its own k points with the global k count
        std::vector<Mat> dmr2(nk), cbr(nk);
        std::vector<std::vector<int>> vr(nk);
        for (int k = 0; k < nk; ++k) {
            const int nb = k < 2 ? n - 1 : n;
            dmr2[k].setZero(nb, nb);
            cbr[k].setZero(nb, nb);
            for (int j = 0; j < nb; ++j)
                for (int i = 0; i < nb; ++i) {
                    dmr2[k](i, j) = X[(size_t)k * n * n + i + (size_t)j * n];
                    cbr[k](i, j) = C[(size_t)k * n * n + i + (size_t)j * n];
                }
            vr[k].assign(v2k[k].begin(), v2k[k].begin() + nb);
        }
        const std::vector<Mat> dmr2_in = dmr2;  // the adapter overwrites densityMatDFT in place (DFT basis)
        std::vector<cplx> rdmr[3];
        dmftgpu_shim::upfolding_density(h, gu, dmr2, ko, hri, mu, so, tr, cbr, eps, vr, kp, rdmr);
        std::vector<cplx> osum(3 * (size_t)nr, 0.0), od2((size_t)nk * no * no);
        for (int part = 0; part < 2; ++part) {
            const int nb = part == 0 ? n - 1 : n, k0 = 2 * part, kn = 2;
            std::vector<cplx> Xp((size_t)kn * nb * nb), Cp((size_t)kn * nb * nb), part_dmr(3 * (size_t)nr);
            std::vector<int> vp((size_t)kn * nb);
            for (int k = 0; k < kn; ++k) {
                std::copy(dmr2_in[k0 + k].v.begin(), dmr2_in[k0 + k].v.end(), Xp.begin() + (size_t)k * nb * nb);
                std::copy(cbr[k0 + k].v.begin(), cbr[k0 + k].v.end(), Cp.begin() + (size_t)k * nb * nb);
                for (int i = 0; i < nb; ++i) vp[(size_t)k * nb + i] = vr[k0 + k][i];
            }
            orc_upfold_density(nb, no, kn, nk, gu.beta, mu, Xp.data(), Cp.data(), e.data() + (size_t)k0 * no, vp.data(),
                               T.data() + (size_t)k0 * no * no, K.data() + (size_t)k0 * no * no,
                               S.data() + (size_t)k0 * no * no, nr, hri.v.data(), kp.data() + 3 * (size_t)k0,
                               od2.data() + (size_t)k0 * no * no, part_dmr.data());
            for (size_t q = 0; q < osum.size(); ++q) osum[q] += part_dmr[q];
        }
        std::vector<cplx> gd2((size_t)nk * no * no), gr2(3 * (size_t)nr);
        for (int k = 0; k < nk; ++k) std::copy(dmr2[k].v.begin(), dmr2[k].v.end(), gd2.begin() + (size_t)k * no * no);
        for (int x = 0; x < 3; ++x) std::copy(rdmr[x].begin(), rdmr[x].end(), gr2.begin() + (size_t)x * nr);
        check("ragged upfolding D_k", relerr(gd2.data(), od2.data(), od2.size()), 1e-13);
        check("ragged upfolding DM(R)", relerr(gr2.data(), osum.data(), osum.size()), 1e-13);
    }

    // qs_spectra band loop: 3 k points of a path, a 4-band model (2 correlated) inside 6 KS bands;
    // H(k) = C diag(eps of the window bands) C^H with C unitary (a Householder reflection), as the
    // downfolded model's eigenvalues are the window's KS energies
    {
        const int nb = 4, ncq = 2, no = 6, nkp = 3, ne = 40;
        const double e0q = -3.0, deq = 0.15, eta = 0.05, muq = 0.1;
        std::vector<std::vector<double>> eks(nkp, std::vector<double>(no));
        std::vector<cplx> hq((size_t)nkp * nb * nb);
        for (int k = 0; k < nkp; ++k) {
            for (int b = 0; b < no; ++b) eks[k][b] = -4.0 + 8.0 * (b + 0.5 * (urand() + 0.5)) / no;
            std::vector<cplx> u(nb);
            double uu = 0;
            for (int i = 0; i < nb; ++i) { u[i] = cplx(urand(), urand()); uu += std::norm(u[i]); }
            for (int j = 0; j < nb; ++j)
                for (int i = 0; i < nb; ++i) {
                    cplx s = 0.0;
                    for (int l = 0; l < nb; ++l) {  // C = I - 2 u u^H / |u|^2; H = C diag(eps_{1..4}) C^H
                        const cplx cil = (i == l ? 1.0 : 0.0) - 2.0 * u[i] * std::conj(u[l]) / uu;
                        const cplx cjl = (j == l ? 1.0 : 0.0) - 2.0 * u[j] * std::conj(u[l]) / uu;
                        s += cil * eks[k][1 + l] * std::conj(cjl);
                    }
                    hq[(size_t)k * nb * nb + i + (size_t)j * nb] = s;
                }
        }
        std::vector<Mat> sr(ne), s0(ne);
        std::vector<cplx> hs((size_t)ne * ncq * ncq), h0((size_t)ne * ncq * ncq, 0.0);
        for (int n = 0; n < ne; ++n) {
            sr[n].setZero(ncq, ncq);
            s0[n].setZero(ncq, ncq);
            const cplx E(e0q + n * deq, 0.0);
            sr[n](0, 0) = 0.4 / (E - 0.3 + cplx(0.0, 0.6));
            sr[n](1, 1) = 0.3 / (E + 0.5 + cplx(0.0, 0.4));
            sr[n](0, 1) = sr[n](1, 0) = 0.1 / (E + cplx(0.0, 0.5));
            std::copy(sr[n].v.begin(), sr[n].v.end(), hs.begin() + (size_t)n * ncq * ncq);
        }
        dmftgpu_shim::DevBuf dh(hq.size()), dm(4 * (size_t)ncq * ncq);
        dh.upload(hq.data());
        std::vector<cplx> zm(dm.n, 0.0);
        dm.upload(zm.data());
        dmftgpu_lattice Lq{nb, ncq, 1, nkp, nkp, 16, 2, 40.0};
        dmftgpu_shim::die(h, dmftgpu_prepare(h, &Lq, (const double*)dh.p, (const double*)dm.p), "prepare path");
        double err0 = 0, errs = 0, errp = 0, scale = 0;
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<std::vector<double>> sw;
            std::vector<std::vector<std::vector<double>>> pw;
            dmftgpu_shim::qs_band_spectra(h, DMFTGPU_SCHUR_F64, ncq, nkp, ne, e0q, deq, eta, muq, pass ? sr : s0, eks,
                                          no, sw, pw);
            for (int k = 0; k < nkp; ++k) {
                std::vector<cplx> gr((size_t)ne * ncq * ncq), tr(ne), g0((size_t)ne * ncq * ncq), t0(ne);
                orc_lattice L1{nb, ncq, 1, 1, 1, ne, 1, 1.0};
                orc_greenftn_real(&L1, hq.data() + (size_t)k * nb * nb, ne, e0q, deq, eta, pass ? hs.data() : h0.data(),
                                  muq, gr.data(), tr.data(), 1);
                orc_greenftn_real(&L1, hq.data() + (size_t)k * nb * nb, ne, e0q, deq, eta, h0.data(), muq, g0.data(),
                                  t0.data(), 1);
                for (int n = 0; n < ne; ++n) {
                    const cplx z(e0q + n * deq + muq, eta);
                    cplx full = 0.0;
                    for (int b = 0; b < no; ++b) full += 1.0 / (z - eks[k][b]);
                    const double ref_free = -full.imag() / PI;                          // Sigma = 0: closed form
                    const double ref = -(full + tr[n] - t0[n]).imag() / PI;
                    scale = std::fmax(scale, std::fabs(ref));
                    if (!pass) err0 = std::fmax(err0, std::fabs(sw[k][n] - ref_free));
                    else errs = std::fmax(errs, std::fabs(sw[k][n] - ref));
                    for (int i = 0; i < ncq; ++i)
                        errp = std::fmax(errp, std::fabs(pw[k][n][i] + gr[(size_t)n * ncq * ncq + i + (size_t)i * ncq].imag() / PI));
                }
            }
        }
        check("qs_spectra bands, Sigma=0", err0 / scale, 1e-13);
        check("qs_spectra bands", errs / scale, 1e-12);
        check("qs_spectra partial weights", errp / scale, 1e-12);
    }

    dmftgpu_destroy(h);
    std::printf(failures ? "SHIM FAILED\n" : "SHIM OK\n");
    return failures ? 1 : 0;
}
