// Small dense complex linear algebra on the host (d <= 32): the tail-fit and
// Fourier-transform set-up works on a handful of d x d matrices per call.
// Column-major, element (i,j) at i + j*d.
#pragma once
#include <cmath>
#include <complex>
#include <vector>

namespace hl {

using cx = std::complex<double>;
using Mat = std::vector<cx>;

inline Mat zeros(int d) { return Mat((size_t)d * d, cx(0.0, 0.0)); }
inline Mat eye(int d)
{
    Mat m = zeros(d);
    for (int i = 0; i < d; ++i) m[i + (size_t)i * d] = 1.0;
    return m;
}
inline Mat mul(int d, const Mat& A, const Mat& B)
{
    Mat C = zeros(d);
    for (int j = 0; j < d; ++j)
        for (int l = 0; l < d; ++l) {
            const cx b = B[l + (size_t)j * d];
            for (int i = 0; i < d; ++i) C[i + (size_t)j * d] += A[i + (size_t)l * d] * b;
        }
    return C;
}
inline Mat adj(int d, const Mat& A)
{
    Mat C = zeros(d);
    for (int j = 0; j < d; ++j)
        for (int i = 0; i < d; ++i) C[i + (size_t)j * d] = std::conj(A[j + (size_t)i * d]);
    return C;
}
inline Mat lin(const Mat& A, cx a, const Mat& B, cx b)
{
    Mat C(A.size());
    for (size_t e = 0; e < A.size(); ++e) C[e] = a * A[e] + b * B[e];
    return C;
}

// Hermitian eigen-decomposition (cyclic Jacobi), eigenvalues ascending,
// V columns are eigenvectors.  The input is symmetrised as (A + A^H)/2.
inline void eigh(int d, const Mat& Ain, std::vector<double>& w, Mat& V)
{
    Mat a = zeros(d);
    V = eye(d);
    for (int j = 0; j < d; ++j)
        for (int i = 0; i < d; ++i) a[i + (size_t)j * d] = 0.5 * (Ain[i + (size_t)j * d] + std::conj(Ain[j + (size_t)i * d]));
    double nrm = 0.0;
    for (auto& x : a) nrm += std::norm(x);
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int q = 1; q < d; ++q)
            for (int p = 0; p < q; ++p) off += 2.0 * std::norm(a[p + (size_t)q * d]);
        if (off <= 1e-34 * nrm || off == 0.0) break;
        for (int p = 0; p < d - 1; ++p)
            for (int q = p + 1; q < d; ++q) {
                const cx b = a[p + (size_t)q * d];
                const double ab = std::abs(b);
                if (ab == 0.0) continue;
                const double app = a[p + (size_t)p * d].real(), aqq = a[q + (size_t)q * d].real();
                const cx ph = b / ab, phc = std::conj(ph);
                const double zeta = (aqq - app) / (2.0 * ab);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int i = 0; i < d; ++i) {
                    cx xp = a[i + (size_t)p * d], xq = a[i + (size_t)q * d];
                    a[i + (size_t)p * d] = c * xp - s * phc * xq;
                    a[i + (size_t)q * d] = s * xp + c * phc * xq;
                    cx vp = V[i + (size_t)p * d], vq = V[i + (size_t)q * d];
                    V[i + (size_t)p * d] = c * vp - s * phc * vq;
                    V[i + (size_t)q * d] = s * vp + c * phc * vq;
                }
                for (int j = 0; j < d; ++j) {
                    cx xp = a[p + (size_t)j * d], xq = a[q + (size_t)j * d];
                    a[p + (size_t)j * d] = c * xp - s * ph * xq;
                    a[q + (size_t)j * d] = s * xp + c * ph * xq;
                }
                a[p + (size_t)q * d] = 0.0;
                a[q + (size_t)p * d] = 0.0;
                a[p + (size_t)p * d] = a[p + (size_t)p * d].real();
                a[q + (size_t)q * d] = a[q + (size_t)q * d].real();
            }
    }
    std::vector<int> ord(d);
    w.assign(d, 0.0);
    for (int i = 0; i < d; ++i) ord[i] = i;
    for (int i = 0; i < d; ++i) {
        int m = i;
        for (int j = i + 1; j < d; ++j)
            if (a[ord[j] + (size_t)ord[j] * d].real() < a[ord[m] + (size_t)ord[m] * d].real()) m = j;
        std::swap(ord[i], ord[m]);
    }
    Mat Vs = zeros(d);
    for (int j = 0; j < d; ++j) {
        w[j] = a[ord[j] + (size_t)ord[j] * d].real();
        for (int i = 0; i < d; ++i) Vs[i + (size_t)j * d] = V[i + (size_t)ord[j] * d];
    }
    V = Vs;
}

inline double min_eig(int d, const Mat& A)
{
    std::vector<double> w;
    Mat V;
    eigh(d, A, w, V);
    return w[0];
}

// U diag(f(w)) U^H of a Hermitian matrix
template <class F>
inline Mat herm_func(int d, const Mat& A, F f)
{
    std::vector<double> w;
    Mat V;
    eigh(d, A, w, V);
    Mat R = zeros(d);
    for (int l = 0; l < d; ++l) {
        const double fl = f(w[l]);
        for (int j = 0; j < d; ++j)
            for (int i = 0; i < d; ++i) R[i + (size_t)j * d] += V[i + (size_t)l * d] * fl * std::conj(V[j + (size_t)l * d]);
    }
    return R;
}

// Partial-pivot LU inverse; returns false if singular.
inline bool inverse(int d, const Mat& A, Mat& X)
{
    Mat lu = A;
    std::vector<int> perm(d);
    for (int i = 0; i < d; ++i) perm[i] = i;
    for (int k = 0; k < d; ++k) {
        int p = k;
        double best = std::abs(lu[k + (size_t)k * d]);
        for (int i = k + 1; i < d; ++i)
            if (std::abs(lu[i + (size_t)k * d]) > best) {
                best = std::abs(lu[i + (size_t)k * d]);
                p = i;
            }
        if (best == 0.0) return false;
        if (p != k) {
            for (int j = 0; j < d; ++j) std::swap(lu[k + (size_t)j * d], lu[p + (size_t)j * d]);
            std::swap(perm[k], perm[p]);
        }
        const cx inv = 1.0 / lu[k + (size_t)k * d];
        for (int i = k + 1; i < d; ++i) lu[i + (size_t)k * d] *= inv;
        for (int j = k + 1; j < d; ++j) {
            const cx u = lu[k + (size_t)j * d];
            for (int i = k + 1; i < d; ++i) lu[i + (size_t)j * d] -= lu[i + (size_t)k * d] * u;
        }
    }
    X = zeros(d);
    for (int c = 0; c < d; ++c) {
        cx* x = &X[(size_t)c * d];
        for (int i = 0; i < d; ++i) x[i] = (perm[i] == c) ? 1.0 : 0.0;
        for (int i = 0; i < d; ++i)
            for (int j = 0; j < i; ++j) x[i] -= lu[i + (size_t)j * d] * x[j];
        for (int i = d - 1; i >= 0; --i) {
            for (int j = i + 1; j < d; ++j) x[i] -= lu[i + (size_t)j * d] * x[j];
            x[i] /= lu[i + (size_t)i * d];
        }
    }
    return true;
}

}  // namespace hl
