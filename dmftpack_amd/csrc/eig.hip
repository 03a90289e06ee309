// Batched Hermitian eigen-decomposition, one wavefront per matrix.
//
// Replaces the per-k SelfAdjointEigenSolver calls of the static part of
// GreenFtn_w (GreenFtn.cpp:197-215) and the small Hermitian problems of the
// Fourier transform.  Cyclic Jacobi with complex rotations on a matrix held
// in LDS; every rotation is applied by up to 64 lanes in parallel (lane i owns
// row/column i).  Eigenvalues are returned in ascending order like Eigen.
#include "lattice_kernels.hpp"

namespace dg {

__global__ __launch_bounds__(64) void k_herm_eig(const cd* __restrict__ X, int ldx, int off, int m,
                                                 const cd* __restrict__ add, int nadd, double* __restrict__ w,
                                                 cd* __restrict__ V, int batch)
{
    extern __shared__ cd sm[];
    cd* A = sm;          // m*m
    cd* Vs = sm + m * m;  // m*m
    const int b = blockIdx.x;
    if (b >= batch) return;
    const int lane = threadIdx.x;
    const cd* Xb = X + (size_t)b * ldx * ldx;
    for (int e = lane; e < m * m; e += 64) {
        int i = e % m, j = e / m;
        cd v = Xb[(size_t)(off + i) + (size_t)(off + j) * ldx];
        cd vt = Xb[(size_t)(off + j) + (size_t)(off + i) * ldx];
        if (add && i < nadd && j < nadd) {
            v = cadd(v, add[i + j * nadd]);
            vt = cadd(vt, add[j + i * nadd]);
        }
        // Hermitian part (A + A^H)/2
        A[e] = cd{0.5 * (v.x + vt.x), 0.5 * (v.y - vt.y)};
        Vs[e] = cd{i == j ? 1.0 : 0.0, 0.0};
    }
    __syncthreads();
    // Frobenius norm^2
    double nrm = 0.0;
    for (int e = lane; e < m * m; e += 64) nrm += A[e].x * A[e].x + A[e].y * A[e].y;
    for (int o = 32; o > 0; o >>= 1) nrm += __shfl_xor(nrm, o);
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off2 = 0.0;
        for (int e = lane; e < m * m; e += 64) {
            int i = e % m, j = e / m;
            if (i != j) off2 += A[e].x * A[e].x + A[e].y * A[e].y;
        }
        for (int o = 32; o > 0; o >>= 1) off2 += __shfl_xor(off2, o);
        if (off2 <= 1e-34 * nrm || off2 == 0.0) break;
        for (int p = 0; p < m - 1; ++p) {
            for (int q = p + 1; q < m; ++q) {
                const cd bpq = A[p + q * m];
                const double ab = sqrt(bpq.x * bpq.x + bpq.y * bpq.y);
                if (ab == 0.0) continue;  // uniform across the wave
                const double app = A[p + p * m].x, aqq = A[q + q * m].x;
                const cd ph = cd{bpq.x / ab, bpq.y / ab};
                const cd phc = cconj(ph);
                const double zeta = (aqq - app) / (2.0 * ab);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double s = t * c;
                __syncthreads();
                if (lane < m) {  // columns p, q of A and V
                    cd xp = A[lane + p * m], xq = A[lane + q * m];
                    A[lane + p * m] = csub(cscale(xp, c), cscale(cmul(phc, xq), s));
                    A[lane + q * m] = cadd(cscale(xp, s), cscale(cmul(phc, xq), c));
                    cd vp = Vs[lane + p * m], vq = Vs[lane + q * m];
                    Vs[lane + p * m] = csub(cscale(vp, c), cscale(cmul(phc, vq), s));
                    Vs[lane + q * m] = cadd(cscale(vp, s), cscale(cmul(phc, vq), c));
                }
                __syncthreads();
                if (lane < m) {  // rows p, q of A
                    cd xp = A[p + lane * m], xq = A[q + lane * m];
                    A[p + lane * m] = csub(cscale(xp, c), cscale(cmul(ph, xq), s));
                    A[q + lane * m] = cadd(cscale(xp, s), cscale(cmul(ph, xq), c));
                }
                __syncthreads();
                if (lane == 0) {
                    A[p + q * m] = cd{0.0, 0.0};
                    A[q + p * m] = cd{0.0, 0.0};
                    A[p + p * m].y = 0.0;
                    A[q + q * m].y = 0.0;
                }
                __syncthreads();
            }
        }
    }
    __syncthreads();
    // rank sort (stable, ascending)
    for (int i = lane; i < m; i += 64) {
        double wi = A[i + i * m].x;
        int r = 0;
        for (int j = 0; j < m; ++j) {
            double wj = A[j + j * m].x;
            r += (wj < wi) || (wj == wi && j < i);
        }
        w[(size_t)b * m + r] = wi;
        if (V)
            for (int e = 0; e < m; ++e) V[(size_t)b * m * m + e + (size_t)r * m] = Vs[e + i * m];
    }
}

hipError_t launch_herm_eig(const cd* X, int ldx, int off, int m, const cd* add, int nadd, double* w, cd* V, int batch,
                           hipStream_t s)
{
    if (batch <= 0 || m <= 0) return hipSuccess;
    size_t shm = sizeof(cd) * 2 * (size_t)m * m;
    hipLaunchKernelGGL(k_herm_eig, dim3(batch), dim3(64), shm, s, X, ldx, off, m, add, nadd, w, V, batch);
    return hipGetLastError();
}

}  // namespace dg
