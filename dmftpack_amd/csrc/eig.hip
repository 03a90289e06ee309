// Batched Hermitian eigen-decomposition, one wavefront per matrix.
//
// Replaces the per-k SelfAdjointEigenSolver calls of the static part of
// GreenFtn_w (GreenFtn.cpp:197-215) and the small Hermitian problems of the
// Fourier transform.  Jacobi with complex rotations on a matrix held in LDS,
// in parallel (round-robin) order: the m/2 disjoint rotations of a round are
// applied together by the 64 lanes.  Eigenvalues are returned in ascending
// order like Eigen.
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
    // Parallel (round-robin) ordering: each round rotates M/2 disjoint pairs at
    // once (A <- J^H A J with J a product of commuting plane rotations), so a
    // sweep is M-1 rounds of wave-wide updates instead of m(m-1)/2 rotations
    // applied one after another.  M = m rounded up to even; pairs touching the
    // padding index are skipped.
    const int M = m + (m & 1);
    const int np = M / 2;
    double* rc = (double*)(Vs + m * m);  // per pair: c, s, phase (x, y) and p, q
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off2 = 0.0;
        for (int e = lane; e < m * m; e += 64) {
            int i = e % m, j = e / m;
            if (i != j) off2 += A[e].x * A[e].x + A[e].y * A[e].y;
        }
        for (int o = 32; o > 0; o >>= 1) off2 += __shfl_xor(off2, o);
        if (off2 <= 1e-34 * nrm || off2 == 0.0) break;
        for (int r = 0; r < M - 1; ++r) {
            if (lane < np) {
                // circle method: (r, M-1) and (r+k, r-k) mod (M-1)
                int a_, b_;
                if (lane == 0) {
                    a_ = r;
                    b_ = M - 1;
                } else {
                    a_ = (r + lane) % (M - 1);
                    b_ = (r - lane + M - 1) % (M - 1);
                }
                const int p = min(a_, b_), q = max(a_, b_);
                double c = 1.0, sn = 0.0, phx = 1.0, phy = 0.0;
                if (q < m) {
                    const cd bpq = A[p + q * m];
                    const double ab = sqrt(bpq.x * bpq.x + bpq.y * bpq.y);
                    if (ab != 0.0) {
                        const double app = A[p + p * m].x, aqq = A[q + q * m].x;
                        phx = bpq.x / ab;
                        phy = bpq.y / ab;
                        const double zeta = (aqq - app) / (2.0 * ab);
                        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                        c = 1.0 / sqrt(1.0 + t * t);
                        sn = t * c;
                    }
                }
                rc[6 * lane] = c;
                rc[6 * lane + 1] = sn;
                rc[6 * lane + 2] = phx;
                rc[6 * lane + 3] = phy;
                rc[6 * lane + 4] = (double)p;
                rc[6 * lane + 5] = (double)(q < m ? q : -1);
            }
            __syncthreads();
            // columns p, q of A and V: A <- A J
            for (int t = lane; t < np * m; t += 64) {
                const int k = t / m, i = t % m;
                const int q = (int)rc[6 * k + 5];
                if (q < 0) continue;
                const int p = (int)rc[6 * k + 4];
                const double c = rc[6 * k], sn = rc[6 * k + 1];
                const cd phc = cd{rc[6 * k + 2], -rc[6 * k + 3]};
                cd xp = A[i + p * m], xq = A[i + q * m];
                A[i + p * m] = csub(cscale(xp, c), cscale(cmul(phc, xq), sn));
                A[i + q * m] = cadd(cscale(xp, sn), cscale(cmul(phc, xq), c));
                cd vp = Vs[i + p * m], vq = Vs[i + q * m];
                Vs[i + p * m] = csub(cscale(vp, c), cscale(cmul(phc, vq), sn));
                Vs[i + q * m] = cadd(cscale(vp, sn), cscale(cmul(phc, vq), c));
            }
            __syncthreads();
            // rows p, q of A: A <- J^H A
            for (int t = lane; t < np * m; t += 64) {
                const int k = t / m, j = t % m;
                const int q = (int)rc[6 * k + 5];
                if (q < 0) continue;
                const int p = (int)rc[6 * k + 4];
                const double c = rc[6 * k], sn = rc[6 * k + 1];
                const cd ph = cd{rc[6 * k + 2], rc[6 * k + 3]};
                cd xp = A[p + j * m], xq = A[q + j * m];
                cd yp = csub(cscale(xp, c), cscale(cmul(ph, xq), sn));
                cd yq = cadd(cscale(xp, sn), cscale(cmul(ph, xq), c));
                // a rotated pair's off-diagonal entries are zero and its
                // diagonal real by construction: store them exactly
                if (sn != 0.0) {
                    if (j == q) {
                        yp = cd{0.0, 0.0};
                        yq.y = 0.0;
                    } else if (j == p) {
                        yq = cd{0.0, 0.0};
                        yp.y = 0.0;
                    }
                }
                A[p + j * m] = yp;
                A[q + j * m] = yq;
            }
            __syncthreads();
        }
    }
    __syncthreads();
    // rank sort (stable, ascending); perm[r] = column of Vs with rank r (the
    // pair table is free now and holds the m <= 32 ints)
    int* perm = (int*)rc;
    for (int i = lane; i < m; i += 64) {
        double wi = A[i + i * m].x;
        int r = 0;
        for (int j = 0; j < m; ++j) {
            double wj = A[j + j * m].x;
            r += (wj < wi) || (wj == wi && j < i);
        }
        w[(size_t)b * m + r] = wi;
        perm[r] = i;
    }
    if (V) {
        __syncthreads();
        for (int e = lane; e < m * m; e += 64) {
            int c = perm[e / m];
            if ((unsigned)c >= (unsigned)m) c = e / m;  // NaN eigenvalues leave ranks unassigned
            V[(size_t)b * m * m + e] = Vs[e % m + c * m];
        }
    }
}

hipError_t launch_herm_eig(const cd* X, int ldx, int off, int m, const cd* add, int nadd, double* w, cd* V, int batch,
                           hipStream_t s)
{
    if (batch <= 0 || m <= 0) return hipSuccess;
    if (m > 32) return hipErrorInvalidValue;  // the pair table holds 16 rotations
    size_t shm = sizeof(cd) * 2 * (size_t)m * m + sizeof(double) * 6 * 17;
    hipLaunchKernelGGL(k_herm_eig, dim3(batch), dim3(64), shm, s, X, ldx, off, m, add, nadd, w, V, batch);
    return hipGetLastError();
}

}  // namespace dg
