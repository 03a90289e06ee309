// Batched Hermitian eigen-decomposition.
//
// Replaces the per-k SelfAdjointEigenSolver calls of the static part of
// GreenFtn_w (GreenFtn.cpp:197-215) and the small Hermitian problems of the
// Fourier transform.  Jacobi with complex rotations in a parallel order: the
// disjoint rotations of a round are applied together.  m <= 16: one matrix per
// 4-, 8- or 16-lane slot of a wavefront, rows in registers (k_herm_eig_slots);
// 17 <= m <= 32: one wavefront per matrix held in LDS (k_herm_eig).
// Eigenvalues are returned in ascending order like Eigen.
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
    double* rc = (double*)(Vs + m * m);  // per pair: c, s, phase (x, y)
    int* pqt = (int*)(rc + 4 * 16);      // per pair: p, q (q = -1: pair skipped)
    // Each lane's work items of a pass (pair k, row or column i) are the same
    // every round: their indices are formed once here instead of by an
    // integer division per item per round.  np * m <= 16 * 32 = 8 * 64.
    const int nitem = np * m;
    int it_k[8], it_i[8];
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int t = lane + 64 * it;
        it_k[it] = t < nitem ? t / m : -1;
        it_i[it] = t < nitem ? t % m : 0;
    }
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
                rc[4 * lane] = c;
                rc[4 * lane + 1] = sn;
                rc[4 * lane + 2] = phx;
                rc[4 * lane + 3] = phy;
                pqt[2 * lane] = p;
                pqt[2 * lane + 1] = q < m ? q : -1;
            }
            __syncthreads();
            // columns p, q of A and V: A <- A J
#pragma unroll
            for (int it = 0; it < 8; ++it) {
                const int k = it_k[it], i = it_i[it];
                if (k < 0) continue;
                const int q = pqt[2 * k + 1];
                if (q < 0) continue;
                const int p = pqt[2 * k];
                const double c = rc[4 * k], sn = rc[4 * k + 1];
                const cd phc = cd{rc[4 * k + 2], -rc[4 * k + 3]};
                cd xp = A[i + p * m], xq = A[i + q * m];
                A[i + p * m] = csub(cscale(xp, c), cscale(cmul(phc, xq), sn));
                A[i + q * m] = cadd(cscale(xp, sn), cscale(cmul(phc, xq), c));
                cd vp = Vs[i + p * m], vq = Vs[i + q * m];
                Vs[i + p * m] = csub(cscale(vp, c), cscale(cmul(phc, vq), sn));
                Vs[i + q * m] = cadd(cscale(vp, sn), cscale(cmul(phc, vq), c));
            }
            __syncthreads();
            // rows p, q of A: A <- J^H A
#pragma unroll
            for (int it = 0; it < 8; ++it) {
                const int k = it_k[it], j = it_i[it];
                if (k < 0) continue;
                const int q = pqt[2 * k + 1];
                if (q < 0) continue;
                const int p = pqt[2 * k];
                const double c = rc[4 * k], sn = rc[4 * k + 1];
                const cd ph = cd{rc[4 * k + 2], rc[4 * k + 3]};
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

// ---------------------------------------------------------------------------
// m <= 16: one matrix per M-lane slot of the wavefront (M = 4, 8 or 16), with
// row j of A and of V in the registers of lane j, stored in XOR order:
// register r holds column j^r, so the diagonal is register 0 of every lane.
// A sweep is the M-1 rounds s = 1..M-1 that pair every index j with j^s (each
// round a perfect matching of 0..M-1, and together all M(M-1)/2 pairs once),
// and in round s a lane's partner entry A[j][j^s] is register s: with the
// rounds unrolled every register index is a compile-time constant.
//   column update A <- A J, V <- V J   lane-local (columns j^r and j^r^s are
//                                      registers r and r^s of the same lane);
//   row update    A <- J^H A           one exchange with lane j^s (ds_swizzle,
//                                      partner register r^s into register r).
// The rotation of pair (p, q), p < q, is the one of the LDS kernel below
// (zero A[p][q] with a real diagonal); both lanes compute it from the same
// three inputs, so they agree bit for bit.
// ---------------------------------------------------------------------------

// value of lane (lane ^ S) within each 32-lane group
template <int S>
__device__ __forceinline__ double swz_xor(double v)
{
    constexpr int pat = 0x1F | (S << 10);  // bit-mask mode: and 0x1F, or 0, xor S
    const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), pat);
    const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), pat);
    return __hiloint2double(hi, lo);
}
template <int S>
__device__ __forceinline__ cd swz_xor(cd v)
{
    return cd{swz_xor<S>(v.x), swz_xor<S>(v.y)};
}

// al*x + be*y
__device__ __forceinline__ cd lin2(cd al, cd x, cd be, cd y) { return cfma(be, y, cmul(al, x)); }
// conj(al)*x + conj(be)*y
__device__ __forceinline__ cd lin2c(cd al, cd x, cd be, cd y) { return cfmac(be, y, cmulc(x, al)); }

template <int S>
constexpr int high_bit()
{
    return S >= 8 ? 8 : S >= 4 ? 4 : S >= 2 ? 2 : 1;
}

// Round s = S of a sweep.  tab: this slot's M entries (row coefficients of
// each index: new row j = al_j B_j + be_j B_{j^S}, B = A J).
template <int M, bool WANTV, int S>
__device__ __forceinline__ void eig_round(cd (&a)[M], cd (&v)[M], int j, int m, bool conv, double4* tab)
{
    constexpr int HS = high_bit<S>();
    // an opaque copy of j: without it the table addresses of every round are
    // hoisted out of the sweep loop and held in registers
    asm volatile("" : "+v"(j));
    const bool isp = (j & HS) == 0;  // j is the lower index of its pair
    const double own0 = a[0].x;
    const double oth0 = swz_xor<S>(own0);
    const cd othS = swz_xor<S>(a[S]);
    const double app = isp ? own0 : oth0, aqq = isp ? oth0 : own0;
    const cd bpq = isp ? a[S] : othS;  // A[p][q], held by lane p
    double c = 1.0, sn = 0.0, phx = 1.0, phy = 0.0;
    if (!conv && (isp ? j ^ S : j) < m) {  // max(j, j^S) < m
        const double ab = sqrt(bpq.x * bpq.x + bpq.y * bpq.y);
        if (ab != 0.0) {
            phx = bpq.x / ab;
            phy = bpq.y / ab;
            const double zeta = (aqq - app) / (2.0 * ab);
            const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = t * c;
        }
    }
    // rows: p' = c B_p - sn ph B_q,  q' = sn B_p + c ph B_q
    const cd al = isp ? cd{c, 0.0} : cd{c * phx, c * phy};
    const cd be = isp ? cd{-sn * phx, -sn * phy} : cd{sn, 0.0};
    tab[j] = double4{al.x, al.y, be.x, be.y};
    __syncthreads();
    // columns: the coefficients of column c are conj(al_c), conj(be_c)
#pragma unroll
    for (int r = 0; r < M; ++r) {
        if (r & HS) continue;
        const int r2 = r ^ S;
        const double4 t1 = tab[j ^ r], t2 = tab[j ^ r2];
        const cd a1 = cd{t1.x, t1.y}, b1 = cd{t1.z, t1.w}, a2 = cd{t2.x, t2.y}, b2 = cd{t2.z, t2.w};
        const cd x = a[r], y = a[r2];
        a[r] = lin2c(a1, x, b1, y);
        a[r2] = lin2c(a2, y, b2, x);
        if constexpr (WANTV) {
            const cd vx = v[r], vy = v[r2];
            v[r] = lin2c(a1, vx, b1, vy);
            v[r2] = lin2c(a2, vy, b2, vx);
            // computed here: sunk to its next use, it would keep the table
            // entries live across the rest of the round
            asm volatile("" : "+v"(v[r].x), "+v"(v[r].y), "+v"(v[r2].x), "+v"(v[r2].y));
        }
    }
    // rows: partner register r^S into register r
#pragma unroll
    for (int r = 0; r < M; ++r) {
        if (r & HS) continue;
        const int r2 = r ^ S;
        const cd xr = swz_xor<S>(a[r2]), xr2 = swz_xor<S>(a[r]);
        a[r] = lin2(al, a[r], be, xr);
        a[r2] = lin2(al, a[r2], be, xr2);
    }
    // a rotated pair's off-diagonal entry is zero and its diagonal real by
    // construction: store them exactly
    if (sn != 0.0) {
        a[S] = cd{0.0, 0.0};
        a[0].y = 0.0;
    }
    __syncthreads();  // the table is rewritten by the next round
    // keep the rounds apart: interleaving them only raises register pressure
    __builtin_amdgcn_sched_barrier(0);
}

template <int M, bool WANTV, int S>
__device__ __forceinline__ void eig_sweep(cd (&a)[M], cd (&v)[M], int j, int m, bool conv, double4* tab)
{
    eig_round<M, WANTV, S>(a, v, j, m, conv, tab);
    if constexpr (S + 1 < M) eig_sweep<M, WANTV, S + 1>(a, v, j, m, conv, tab);
}

template <int M>
__device__ __forceinline__ double slot_sum(double x)
{
#pragma unroll
    for (int o = 1; o < M; o <<= 1) x += __shfl_xor(x, o);
    return x;
}

template <int M, bool WANTV>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_herm_eig_slots(
    const cd* __restrict__ X, int ldx, int off, int m, const cd* __restrict__ add, int nadd, double* __restrict__ w,
    cd* __restrict__ V, int batch)
{
    constexpr int NS = 64 / M;  // matrices per wavefront
    __shared__ double4 tab[64];
    __shared__ double wl[64];
    __shared__ int rk[64];
    const int lane = threadIdx.x;
    const int j = lane % M, slot = lane / M;
    const int b = blockIdx.x * NS + slot;
    const bool live = b < batch && j < m;
    cd a[M], v[M];
    const cd* Xb = X + (size_t)(b < batch ? b : 0) * ldx * ldx;
#pragma unroll
    for (int r = 0; r < M; ++r) {
        const int c = j ^ r;
        cd h = cd{0.0, 0.0};
        if (live && c < m) {
            cd x = Xb[(size_t)(off + j) + (size_t)(off + c) * ldx];
            cd xt = Xb[(size_t)(off + c) + (size_t)(off + j) * ldx];
            if (add && j < nadd && c < nadd) {
                x = cadd(x, add[j + c * nadd]);
                xt = cadd(xt, add[c + j * nadd]);
            }
            // Hermitian part (A + A^H)/2
            h = cd{0.5 * (x.x + xt.x), 0.5 * (x.y - xt.y)};
        }
        a[r] = h;
        v[r] = cd{r == 0 && live ? 1.0 : 0.0, 0.0};
    }
    double nrm = 0.0;
#pragma unroll
    for (int r = 0; r < M; ++r) nrm += a[r].x * a[r].x + a[r].y * a[r].y;
    nrm = slot_sum<M>(nrm);
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off2 = 0.0;
#pragma unroll
        for (int r = 1; r < M; ++r) off2 += a[r].x * a[r].x + a[r].y * a[r].y;
        off2 = slot_sum<M>(off2);
        const bool conv = off2 <= 1e-34 * nrm || off2 == 0.0;
        if (__all(conv)) break;
        eig_sweep<M, WANTV, 1>(a, v, j, m, conv, tab + slot * M);
    }
    // stable ascending rank; NaN sorts last so the ranks stay a permutation
    const double wj = a[0].x;
    wl[lane] = wj;
    __syncthreads();
    int rank = 0;
    const bool nj = wj != wj;
    for (int i = 0; i < m; ++i) {
        const double wi = wl[slot * M + i];
        const bool ni = wi != wi;
        rank += ni ? (nj && i < j) : (nj || wi < wj || (wi == wj && i < j));
    }
    rk[lane] = rank;
    __syncthreads();
    if (!live) return;
    w[(size_t)b * m + rank] = wj;
    if constexpr (WANTV) {
        cd* Vb = V + (size_t)b * m * m;
#pragma unroll
        for (int r = 0; r < M; ++r) {
            const int c = j ^ r;
            if (c < m) Vb[j + (size_t)rk[slot * M + c] * m] = v[r];
        }
    }
}

template <int M>
static hipError_t launch_slots(const cd* X, int ldx, int off, int m, const cd* add, int nadd, double* w, cd* V,
                               int batch, hipStream_t s)
{
    const int nb = (batch + 64 / M - 1) / (64 / M);
    if (V)
        hipLaunchKernelGGL((k_herm_eig_slots<M, true>), dim3(nb), dim3(64), 0, s, X, ldx, off, m, add, nadd, w, V, batch);
    else
        hipLaunchKernelGGL((k_herm_eig_slots<M, false>), dim3(nb), dim3(64), 0, s, X, ldx, off, m, add, nadd, w, V,
                           batch);
    return hipGetLastError();
}

hipError_t launch_herm_eig(const cd* X, int ldx, int off, int m, const cd* add, int nadd, double* w, cd* V, int batch,
                           hipStream_t s)
{
    if (batch <= 0 || m <= 0) return hipSuccess;
    if (m <= 4) return launch_slots<4>(X, ldx, off, m, add, nadd, w, V, batch, s);
    if (m <= 8) return launch_slots<8>(X, ldx, off, m, add, nadd, w, V, batch, s);
    if (m <= 16) return launch_slots<16>(X, ldx, off, m, add, nadd, w, V, batch, s);
    if (m > 32) return hipErrorInvalidValue;  // the pair table holds 16 rotations
    size_t shm = sizeof(cd) * 2 * (size_t)m * m + sizeof(double) * 4 * 16 + sizeof(int) * 2 * 16;
    hipLaunchKernelGGL(k_herm_eig, dim3(batch), dim3(64), shm, s, X, ldx, off, m, add, nadd, w, V, batch);
    return hipGetLastError();
}

}  // namespace dg
