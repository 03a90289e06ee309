// Device-side helpers shared by the lattice, Fourier and moment kernels.
// Complex numbers are plain {re, im} pairs of doubles (16-byte aligned so one
// element moves with a single 128-bit load/store); no std::complex on device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dg {

struct __align__(16) cd {
    double x, y;
};
struct __align__(8) cf {
    float x, y;
};

// Matches the reference's pi (dmft_common.h:43).
constexpr double kPi = 3.1415926535897932384626;

__device__ __forceinline__ cd mk(double x, double y) { return cd{x, y}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cd cneg(cd a) { return cd{-a.x, -a.y}; }
__device__ __forceinline__ cd cconj(cd a) { return cd{a.x, -a.y}; }
__device__ __forceinline__ cd cscale(cd a, double s) { return cd{a.x * s, a.y * s}; }
__device__ __forceinline__ cd cmul(cd a, cd b) { return cd{fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x)}; }
// a * conj(b)
__device__ __forceinline__ cd cmulc(cd a, cd b) { return cd{fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -a.x * b.y)}; }
// acc + a*b
__device__ __forceinline__ cd cfma(cd a, cd b, cd acc)
{
    return cd{fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y))};
}
// acc - a*b
__device__ __forceinline__ cd cfms(cd a, cd b, cd acc)
{
    return cd{fma(-a.x, b.x, fma(a.y, b.y, acc.x)), fma(-a.x, b.y, fma(-a.y, b.x, acc.y))};
}
// acc + conj(a)*b
__device__ __forceinline__ cd cfmac(cd a, cd b, cd acc)
{
    return cd{fma(a.x, b.x, fma(a.y, b.y, acc.x)), fma(a.x, b.y, fma(-a.y, b.x, acc.y))};
}
// 1/q for finite q > 0 away from the denormal range: hardware reciprocal
// estimate refined by two Newton-Raphson steps r <- r + r (1 - q r), each of
// which doubles the correct bits (~1 ulp result, 5 instructions instead of the
// IEEE division sequence).
__device__ __forceinline__ double rcp_pos(double q)
{
    double r = __builtin_amdgcn_rcp(q);
    r = fma(r, fma(-q, r, 1.0), r);
    return fma(r, fma(-q, r, 1.0), r);
}
__device__ __forceinline__ cd crcp(cd a)
{
    double d = 1.0 / fma(a.x, a.x, a.y * a.y);
    return cd{a.x * d, -a.y * d};
}
__device__ __forceinline__ cd csel(bool c, cd a, cd b) { return c ? a : b; }

// float complex
__device__ __forceinline__ cf fmk(float x, float y) { return cf{x, y}; }
__device__ __forceinline__ cf fmul(cf a, cf b) { return cf{fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x)}; }
__device__ __forceinline__ cf ffms(cf a, cf b, cf acc)
{
    return cf{fmaf(-a.x, b.x, fmaf(a.y, b.y, acc.x)), fmaf(-a.x, b.y, fmaf(-a.y, b.x, acc.y))};
}
// acc + a*b
__device__ __forceinline__ cf ffma(cf a, cf b, cf acc)
{
    return cf{fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y))};
}
__device__ __forceinline__ cf frcp(cf a)
{
    float d = __frcp_rn(fmaf(a.x, a.x, a.y * a.y));
    return cf{a.x * d, -a.y * d};
}

// Orders LDS traffic between lanes of one wavefront (lock-step execution, LDS
// operations of a wave complete in issue order; the fences stop the compiler
// from moving loads/stores across the point).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ cd ld(const cd* p) { return *p; }

// Per-lane sums in LDS: an LDS add without return (ds_add_f64) instead of a
// read, add and write.  Nothing waits for the old value, so a sum does not
// stall the wave; each lane owns its entries, so the adds land in program
// order and the result is the one of the plain read-add-write.
__device__ __forceinline__ void lds_add(double* p, double v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(cd* p, cd v)
{
    lds_add(&p->x, v.x);
    lds_add(&p->y, v.y);
}

// Matsubara frequency w_n = pi(2n+1)/beta (GreenFtn.cpp:219).
__device__ __forceinline__ double matsubara(int n, double beta) { return kPi * (2.0 * n + 1.0) / beta; }

// Sigma(iw_n) element (i,j) for frequency index nw: SelfE_w for nw < n_freq,
// otherwise the moment expansion m0 + m1/iw + m2/iw^2 + m3/iw^3
// (GreenFtn.cpp:224-240).  Column-major nc x nc.
__device__ __forceinline__ cd sigma_elem(const cd* __restrict__ sigma, const cd* __restrict__ mom, int nc, int n_freq,
                                         double beta, int nw, int i, int j)
{
    const int e = i + j * nc;
    if (nw < n_freq) return sigma[(size_t)nw * nc * nc + e];
    const double w = matsubara(nw, beta);
    // 1/iw = -i/w, 1/iw^2 = -1/w^2, 1/iw^3 = i/w^3
    const double r1 = 1.0 / w, r2 = r1 * r1, r3 = r2 * r1;
    const size_t m2 = (size_t)nc * nc;
    cd m0 = mom[e], m1 = mom[m2 + e], mm2 = mom[2 * m2 + e], m3 = mom[3 * m2 + e];
    cd v = m0;
    v = cadd(v, cd{m1.y * r1, -m1.x * r1});
    v = cadd(v, cd{-mm2.x * r2, -mm2.y * r2});
    v = cadd(v, cd{-m3.y * r3, m3.x * r3});
    return v;
}

}  // namespace dg
