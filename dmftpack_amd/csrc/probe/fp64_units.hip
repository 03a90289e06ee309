// FP64 throughput of the MI355X matrix cores (v_mfma_f64_16x16x4) and of the
// vector FMA, alone and issued together by the same wavefronts.  Not part of
// the library: a measurement behind the choice of unit for the FP64 kernels.
//
//   hipcc -O3 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -o fp64_units dmftpack_amd/csrc/probe/fp64_units.hip && ./fp64_units
//
// Every kernel works in registers only (one store per lane at the end), 64
// threads per workgroup, grids of 1024..8192 workgroups (1..8 per SIMD).
#include <hip/hip_runtime.h>

#include <cstdio>

typedef double dv4 __attribute__((ext_vector_type(4)));

// per iteration: 4 independent MFMA accumulators (if M) and NV x 8 independent
// FMA chains, interleaved in program order
template <bool M, int NV>
__global__ __launch_bounds__(64) void k_units(double* out, int iters, double a0, double b0)
{
    dv4 acc[4];
    for (int c = 0; c < 4; ++c) acc[c] = dv4{0, 0, 0, 0};
    double v[8];
    for (int c = 0; c < 8; ++c) v[c] = c;
    const double a = a0 + threadIdx.x * 1e-3, b = b0 - threadIdx.x * 1e-3;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (M) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
#pragma unroll
            for (int q = 0; q < NV; ++q)
#pragma unroll
                for (int r = 0; r < 2; ++r) v[(2 * c + r) & 7] = __builtin_fma(a, v[(2 * c + r) & 7], b);
        }
    }
    double s = 0;
    for (int c = 0; c < 4; ++c) s += acc[c][0] + acc[c][1] + acc[c][2] + acc[c][3];
    for (int c = 0; c < 8; ++c) s += v[c];
    out[blockIdx.x * 64 + threadIdx.x] = s;
}

template <typename F>
static float time_ms(F launch)
{
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    launch();
    (void)hipDeviceSynchronize();
    (void)hipEventRecord(e0);
    launch();
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

int main()
{
    const int iters = 2000;
    double* out = nullptr;
    if (hipMalloc(&out, (size_t)8192 * 64 * sizeof(double)) != hipSuccess) return 1;
    printf("grid  waves/SIMD  mfma_ms  mfma_TF  fma_ms  fma_TF  both_ms  (4 MFMA + 64 FMA per iteration)\n");
    for (int g : {1024, 2048, 4096, 8192}) {
        const float tm = time_ms([&] { hipLaunchKernelGGL((k_units<true, 0>), dim3(g), dim3(64), 0, 0, out, iters, 1.0, 0.5); });
        const float tv = time_ms([&] { hipLaunchKernelGGL((k_units<false, 8>), dim3(g), dim3(64), 0, 0, out, iters, 1.0, 0.5); });
        const float tb = time_ms([&] { hipLaunchKernelGGL((k_units<true, 8>), dim3(g), dim3(64), 0, 0, out, iters, 1.0, 0.5); });
        const double fm = (double)g * iters * 4 * 16 * 16 * 4 * 2;  // MFMA FLOPs
        const double fv = (double)g * 64 * iters * 4 * 8 * 2 * 2;   // FMA FLOPs (8 x 2 per c, 4 c)
        printf("%5d  %10.0f  %7.3f  %7.1f  %6.3f  %6.1f  %7.3f\n", g, g / 1024.0, tm, fm / (tm * 1e-3) / 1e12, tv,
               fv / (tv * 1e-3) / 1e12, tb);
    }
    (void)hipFree(out);
    return 0;
}
