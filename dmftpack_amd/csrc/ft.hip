// Matsubara tails and Fourier transforms (placeholder until the kernels land).
#include "api_internal.hpp"

extern "C" int dmftgpu_moments(dmftgpu_handle h, int, int, double, const double*, int, double*, int*)
{
    return dgi::fail(h, DMFTGPU_ERR_UNSUPPORTED, "moments not built yet");
}
extern "C" int dmftgpu_ft_w2tau(dmftgpu_handle h, int, int, int, double, const double*, const double*, double*, int*)
{
    return dgi::fail(h, DMFTGPU_ERR_UNSUPPORTED, "ft_w2tau not built yet");
}
extern "C" int dmftgpu_ft_tau2w(dmftgpu_handle h, int, int, int, double, const double*, double*, int*)
{
    return dgi::fail(h, DMFTGPU_ERR_UNSUPPORTED, "ft_tau2w not built yet");
}
extern "C" int dmftgpu_density_matrix(dmftgpu_handle h, const double*, double, double*)
{
    return dgi::fail(h, DMFTGPU_ERR_UNSUPPORTED, "density_matrix not built yet");
}
