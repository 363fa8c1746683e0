// Elementwise SIRT steps (weighted residual, weighted update).
#include "vir_common.h"

namespace vir {

__global__ void k_sirt_residual(const float* __restrict__ b, float* __restrict__ r, const float* __restrict__ W,
                                int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SirtEpi E{b, W, nullptr, 0.0f};
    r[i] = sirt_residual(E, i, r[i]);
}

__global__ void k_sirt_update(float* __restrict__ x, const float* __restrict__ u, const float* __restrict__ W,
                              float gamma, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SirtEpi E{nullptr, W, x, gamma};
    sirt_update(E, i, u[i]);
}

}  // namespace vir

using namespace vir;

extern "C" {

int vir_sirt_residual(const float* b, float* r, const float* W, int64_t n, void* stream) {
    if (n < 0) return fail_arg("n < 0");
    if (n == 0) return VIR_OK;
    if (!b || !r || !W) return fail_arg("NULL pointer argument");
    hipLaunchKernelGGL(k_sirt_residual, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), b, r, W,
                       n);
    return check_hip(hipGetLastError(), "vir_sirt_residual launch");
}

int vir_sirt_update(float* x, const float* u, const float* W, float gamma, int64_t n, void* stream) {
    if (n < 0) return fail_arg("n < 0");
    if (n == 0) return VIR_OK;
    if (!x || !u || !W) return fail_arg("NULL pointer argument");
    hipLaunchKernelGGL(k_sirt_update, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), x, u, W,
                       gamma, n);
    return check_hip(hipGetLastError(), "vir_sirt_update launch");
}

}  // extern "C"
