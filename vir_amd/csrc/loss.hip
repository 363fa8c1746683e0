// Loss terms for gradient-based reconstruction, each value reduced
// deterministically (per-block partial sums in double into `work`, then one
// block sums them in order).
//
// Least squares:  r = a - b,  value = 0.5 * sum r^2.
//
// Smoothed isotropic total variation and its gradient, fused in one pass:
//
//   TV(x) = sum_i ( sqrt(|D x|_i^2 + eps^2) - eps ),
//   (D x)_i = forward differences along x, y, z, zero at the last index,
//
// over volumes (nbatch, nx, ny, nz), z fastest.  The gradient at voxel i is
//   -(dx_i + dy_i + dz_i)/n_i + dx_{i-ex}/n_{i-ex} + dy_{i-ey}/n_{i-ey}
//   + dz_{i-ez}/n_{i-ez},
// each voxel recomputing its three lower neighbours' norms (the stencil
// reads stay in L2; one HBM read of x and one write of the gradient).
#include <cmath>

#include "vir_common.h"

namespace vir {

constexpr int kTvBlocks = 2048;

struct TvShape {
    int64_t n;  // nbatch * nx * ny * nz
    int nx, ny, nz;
    float eps, weight;
};

__device__ __forceinline__ float tv_norm(const float* __restrict__ x, const TvShape& S, int64_t i, int ix,
                                         int iy, int iz, float& dx, float& dy, float& dz) {
    const float v = x[i];
    const int64_t sy = S.nz, sx = (int64_t)S.ny * S.nz;
    dx = ix + 1 < S.nx ? x[i + sx] - v : 0.0f;
    dy = iy + 1 < S.ny ? x[i + sy] - v : 0.0f;
    dz = iz + 1 < S.nz ? x[i + 1] - v : 0.0f;
    return sqrtf(dx * dx + dy * dy + dz * dz + S.eps * S.eps);
}

__device__ __forceinline__ void block_partial(double part, double* __restrict__ work) {
    __shared__ double red[4];
    for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) work[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_tv(TvShape S, const float* __restrict__ x, float* __restrict__ grad,
                                            int accumulate, double* __restrict__ work) {
    double part = 0.0;
    const int64_t sy = S.nz, sx = (int64_t)S.ny * S.nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < S.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int iz = (int)(i % S.nz);
        const int iy = (int)((i / S.nz) % S.ny);
        const int ix = (int)((i / sx) % S.nx);
        float dx, dy, dz;
        const float n = tv_norm(x, S, i, ix, iy, iz, dx, dy, dz);
        part += (double)(n - S.eps);
        if (grad) {
            float g = -(dx + dy + dz) / n;
            float ax, ay, az;
            if (ix > 0) {
                const float m = tv_norm(x, S, i - sx, ix - 1, iy, iz, ax, ay, az);
                g += ax / m;
            }
            if (iy > 0) {
                const float m = tv_norm(x, S, i - sy, ix, iy - 1, iz, ax, ay, az);
                g += ay / m;
            }
            if (iz > 0) {
                const float m = tv_norm(x, S, i - 1, ix, iy, iz - 1, ax, ay, az);
                g += az / m;
            }
            g *= S.weight;
            grad[i] = accumulate ? grad[i] + g : g;
        }
    }
    if (work) block_partial(part, work);
}

__global__ __launch_bounds__(256) void k_sqdist(const float* __restrict__ a, const float* __restrict__ b,
                                                float* __restrict__ r, int64_t n, double* __restrict__ work) {
    double part = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float d = a[i] - b[i];
        if (r) r[i] = d;
        part += (double)d * (double)d;
    }
    if (work) block_partial(part, work);
}

__global__ __launch_bounds__(256) void k_tv_finish(const double* __restrict__ work, int nblk, double scale,
                                                   double* __restrict__ value) {
    __shared__ double red[256];
    double s = 0.0;
    for (int k = threadIdx.x; k < nblk; k += 256) s += work[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) value[0] = red[0] * scale;
}

}  // namespace vir

using namespace vir;

extern "C" {

int vir_tv(const float* x, int64_t nbatch, int32_t nx, int32_t ny, int32_t nz, float eps, float weight,
           float* grad, double* value, double* work, uint32_t flags, void* stream) {
    if (nbatch < 0 || nx < 1 || ny < 1 || nz < 1) return fail_arg("bad volume shape");
    if (!(eps > 0.0f)) return fail_arg("eps must be > 0");
    if (value && !work) return fail_arg("value needs a work buffer of VIR_TV_WORK doubles");
    const int64_t n = nbatch * nx * ny * nz;
    hipStream_t st = as_stream(stream);
    if (n == 0) {
        if (value) {
            hipLaunchKernelGGL(k_tv_finish, dim3(1), dim3(256), 0, st, work, 0, 0.0, value);
            return check_hip(hipGetLastError(), "vir_tv launch");
        }
        return VIR_OK;
    }
    if (!x) return fail_arg("NULL pointer argument");
    TvShape S{n, nx, ny, nz, eps, weight};
    const int64_t want = (n + 255) / 256;
    const int nblk = (int)(want < kTvBlocks ? want : kTvBlocks);
    const int acc = (flags & VIR_ACCUMULATE) ? 1 : 0;
    hipLaunchKernelGGL(k_tv, dim3(nblk), dim3(256), 0, st, S, x, grad, acc, value ? work : nullptr);
    int rc = check_hip(hipGetLastError(), "vir_tv launch");
    if (rc || !value) return rc;
    hipLaunchKernelGGL(k_tv_finish, dim3(1), dim3(256), 0, st, work, nblk, (double)weight, value);
    return check_hip(hipGetLastError(), "vir_tv launch");
}

int vir_sqdist(const float* a, const float* b, float* r, int64_t n, double* value, double* work,
               void* stream) {
    if (n < 0) return fail_arg("n must be >= 0");
    if (value && !work) return fail_arg("value needs a work buffer of VIR_TV_WORK doubles");
    if (n > 0 && (!a || !b)) return fail_arg("NULL pointer argument");
    hipStream_t st = as_stream(stream);
    int nblk = 0;
    if (n > 0) {
        const int64_t want = (n + 255) / 256;
        nblk = (int)(want < kTvBlocks ? want : kTvBlocks);
        hipLaunchKernelGGL(k_sqdist, dim3(nblk), dim3(256), 0, st, a, b, r, n, value ? work : nullptr);
        int rc = check_hip(hipGetLastError(), "vir_sqdist launch");
        if (rc) return rc;
    }
    if (!value) return VIR_OK;
    hipLaunchKernelGGL(k_tv_finish, dim3(1), dim3(256), 0, st, work, nblk, 0.5, value);
    return check_hip(hipGetLastError(), "vir_sqdist launch");
}

}  // extern "C"
