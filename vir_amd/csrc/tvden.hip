// Total-variation denoising by Chambolle's projection algorithm
// (A. Chambolle, "An algorithm for total variation minimization and
// applications", J. Math. Imaging Vis. 20 (2004) 89-97, eq. (9)):
//
//   u = argmin_u ||u - g||^2 / 2 + w TV(u),  TV(u) = sum_i |(grad u)_i|,
//
// grad = forward differences along each axis (zero at the last index),
// div = -grad^T.  In the dual variable q = w p:
//
//   v = div q - g
//   q <- (q + tau grad v) / (1 + (tau / w) |grad v|)
//   u  = g - div q
//
// One iteration is two launches over the voxels (grid-stride, 256 threads):
// the residual v into a scratch image, then the dual step, in place on q
// (each q element is read and written by its own thread; v is not written).
// q holds one plane per axis, plane a at offset a * n; its values at an
// axis' last index are taken as zero.  The step also keeps the largest
// |q_new - q_old| per block for the host's stopping rule.
#include <cmath>

#include "vir_common.h"

namespace vir {

constexpr int kTvdBlocks = 2048;  // VIR_TV_WORK

struct TvdShape {
    int64_t n;           // voxels
    int n0, n1, n2;      // sizes along axes 0, 1, 2 (1 for missing axes)
    int64_t s0, s1;      // strides of axes 0 and 1 (axis 2: 1)
};

template <int ND>
__device__ __forceinline__ void tvd_coords(const TvdShape& S, int64_t j, int (&i)[3]) {
    i[2] = (int)(j % S.n2);
    const int64_t r = j / S.n2;
    i[1] = (int)(r % S.n1);
    i[0] = (int)(r / S.n1);
}

// v = sign (div q - g)
template <int ND>
__global__ __launch_bounds__(256) void k_tvd_residual(TvdShape S, const float* __restrict__ q,
                                                     const float* __restrict__ g, float sign,
                                                     float* __restrict__ v) {
    const int nn[3] = {S.n0, S.n1, S.n2};
    const int64_t st[3] = {S.s0, S.s1, 1};
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < S.n;
         j += (int64_t)gridDim.x * blockDim.x) {
        int i[3];
        tvd_coords<ND>(S, j, i);
        float d = 0.0f;
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            const float* __restrict__ qa = q + a * S.n;
            if (i[a] < nn[a] - 1) d += qa[j];
            if (i[a] > 0) d -= qa[j - st[a]];
        }
        v[j] = sign * (d - g[j]);
    }
}

template <int ND>
__global__ __launch_bounds__(256) void k_tvd_step(TvdShape S, const float* __restrict__ v, float* __restrict__ q,
                                                 float tau, float tau_w, float* __restrict__ dmax) {
    const int nn[3] = {S.n0, S.n1, S.n2};
    const int64_t st[3] = {S.s0, S.s1, 1};
    float m = 0.0f;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < S.n;
         j += (int64_t)gridDim.x * blockDim.x) {
        int i[3];
        tvd_coords<ND>(S, j, i);
        const float vj = v[j];
        float gr[ND];
        float n2 = 0.0f;
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            gr[a] = i[a] < nn[a] - 1 ? v[j + st[a]] - vj : 0.0f;
            n2 += gr[a] * gr[a];
        }
        const float den = 1.0f + tau_w * sqrtf(n2);
#pragma unroll
        for (int a = 0; a < ND; ++a) {
            float* __restrict__ qa = q + a * S.n;
            const float old = qa[j];
            const float nw = (old + tau * gr[a]) / den;
            qa[j] = nw;
            m = fmaxf(m, fabsf(nw - old));
        }
    }
    if (!dmax) return;
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) dmax[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

__global__ __launch_bounds__(256) void k_tvd_max(float* __restrict__ dmax, int nblk) {
    __shared__ float red[256];
    float m = 0.0f;
    for (int k = threadIdx.x; k < nblk; k += 256) m = fmaxf(m, dmax[k]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) dmax[0] = red[0];
}

template <int ND>
static int tvd_run(const TvdShape& S, const float* g, float* q, float* v, float* u, float weight, float tau,
                   int iters, float* dmax, hipStream_t st) {
    const int64_t want = (S.n + 255) / 256;
    const int nblk = (int)(want < kTvdBlocks ? want : kTvdBlocks);
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(k_tvd_residual<ND>, dim3(nblk), dim3(256), 0, st, S, q, g, 1.0f, v);
        float* dm = (dmax && it == iters - 1) ? dmax : nullptr;
        hipLaunchKernelGGL(k_tvd_step<ND>, dim3(nblk), dim3(256), 0, st, S, v, q, tau, tau / weight, dm);
    }
    if (dmax && iters > 0) hipLaunchKernelGGL(k_tvd_max, dim3(1), dim3(256), 0, st, dmax, nblk);
    if (u) hipLaunchKernelGGL(k_tvd_residual<ND>, dim3(nblk), dim3(256), 0, st, S, q, g, -1.0f, u);
    return check_hip(hipGetLastError(), "vir_tv_chambolle launch");
}

}  // namespace vir

using namespace vir;

extern "C" {

int vir_tv_chambolle(int32_t ndim, int32_t nx, int32_t ny, int32_t nz, const float* g, float* q, float* v,
                     float* u, float weight, float tau, int32_t iters, float* dmax, void* stream) {
    if (ndim < 1 || ndim > 3) return fail_arg("ndim must be 1, 2 or 3");
    if (nx < 1 || ny < 1 || nz < 1) return fail_arg("sizes must be >= 1");
    if ((ndim < 2 && ny != 1) || (ndim < 3 && nz != 1)) return fail_arg("sizes of missing axes must be 1");
    if (!(weight > 0.0f) || !std::isfinite(weight)) return fail_arg("weight must be > 0");
    if (!(tau > 0.0f) || !std::isfinite(tau)) return fail_arg("tau must be > 0");
    if (iters < 0) return fail_arg("iters must be >= 0");
    if (!g || !q || (iters > 0 && !v)) return fail_arg("NULL pointer argument");
    TvdShape S;
    S.n0 = nx; S.n1 = ny; S.n2 = nz;
    S.n = (int64_t)nx * ny * nz;
    S.s1 = nz;
    S.s0 = (int64_t)ny * nz;
    hipStream_t st = as_stream(stream);
    if (ndim == 1) return tvd_run<1>(S, g, q, v, u, weight, tau, iters, dmax, st);
    if (ndim == 2) return tvd_run<2>(S, g, q, v, u, weight, tau, iters, dmax, st);
    return tvd_run<3>(S, g, q, v, u, weight, tau, iters, dmax, st);
}

}  // extern "C"
