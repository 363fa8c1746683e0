// Chambolle-Pock primal-dual iteration for TV-regularised reconstruction
// (tv.py:229-294), the three elementwise steps fused into kernels:
//
//   dual TV     z_c += sigma_z * (roll(ubar, -1, c) - ubar)   per axis c
//               z    = z / max(1, |z|)                         (projection onto
//                                                              the unit ball,
//                                                              the intent of
//                                                              pyT2, tv.py:24-95)
//   dual data   q    = (q + sigma_q * (A ubar - g)) / (1 + sigma_q * lam)
//   primal      u'   = u + tau * (div z - A^T q),  div z = sum_c z_c - roll(z_c, 1, c)
//               ubar = u' + theta * (u' - u)
//
// with periodic differences as the reference's np.roll (tv.py:207-224).
// Volumes (nx, ny, nz) C-order, nz = 1 for 2-D; z holds ndim planes of the
// volume's size.
#include "vir_common.h"

namespace vir {

struct PdShape {
    int nx, ny, nz, ndim;
    int64_t n;
};

__device__ __forceinline__ void pd_coords(const PdShape& S, int64_t i, int& x, int& y, int& z) {
    z = (int)(i % S.nz);
    y = (int)((i / S.nz) % S.ny);
    x = (int)(i / ((int64_t)S.nz * S.ny));
}

__device__ __forceinline__ int64_t pd_index(const PdShape& S, int x, int y, int z) {
    return ((int64_t)x * S.ny + y) * S.nz + z;
}

__global__ __launch_bounds__(256) void k_pdhg_dual_tv(PdShape S, const float* __restrict__ ubar,
                                                      float* __restrict__ zf, float sigma) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    int x, y, z;
    pd_coords(S, i, x, y, z);
    const float u0 = ubar[i];
    float g[3];
    g[0] = ubar[pd_index(S, x + 1 < S.nx ? x + 1 : 0, y, z)] - u0;
    g[1] = ubar[pd_index(S, x, y + 1 < S.ny ? y + 1 : 0, z)] - u0;
    g[2] = S.ndim == 3 ? ubar[pd_index(S, x, y, z + 1 < S.nz ? z + 1 : 0)] - u0 : 0.0f;
    float v[3], nn = 0.0f;
    for (int c = 0; c < S.ndim; ++c) {
        v[c] = zf[c * S.n + i] + sigma * g[c];
        nn += v[c] * v[c];
    }
    const float scale = nn > 1.0f ? 1.0f / sqrtf(nn) : 1.0f;
    for (int c = 0; c < S.ndim; ++c) zf[c * S.n + i] = v[c] * scale;
}

__global__ __launch_bounds__(256) void k_pdhg_dual_data(int64_t n, const float* __restrict__ Au,
                                                        const float* __restrict__ g,
                                                        const float* __restrict__ sigma_q, float lam,
                                                        float* __restrict__ q) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float s = sigma_q[i];
    q[i] = (q[i] + s * (Au[i] - g[i])) / (1.0f + s * lam);
}

__global__ __launch_bounds__(256) void k_pdhg_primal(PdShape S, const float* __restrict__ zf,
                                                     const float* __restrict__ ATq,
                                                     const float* __restrict__ tau, float theta,
                                                     float* __restrict__ u, float* __restrict__ ubar) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    int x, y, z;
    pd_coords(S, i, x, y, z);
    float dv = zf[i] - zf[pd_index(S, x > 0 ? x - 1 : S.nx - 1, y, z)];
    dv += zf[S.n + i] - zf[S.n + pd_index(S, x, y > 0 ? y - 1 : S.ny - 1, z)];
    if (S.ndim == 3) dv += zf[2 * S.n + i] - zf[2 * S.n + pd_index(S, x, y, z > 0 ? z - 1 : S.nz - 1)];
    const float u0 = u[i];
    const float u1 = u0 + tau[i] * (dv - ATq[i]);
    u[i] = u1;
    ubar[i] = u1 + theta * (u1 - u0);
}

static int pd_shape(int32_t nx, int32_t ny, int32_t nz, PdShape& S) {
    if (nx < 1 || ny < 1 || nz < 1) return fail_arg("bad volume shape");
    S.nx = nx; S.ny = ny; S.nz = nz;
    S.ndim = nz > 1 ? 3 : 2;
    S.n = (int64_t)nx * ny * nz;
    return VIR_OK;
}

}  // namespace vir

using namespace vir;

extern "C" {

int vir_pdhg_dual_tv(int32_t nx, int32_t ny, int32_t nz, const float* ubar, float* z, float sigma_z,
                     void* stream) {
    PdShape S;
    int rc = pd_shape(nx, ny, nz, S);
    if (rc) return rc;
    if (!ubar || !z) return fail_arg("NULL pointer argument");
    hipLaunchKernelGGL(k_pdhg_dual_tv, dim3((unsigned)((S.n + 255) / 256)), dim3(256), 0, as_stream(stream), S, ubar,
                       z, sigma_z);
    return check_hip(hipGetLastError(), "vir_pdhg_dual_tv launch");
}

int vir_pdhg_dual_data(int64_t n, const float* Au, const float* g, const float* sigma_q, float lam, float* q,
                       void* stream) {
    if (n < 0) return fail_arg("n must be >= 0");
    if (n == 0) return VIR_OK;
    if (!Au || !g || !sigma_q || !q) return fail_arg("NULL pointer argument");
    hipLaunchKernelGGL(k_pdhg_dual_data, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), n, Au,
                       g, sigma_q, lam, q);
    return check_hip(hipGetLastError(), "vir_pdhg_dual_data launch");
}

int vir_pdhg_primal(int32_t nx, int32_t ny, int32_t nz, const float* z, const float* ATq, const float* tau,
                    float theta, float* u, float* ubar, void* stream) {
    PdShape S;
    int rc = pd_shape(nx, ny, nz, S);
    if (rc) return rc;
    if (!z || !ATq || !tau || !u || !ubar) return fail_arg("NULL pointer argument");
    hipLaunchKernelGGL(k_pdhg_primal, dim3((unsigned)((S.n + 255) / 256)), dim3(256), 0, as_stream(stream), S, z,
                       ATq, tau, theta, u, ubar);
    return check_hip(hipGetLastError(), "vir_pdhg_primal launch");
}

}  // extern "C"
