// Ellipse / ellipsoid phantoms evaluated on the device (synthetic inputs in
// the style of the reference's phantoms.py).
//
// Voxel (i, j, k) of an (nx, ny, nz) grid sits at FOV-normalised coordinates
//   p = ((i + 0.5)/nx * 2 - 1, (j + 0.5)/ny * 2 - 1, (k + 0.5)/nz * 2 - 1)
// and takes sum_e rho_e * [ |((p - c_e) R_e) / a_e|^2 <= 1 ].  Each ellipsoid
// is 16 floats: centre (3), semi-axes (3), row-major rotation R (9), rho.
// For 2-D (nz == 1) only x and y enter; the rotation's upper-left 2x2 is
// used and the z axis length is ignored.
#include "vir_common.h"

namespace vir {

__global__ __launch_bounds__(256) void k_phantom(int nx, int ny, int nz, int nell, const float* __restrict__ prm,
                                                 float* __restrict__ out) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int k = (int)(t % nz);
    const int j = (int)((t / nz) % ny);
    const int i = (int)(t / ((int64_t)nz * ny));
    const float px = ((float)i + 0.5f) / (float)nx * 2.0f - 1.0f;
    const float py = ((float)j + 0.5f) / (float)ny * 2.0f - 1.0f;
    const float pz = nz > 1 ? ((float)k + 0.5f) / (float)nz * 2.0f - 1.0f : 0.0f;
    float acc = 0.0f;
    for (int e = 0; e < nell; ++e) {
        const float* q = prm + 16 * e;
        const float dx = px - q[0], dy = py - q[1], dz = nz > 1 ? pz - q[2] : 0.0f;
        // row vector (dx, dy, dz) times R
        const float r0 = dx * q[6] + dy * q[9] + dz * q[12];
        const float r1 = dx * q[7] + dy * q[10] + dz * q[13];
        const float r2 = dx * q[8] + dy * q[11] + dz * q[14];
        const float u = r0 / q[3], v = r1 / q[4], w = nz > 1 ? r2 / q[5] : 0.0f;
        if (u * u + v * v + w * w <= 1.0f) acc += q[15];
    }
    out[t] = acc;
}

}  // namespace vir

using namespace vir;

extern "C" {

int vir_phantom(int32_t nx, int32_t ny, int32_t nz, int32_t nell, const float* params, float* out,
                void* stream) {
    if (nx < 1 || ny < 1 || nz < 1 || nell < 0) return fail_arg("bad sizes");
    if (!out || (nell > 0 && !params)) return fail_arg("NULL pointer argument");
    const int64_t n = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(k_phantom, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), nx, ny, nz,
                       nell, params, out);
    return check_hip(hipGetLastError(), "vir_phantom launch");
}

}  // extern "C"
