// Shared host/device helpers for libvir_amd: error state, launch helpers,
// grid-parameter construction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/vir_amd.h"
#include "siddon_ray.h"

namespace vir {

void set_error(const std::string& msg);
int fail_arg(const std::string& msg);
int check_hip(hipError_t e, const char* what);

// Block -> XCD placement: the dispatcher sends block b to XCD b % 8.  Give
// each XCD a contiguous range of logical blocks so neighbouring rays/bricks
// share one L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nblocks) {
    int64_t x = bid % 8, i = bid / 8;
    int64_t q = nblocks / 8, r = nblocks % 8;
    return x * q + (x < r ? x : r) + i;
}

int make_grid(const vir_grid* gp, GridK& g);

// SIRT steps fused into a projector's write-back (sys_mat/sirt_dd.py:51-72),
// with the arithmetic of the elementwise kernels vir_sirt_residual /
// vir_sirt_update, so fused and separate iterations give the same bits.
//   residual (b != NULL): out = (b - v) / w where w != 0, else v
//   update   (x != NULL): x = x + gamma * (w != 0 ? v / w : v); out unused
struct SirtEpi {
    const float* b;
    const float* w;
    float* x;
    float gamma;
};

__device__ __forceinline__ float sirt_residual(const SirtEpi& E, int64_t i, float v) {
    const float w = E.w[i];
    return w != 0.0f ? (E.b[i] - v) / w : v;
}

__device__ __forceinline__ void sirt_update(const SirtEpi& E, int64_t i, float u) {
    const float w = E.w[i];
    const float q = w != 0.0f ? u / w : u;
    E.x[i] = E.x[i] + E.gamma * q;
}

// The epilogue of a batch item whose per-item arrays (b, x) start `off`
// elements further on; the weights are shared by the batch.
__host__ __device__ inline SirtEpi sirt_item(const SirtEpi& E, int64_t off) {
    return SirtEpi{E.b ? E.b + off : nullptr, E.w, E.x ? E.x + off : nullptr, E.gamma};
}

// Stores v at out[i], or the fused SIRT step of v.
__device__ __forceinline__ void sirt_store(const SirtEpi& E, float* out, int64_t i, float v) {
    if (E.x) sirt_update(E, i, v);
    else out[i] = E.b ? sirt_residual(E, i, v) : v;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace vir
