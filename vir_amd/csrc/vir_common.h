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

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace vir
