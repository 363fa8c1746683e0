// Stored-matrix (CSR) projection operators and the deterministic CSR -> CSC
// transpose used for atomic-free backprojection.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "vir_common.h"

namespace vir {

// G lanes per row: strided loads of (ind, len), f64 partial sums, shuffle
// reduction.  G divides the wavefront (64).
template <int G>
__global__ __launch_bounds__(256) void k_csr_fp(const int32_t* __restrict__ ind, const float* __restrict__ len,
                                               const int64_t* __restrict__ off, int64_t nrows,
                                               const float* __restrict__ vol, float* __restrict__ sino,
                                               int accumulate) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t row = gid / G;
    int sub = (int)(gid % G);
    double s = 0.0;
    if (row < nrows) {
        int64_t e0 = off[row], e1 = off[row + 1];
        for (int64_t e = e0 + sub; e < e1; e += G) {
            int32_t v = ind[e];
            if (v >= 0) s += (double)(vol[v] * len[e]);
        }
    }
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) s += __shfl_xor(s, m, G);
    if (row < nrows && sub == 0) {
        float r = (float)s;
        if (accumulate) r += sino[row];
        sino[row] = r;
    }
}

// One lane per row, f32 sums in entry order: the rounding sequence of
// forward_proj_ravel_rays_struct (projection/projection_c.c:101-157).
__global__ void k_csr_fp_seq(const int32_t* __restrict__ ind, const float* __restrict__ len,
                             const int64_t* __restrict__ off, int64_t nrows, const float* __restrict__ vol,
                             float* __restrict__ sino, int accumulate) {
    int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nrows) return;
    float s = accumulate ? sino[row] : 0.0f;
    for (int64_t e = off[row]; e < off[row + 1]; ++e) {
        int32_t v = ind[e];
        if (v >= 0) s += vol[v] * len[e];
    }
    sino[row] = s;
}

__global__ void k_keys(const int32_t* __restrict__ ind, int64_t nnz, int32_t sentinel, int32_t* __restrict__ key,
                       int32_t* __restrict__ val) {
    int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    int32_t v = ind[e];
    key[e] = (v >= 0 && v < sentinel) ? v : sentinel;
    val[e] = (int32_t)e;
}

// colptr[v] = first position of a key >= v in the sorted keys (lower bound)
__global__ void k_colptr(const int32_t* __restrict__ skey, int64_t nnz, int64_t nvox, int64_t* __restrict__ colptr) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v > nvox) return;
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if ((int64_t)skey[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    colptr[v] = lo;
}

__global__ void k_gather_csc(const int32_t* __restrict__ sval, const int64_t* __restrict__ off, int64_t nrows,
                             const float* __restrict__ len, int64_t n, int32_t* __restrict__ row,
                             float* __restrict__ clen) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t e = sval[i];
    // row of entry e: last r with off[r] <= e
    int64_t lo = 0, hi = nrows;
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) >> 1;
        if (off[mid] <= e) lo = mid;
        else hi = mid;
    }
    row[i] = (int32_t)lo;
    clen[i] = len[e];
}

template <int G>
__global__ __launch_bounds__(256) void k_csc_bp(const int64_t* __restrict__ colptr, const int32_t* __restrict__ row,
                                               const float* __restrict__ clen, int64_t nvox,
                                               const float* __restrict__ sino, float* __restrict__ vol,
                                               int accumulate) {
    int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t v = gid / G;
    int sub = (int)(gid % G);
    double s = 0.0;
    if (v < nvox) {
        for (int64_t e = colptr[v] + sub; e < colptr[v + 1]; e += G) s += (double)(sino[row[e]] * clen[e]);
    }
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) s += __shfl_xor(s, m, G);
    if (v < nvox && sub == 0) {
        float r = (float)s;
        if (accumulate) r += vol[v];
        vol[v] = r;
    }
}

// One lane per voxel, f32 sums in (ray, entry) order: the rounding sequence of
// back_proj_ravel_rays_struct (projection/projection_c.c:176-208).
__global__ void k_csc_bp_seq(const int64_t* __restrict__ colptr, const int32_t* __restrict__ row,
                             const float* __restrict__ clen, int64_t nvox, const float* __restrict__ sino,
                             float* __restrict__ vol, int accumulate) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    float s = accumulate ? vol[v] : 0.0f;
    for (int64_t e = colptr[v]; e < colptr[v + 1]; ++e) s += sino[row[e]] * clen[e];
    vol[v] = s;
}

static inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace vir

using namespace vir;

extern "C" {

int vir_csr_fp(const int32_t* ind, const float* len, const int64_t* off, int64_t nrays, const float* vol,
               float* sino, uint32_t flags, void* stream) {
    if (nrays < 0) return fail_arg("nrays < 0");
    if (nrays == 0) return VIR_OK;
    if (!ind || !len || !off || !vol || !sino) return fail_arg("NULL pointer argument");
    hipStream_t st = as_stream(stream);
    int acc = (flags & VIR_ACCUMULATE) ? 1 : 0;
    if (flags & VIR_SEQUENTIAL) {
        hipLaunchKernelGGL(k_csr_fp_seq, dim3(nblk(nrays, 256)), dim3(256), 0, st, ind, len, off, nrays, vol, sino,
                           acc);
    } else {
        hipLaunchKernelGGL(k_csr_fp<16>, dim3(nblk(nrays * 16, 256)), dim3(256), 0, st, ind, len, off, nrays, vol,
                           sino, acc);
    }
    return check_hip(hipGetLastError(), "vir_csr_fp launch");
}

int vir_csr_transpose(const int32_t* ind, const float* len, const int64_t* off, int64_t nrays, int64_t nnz,
                      int64_t nvox, int64_t* colptr, int32_t* row, float* clen, void* stream) {
    if (nrays < 0 || nnz < 0 || nvox < 1) return fail_arg("bad sizes");
    if (nnz >= (int64_t)1 << 31) return fail_arg("nnz must be < 2^31 for the stored-matrix transpose");
    if (nvox >= (int64_t)1 << 31) return fail_arg("nvox must be < 2^31");
    if (!colptr || (nnz > 0 && (!ind || !len || !off || !row || !clen))) return fail_arg("NULL pointer argument");
    hipStream_t st = as_stream(stream);
    int rc;
    if (nnz == 0) {
        rc = check_hip(hipMemsetAsync(colptr, 0, sizeof(int64_t) * (nvox + 1), st), "memset colptr");
        if (rc) return rc;
        return check_hip(hipStreamSynchronize(st), "sync");
    }
    int32_t *k_in = nullptr, *k_out = nullptr, *v_in = nullptr, *v_out = nullptr;
    void* tmp = nullptr;
    size_t tmpsz = 0;
    int bits = 1;
    while (((int64_t)1 << bits) <= nvox) ++bits;
    rc = check_hip(hipMalloc(&k_in, sizeof(int32_t) * nnz * 4), "hipMalloc transpose buffers");
    if (rc) return rc;
    k_out = k_in + nnz;
    v_in = k_out + nnz;
    v_out = v_in + nnz;
    hipLaunchKernelGGL(k_keys, dim3(nblk(nnz, 256)), dim3(256), 0, st, ind, nnz, (int32_t)nvox, k_in, v_in);
    rc = check_hip(rocprim::radix_sort_pairs(nullptr, tmpsz, k_in, k_out, v_in, v_out, (size_t)nnz, 0, bits, st),
                   "radix_sort_pairs size");
    if (!rc) rc = check_hip(hipMalloc(&tmp, tmpsz ? tmpsz : 1), "hipMalloc sort temp");
    if (!rc)
        rc = check_hip(rocprim::radix_sort_pairs(tmp, tmpsz, k_in, k_out, v_in, v_out, (size_t)nnz, 0, bits, st),
                       "radix_sort_pairs");
    if (!rc) {
        hipLaunchKernelGGL(k_colptr, dim3(nblk(nvox + 1, 256)), dim3(256), 0, st, k_out, nnz, nvox, colptr);
        hipLaunchKernelGGL(k_gather_csc, dim3(nblk(nnz, 256)), dim3(256), 0, st, v_out, off, nrays, len, nnz, row,
                           clen);
        rc = check_hip(hipGetLastError(), "transpose kernels");
    }
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "transpose sync");
    if (tmp) hipFree(tmp);
    hipFree(k_in);
    return rc;
}

int vir_csc_bp(const int64_t* colptr, const int32_t* row, const float* clen, int64_t nvox, const float* sino,
               float* vol, uint32_t flags, void* stream) {
    if (nvox < 0) return fail_arg("nvox < 0");
    if (nvox == 0) return VIR_OK;
    if (!colptr || !row || !clen || !sino || !vol) return fail_arg("NULL pointer argument");
    hipStream_t st = as_stream(stream);
    int acc = (flags & VIR_ACCUMULATE) ? 1 : 0;
    if (flags & VIR_SEQUENTIAL) {
        hipLaunchKernelGGL(k_csc_bp_seq, dim3(nblk(nvox, 256)), dim3(256), 0, st, colptr, row, clen, nvox, sino, vol,
                           acc);
    } else {
        hipLaunchKernelGGL(k_csc_bp<8>, dim3(nblk(nvox * 8, 256)), dim3(256), 0, st, colptr, row, clen, nvox, sino,
                           vol, acc);
    }
    return check_hip(hipGetLastError(), "vir_csc_bp launch");
}

}  // extern "C"
