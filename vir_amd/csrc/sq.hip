// Distance-driven projectors of the translational "square" geometry
// (sys_mat/dd.py:903-1313) for gfx950.
//
// A source translates along each of the four sides of a square field of view
// (nsrc_p positions along the side, nsrc_z in height); the flat detector sits
// opposite it.  For the side looking along -o with the parallel axis p, a
// voxel slice at orthogonal boundary o_l (the lower boundary of the slice,
// dd.py:981) is magnified by M = DSD / (DSO - o_l); voxel p-boundaries
// project to M*(p - src_p) + ssrc_p and z-boundaries to M*(z - src_z) +
// ssrc_z, overlapped with the detector bins (boundspace(nu, du, su+ssrc_p),
// boundspace(nv, dv, sv+ssrc_z)).  The four sides are the volume rotated by
// multiples of 90 degrees (dd.py:1003-1006, 1042-1045):
//   side 0: img[io, ip, iz]            side 1: img[n-1-ip, io, iz]
//   side 2: img[n-1-io, n-1-ip, iz]    side 3: img[ip, n-1-io, iz]
// Sinogram (nsrc_p, nsrc_z, nv, nu, 4).  The forward projection sums
// value * overlap_u * overlap_v (the reference scales it by
// d_pix^2/(du*dv), dd.py:930); the backprojection is its transpose, unscaled
// (dd.py:1104-1135).
//
// Forward: one thread per (source p, source z, v, u) computing the four sides
// (they share every overlap).  Backward: one thread per voxel visiting the
// bins of its four rotated positions.  Gathers, no atomics.
#include <cmath>

#include "vir_common.h"

namespace vir {

struct SQGeo {
    int n, nz, nu, nv, nsp, nsz;
    float d, du, dv, dsp, dsz, su, sv, ssp, ssz, DSO, DSD;
    float out_scale;
};

// boundspace / censpace of sys_mat/pf.py:42-55 in float32
__device__ __forceinline__ float sq_bnd(int n, float d, float s, int k) {
    return (float)k * d + (s - d * ((float)n * 0.5f));
}
__device__ __forceinline__ float sq_cen(int n, float d, float s, int k) {
    return (float)k * d + (s - d * ((float)(n - 1) * 0.5f));
}

__device__ __forceinline__ float sq_ovl(float a0, float a1, float b0, float b1) {
    return fminf(a1, b1) - fmaxf(a0, b0);
}

__device__ __forceinline__ int64_t sq_vox(const SQGeo& G, int side, int io, int ip, int iz) {
    const int n = G.n;
    int ix, iy;
    switch (side) {
        case 0: ix = io; iy = ip; break;
        case 1: ix = n - 1 - ip; iy = io; break;
        case 2: ix = n - 1 - io; iy = n - 1 - ip; break;
        default: ix = ip; iy = n - 1 - io; break;
    }
    return ((int64_t)ix * n + iy) * G.nz + iz;
}

__global__ __launch_bounds__(256) void k_ddsq_fp(SQGeo G, const float* __restrict__ img, float* __restrict__ sino,
                                                int accumulate) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nsp * G.nsz * G.nv * G.nu) return;
    const int iu = (int)(t % G.nu);
    const int iv = (int)((t / G.nu) % G.nv);
    const int isz = (int)((t / ((int64_t)G.nu * G.nv)) % G.nsz);
    const int isp = (int)(t / ((int64_t)G.nu * G.nv * G.nsz));
    const float ul = sq_bnd(G.nu, G.du, G.su + G.ssp, iu), ur = sq_bnd(G.nu, G.du, G.su + G.ssp, iu + 1);
    const float vl = sq_bnd(G.nv, G.dv, G.sv + G.ssz, iv), vr = sq_bnd(G.nv, G.dv, G.sv + G.ssz, iv + 1);
    const float sp = sq_cen(G.nsp, G.dsp, G.ssp, isp), sz = sq_cen(G.nsz, G.dsz, G.ssz, isz);
    const float p0 = sq_bnd(G.n, G.d, 0.0f, 0), z0 = sq_bnd(G.nz, G.d, 0.0f, 0);
    const float inv_d = 1.0f / G.d;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int io = 0; io < G.n; ++io) {
        const float M = G.DSD / (G.DSO - sq_bnd(G.n, G.d, 0.0f, io));
        if (!(M > 0.0f)) continue;
        const float psrc = M * sp - G.ssp, zsrc = M * sz - G.ssz;
        // voxels whose projected interval can meet the bin (inverse, widened)
        const float a0 = ((ul + psrc) / M - p0) * inv_d, a1 = ((ur + psrc) / M - p0) * inv_d;
        const float b0 = ((vl + zsrc) / M - z0) * inv_d, b1 = ((vr + zsrc) / M - z0) * inv_d;
        if (!(a1 > -2.0f) || !(a0 < (float)G.n + 1.0f) || !(b1 > -2.0f) || !(b0 < (float)G.nz + 1.0f)) continue;
        const int ip0 = (int)fmaxf(floorf(a0) - 1.0f, 0.0f), ip1 = (int)fminf(floorf(a1) + 1.0f, (float)(G.n - 1));
        const int iz0 = (int)fmaxf(floorf(b0) - 1.0f, 0.0f), iz1 = (int)fminf(floorf(b1) + 1.0f, (float)(G.nz - 1));
        for (int iz = iz0; iz <= iz1; ++iz) {
            const float ovv = sq_ovl(M * sq_bnd(G.nz, G.d, 0.0f, iz) - zsrc, M * sq_bnd(G.nz, G.d, 0.0f, iz + 1) - zsrc,
                                     vl, vr);
            if (!(ovv > 0.0f)) continue;
            for (int ip = ip0; ip <= ip1; ++ip) {
                const float ovu = sq_ovl(M * sq_bnd(G.n, G.d, 0.0f, ip) - psrc,
                                         M * sq_bnd(G.n, G.d, 0.0f, ip + 1) - psrc, ul, ur);
                if (!(ovu > 0.0f)) continue;
                const float w = ovu * ovv;
#pragma unroll
                for (int f = 0; f < 4; ++f) acc[f] += (double)(img[sq_vox(G, f, io, ip, iz)] * w);
            }
        }
    }
    float* out = sino + t * 4;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        float r = (float)acc[f] * G.out_scale;
        out[f] = accumulate ? out[f] + r : r;
    }
}

__global__ __launch_bounds__(256) void k_ddsq_bp(SQGeo G, const float* __restrict__ sino, float* __restrict__ img,
                                                int accumulate) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.n * G.n * G.nz) return;
    const int iz = (int)(t % G.nz);
    const int iy = (int)((t / G.nz) % G.n);
    const int ix = (int)(t / ((int64_t)G.nz * G.n));
    const int n = G.n;
    const float u0 = sq_bnd(G.nu, G.du, G.su + G.ssp, 0), v0 = sq_bnd(G.nv, G.dv, G.sv + G.ssz, 0);
    const float inv_du = 1.0f / G.du, inv_dv = 1.0f / G.dv;
    const float zl = sq_bnd(G.nz, G.d, 0.0f, iz), zr = sq_bnd(G.nz, G.d, 0.0f, iz + 1);
    double acc = 0.0;
    for (int f = 0; f < 4; ++f) {
        int io, ip;
        switch (f) {
            case 0: io = ix; ip = iy; break;
            case 1: io = iy; ip = n - 1 - ix; break;
            case 2: io = n - 1 - ix; ip = n - 1 - iy; break;
            default: io = n - 1 - iy; ip = ix; break;
        }
        const float M = G.DSD / (G.DSO - sq_bnd(n, G.d, 0.0f, io));
        if (!(M > 0.0f)) continue;
        const float pl = sq_bnd(n, G.d, 0.0f, ip), pr = sq_bnd(n, G.d, 0.0f, ip + 1);
        for (int isp = 0; isp < G.nsp; ++isp) {
            const float psrc = M * sq_cen(G.nsp, G.dsp, G.ssp, isp) - G.ssp;
            const float a = M * pl - psrc, b = M * pr - psrc;
            if (!(b > u0 - G.du) || !((a - u0) * inv_du < (float)G.nu + 1.0f)) continue;
            const int ju0 = (int)fmaxf(floorf((a - u0) * inv_du) - 1.0f, 0.0f);
            const int ju1 = (int)fminf(floorf((b - u0) * inv_du) + 1.0f, (float)(G.nu - 1));
            for (int isz = 0; isz < G.nsz; ++isz) {
                const float zsrc = M * sq_cen(G.nsz, G.dsz, G.ssz, isz) - G.ssz;
                const float c = M * zl - zsrc, e = M * zr - zsrc;
                if (!(e > v0 - G.dv) || !((c - v0) * inv_dv < (float)G.nv + 1.0f)) continue;
                const int jv0 = (int)fmaxf(floorf((c - v0) * inv_dv) - 1.0f, 0.0f);
                const int jv1 = (int)fminf(floorf((e - v0) * inv_dv) + 1.0f, (float)(G.nv - 1));
                const float* blk = sino + ((int64_t)isp * G.nsz + isz) * G.nv * G.nu * 4;
                for (int jv = jv0; jv <= jv1; ++jv) {
                    const float ovv = sq_ovl(c, e, sq_bnd(G.nv, G.dv, G.sv + G.ssz, jv),
                                             sq_bnd(G.nv, G.dv, G.sv + G.ssz, jv + 1));
                    if (!(ovv > 0.0f)) continue;
                    for (int ju = ju0; ju <= ju1; ++ju) {
                        const float ovu = sq_ovl(a, b, sq_bnd(G.nu, G.du, G.su + G.ssp, ju),
                                                 sq_bnd(G.nu, G.du, G.su + G.ssp, ju + 1));
                        if (!(ovu > 0.0f)) continue;
                        acc += (double)(blk[((int64_t)jv * G.nu + ju) * 4 + f] * (ovu * ovv));
                    }
                }
            }
        }
    }
    float r = (float)acc * G.out_scale;
    if (accumulate) r += img[t];
    img[t] = r;
}

static int sq_setup(const vir_ddsq_geom* g, SQGeo& G) {
    if (!g) return fail_arg("geometry is NULL");
    G.n = g->n; G.nz = g->nz; G.nu = g->nu; G.nv = g->nv; G.nsp = g->nsrc_p; G.nsz = g->nsrc_z;
    G.d = g->d_pix; G.du = g->du; G.dv = g->dv; G.dsp = g->dsrc_p; G.dsz = g->dsrc_z;
    G.su = g->su; G.sv = g->sv; G.ssp = g->ssrc_p; G.ssz = g->ssrc_z; G.DSO = g->DSO; G.DSD = g->DSD;
    if (G.n < 1 || G.nz < 1 || G.nu < 1 || G.nv < 1 || G.nsp < 0 || G.nsz < 0) return fail_arg("bad sizes");
    if (!(G.d > 0.0f) || !(G.du > 0.0f) || !(G.dv > 0.0f)) return fail_arg("spacings must be > 0");
    if (!(G.DSD > 0.0f)) return fail_arg("DSD must be > 0");
    return VIR_OK;
}

}  // namespace vir

using namespace vir;

extern "C" {

int vir_ddsq_fp(const vir_ddsq_geom* geom, const float* img, float* sino, float out_scale, uint32_t flags,
                void* stream) {
    SQGeo G;
    int rc = sq_setup(geom, G);
    if (rc) return rc;
    const int64_t n = (int64_t)G.nsp * G.nsz * G.nv * G.nu;
    if (n == 0) return VIR_OK;
    if (!img || !sino) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    hipLaunchKernelGGL(k_ddsq_fp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), G, img, sino,
                       (flags & VIR_ACCUMULATE) ? 1 : 0);
    return check_hip(hipGetLastError(), "vir_ddsq_fp launch");
}

int vir_ddsq_bp(const vir_ddsq_geom* geom, const float* sino, float* img, float out_scale, uint32_t flags,
                void* stream) {
    SQGeo G;
    int rc = sq_setup(geom, G);
    if (rc) return rc;
    if (!img || (G.nsp * G.nsz > 0 && !sino)) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    const int64_t n = (int64_t)G.n * G.n * G.nz;
    hipLaunchKernelGGL(k_ddsq_bp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), G, sino, img,
                       (flags & VIR_ACCUMULATE) ? 1 : 0);
    return check_hip(hipGetLastError(), "vir_ddsq_bp launch");
}

}  // extern "C"
