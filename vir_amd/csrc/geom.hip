// Circular-orbit ray endpoints generated on the device (proj_geom.py:61-148).
//
// Unrotated endpoints (view-independent):
//   par : src (-src_iso, y, z),  trg (det_iso, y, z)      y in DetsY, z in DetsZ or 0
//   fan : src (-src_iso, 0, z),  trg (det_iso, y, z)      z in DetsZ or 0
//   cone: src (-src_iso, 0, 0),  trg (det_iso, y, z)
// Each view rotates (x, y) by the row-vector product
//   [x y] @ [[c, s], [-s, c]]  =  (x c - y s,  x s + y c)
// evaluated in double the way numpy's matmul does it for an (N, 2) x (2, 2)
// product through OpenBLAS dgemm: the k = 0 product rounded, then the k = 1
// product fused onto it,
//   x' = fma(y, -s, RN(x c)),   y' = fma(y, c, RN(x s)),
// and finally rounded to float32 (the cast siddons() applies).  cos/sin are
// given per view in double (the host's libm values).  Layout: (nviews, nu
// [, nv], 3), z varying fastest, as geom_circular returns.
#include <cmath>

#include "vir_common.h"

namespace vir {

struct GeomArgs {
    int mode;  // 0 par, 1 fan, 2 cone
    int64_t nviews;
    int nu, nv;  // nv = 1 when there is no DetsZ
    int has_z;
    double src_iso, det_iso;
};

__device__ __forceinline__ void rot_store(double x, double y, double z, double c, double s, float* out) {
    const double xc = x * c, xs = x * s;
    const double xr = __builtin_fma(y, -s, xc);
    const double yr = __builtin_fma(y, c, xs);
    out[0] = (float)xr;
    out[1] = (float)yr;
    out[2] = (float)z;
}

__global__ __launch_bounds__(256) void k_geom_circular(GeomArgs A, const double* __restrict__ cs,
                                                       const double* __restrict__ sn,
                                                       const double* __restrict__ dy,
                                                       const double* __restrict__ dz, float* __restrict__ src,
                                                       float* __restrict__ trg) {
    const int64_t per_view = (int64_t)A.nu * A.nv;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per_view * A.nviews) return;
    const int64_t a = t / per_view;
    const int64_t r = t % per_view;
    const int iv = (int)(r % A.nv), iu = (int)(r / A.nv);
    const double c = cs[a], s = sn[a];
    const double y = dy[iu];
    const double z = A.has_z ? dz[iv] : 0.0;
    double sy, sz, tz;
    if (A.mode == 0) {
        sy = y; sz = z; tz = z;
    } else if (A.mode == 1) {
        sy = 0.0; sz = z; tz = z;
    } else {
        sy = 0.0; sz = 0.0; tz = z;
    }
    rot_store(-A.src_iso, sy, sz, c, s, src + 3 * t);
    rot_store(A.det_iso, y, tz, c, s, trg + 3 * t);
}

// Square (four-sided translational) geometry (proj_geom.py:12-57): bin b of
// side k at (b, -d), (-d, b), (-b, d), (d, -b), looking along angle a; the
// target adds the per-angle offsets (D, E) = (cos a d 5, sin a d 5) given in
// double by the host (numpy's values): side 0 (+D, +E), 1 (+E, +D),
// 2 (-D, -E), 3 (-E, -D), in double, then rounded to float32.  Layout
// (nbins, nangs, 4, 3).
__global__ __launch_bounds__(256) void k_geom_square(int64_t nbins, int64_t nangs, const double* __restrict__ bx,
                                                     const double* __restrict__ nbx, double d,
                                                     const double* __restrict__ D,
                                                     const double* __restrict__ E, float* __restrict__ src,
                                                     float* __restrict__ trg) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbins * nangs * 4) return;
    const int k = (int)(t % 4);
    const int64_t a = (t / 4) % nangs, b = t / (4 * nangs);
    const double pb = bx[b], mb = nbx[b], dd = D[a], ee = E[a];
    double sx, sy, tx, ty;
    if (k == 0) { sx = pb; sy = -d; tx = sx + dd; ty = sy + ee; }
    else if (k == 1) { sx = -d; sy = pb; tx = sx + ee; ty = sy + dd; }
    else if (k == 2) { sx = mb; sy = d; tx = sx - dd; ty = sy - ee; }
    else { sx = d; sy = mb; tx = sx - ee; ty = sy - dd; }
    float* s3 = src + 3 * t;
    float* t3 = trg + 3 * t;
    s3[0] = (float)sx; s3[1] = (float)sy; s3[2] = 0.0f;
    t3[0] = (float)tx; t3[1] = (float)ty; t3[2] = 0.0f;
}

}  // namespace vir

using namespace vir;

extern "C" {

int vir_geom_circular(int32_t mode, int64_t nviews, const double* cosv, const double* sinv, int32_t nu,
                      const double* dets_y, int32_t nv, const double* dets_z, double src_iso, double det_iso,
                      float* src, float* trg, void* stream) {
    if (mode < 0 || mode > 2) return fail_arg("mode must be 0 (par), 1 (fan) or 2 (cone)");
    if (nviews < 0 || nu < 1) return fail_arg("bad sizes");
    const int has_z = dets_z != nullptr;
    if (mode == 2 && !has_z) return fail_arg("cone geometry needs DetsZ");
    if (has_z && nv < 1) return fail_arg("nv must be >= 1 with DetsZ");
    GeomArgs A{mode, nviews, nu, has_z ? nv : 1, has_z, src_iso, det_iso};
    const int64_t n = nviews * nu * (int64_t)A.nv;
    if (n == 0) return VIR_OK;
    if (!cosv || !sinv || !dets_y || !src || !trg) return fail_arg("NULL pointer argument");
    hipLaunchKernelGGL(k_geom_circular, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), A,
                       cosv, sinv, dets_y, dets_z, src, trg);
    return check_hip(hipGetLastError(), "vir_geom_circular launch");
}

int vir_geom_square(int64_t nbins, const double* binsx, const double* neg_binsx, int64_t nangs,
                    double dist_y, const double* step_d, const double* step_e, float* src, float* trg,
                    void* stream) {
    if (nbins < 0 || nangs < 0) return fail_arg("bad sizes");
    const int64_t n = nbins * nangs * 4;
    if (n == 0) return VIR_OK;
    if (!binsx || !neg_binsx || !step_d || !step_e || !src || !trg) return fail_arg("NULL pointer argument");
    hipLaunchKernelGGL(k_geom_square, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), nbins,
                       nangs, binsx, neg_binsx, dist_y, step_d, step_e, src, trg);
    return check_hip(hipGetLastError(), "vir_geom_square launch");
}

}  // extern "C"
