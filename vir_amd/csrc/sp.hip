// Separable-footprint prototypes of sys_mat/sp.py, with exact adjoints.
//
// 1. Triangle (linear B-spline) footprint fan beam, sp.py:12-87.  Pixel
//    centre (x, y) = d (i - n/2 + 1/2), bin centre u_b = du (b - nu/2 + 1/2);
//    per view (c, s) = (cos a, sin a):
//        xp = c x + s y,  yp = -s x + c y,  den = DSO - yp,
//        u0 = DSD xp / den,  h = DSD d / den,
//        w  = max(0, 1 - |u_b - u0| / h)          (sampled at bin centres)
//    sino[a, b] = out_scale sum_pixels img w; the reference uses
//    out_scale = d.  The backprojection is the exact transpose (same weight
//    function, same pairs).
//
//    Forward, one thread per (view, bin): w > 0  <=>  |u_b den - DSD xp| <
//    DSD d, and u_b den - DSD xp = A + B t is affine in the position t of the
//    pixel along a line of the grid (B = u_b s - DSD c along x at fixed y,
//    B = -u_b c - DSD s along y at fixed x).  The thread walks the lines of
//    the axis with the larger |B| (>= DSD/sqrt(2): B_x^2 + B_y^2 = u_b^2 +
//    DSD^2), whose pixels meeting the bin are those with t inside
//    (+-DSD d - A)/B, at most ~2 sqrt(2) + 1 per line; one pixel of margin
//    each side, the exact test is w > 0.
//    Backward, one thread per pixel: per view the bins within [u0 - h,
//    u0 + h] (sp.py:70-74) plus one of margin.
//
// 2. The separable toy model of sp.py:93-211 (CTGeometry, footprint1D,
//    forward_project, back_project): with f(r) = max(0, 1 - |r|),
//        proj[v, u] = sum_{x,y,z} vol[x,y,z] f(u/Dd - x/Ds) f(v/Dd - y/Ds) f(z/Ds)
//    evaluated as three contractions (z, then y, then x) in double, and its
//    transpose in the reverse order.  Voxel centres (i - n/2 + 1/2) d_axis,
//    detector centres (j - n/2 + 1/2) pitch (sp.py:105-111).
#include <cmath>

#include "vir_common.h"

namespace vir {

struct SpGeo {
    int nx, ny, nu, na;
    float d, du, DSO, DSD;
    float out_scale;
};

__device__ __forceinline__ float sp_pix(int n, float d, int i) { return d * (((float)i - 0.5f * (float)n) + 0.5f); }

__device__ __forceinline__ float sp_weight(const SpGeo& G, float c, float s, int ix, int iy, float ub) {
    const float x = sp_pix(G.nx, G.d, ix), y = sp_pix(G.ny, G.d, iy);
    const float xp = c * x + s * y, yp = -s * x + c * y;
    const float den = G.DSO - yp;
    const float u0 = G.DSD * xp / den;
    const float h = G.DSD * G.d / den;
    return fmaxf(0.0f, 1.0f - fabsf(ub - u0) / h);
}

__global__ __launch_bounds__(256) void k_sp_fan_fp(SpGeo G, const float* __restrict__ cosv,
                                                  const float* __restrict__ sinv, const float* __restrict__ img,
                                                  float* __restrict__ sino, int accumulate) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.na * G.nu) return;
    const int a = (int)(t / G.nu), b = (int)(t % G.nu);
    const float c = cosv[a], s = sinv[a];
    const float ub = sp_pix(G.nu, G.du, b);
    const float Bx = ub * s - G.DSD * c, By = -ub * c - G.DSD * s;
    const float Dd = G.DSD * G.d;
    double acc = 0.0;
    if (fabsf(Bx) >= fabsf(By)) {
        for (int iy = 0; iy < G.ny; ++iy) {
            const float y = sp_pix(G.ny, G.d, iy);
            const float A = (ub * G.DSO - ub * c * y) - G.DSD * s * y;
            float x0 = (-Dd - A) / Bx, x1 = (Dd - A) / Bx;
            if (x0 > x1) { const float tmp = x0; x0 = x1; x1 = tmp; }
            int k0 = (int)floorf(x0 / G.d + 0.5f * (float)G.nx - 0.5f) - 1;
            int k1 = (int)ceilf(x1 / G.d + 0.5f * (float)G.nx - 0.5f) + 1;
            k0 = max(k0, 0);
            k1 = min(k1, G.nx - 1);
            for (int ix = k0; ix <= k1; ++ix) {
                const float w = sp_weight(G, c, s, ix, iy, ub);
                if (w > 0.0f) acc += (double)(img[(int64_t)ix * G.ny + iy] * w);
            }
        }
    } else {
        for (int ix = 0; ix < G.nx; ++ix) {
            const float x = sp_pix(G.nx, G.d, ix);
            const float A = (ub * G.DSO + ub * s * x) - G.DSD * c * x;
            float y0 = (-Dd - A) / By, y1 = (Dd - A) / By;
            if (y0 > y1) { const float tmp = y0; y0 = y1; y1 = tmp; }
            int k0 = (int)floorf(y0 / G.d + 0.5f * (float)G.ny - 0.5f) - 1;
            int k1 = (int)ceilf(y1 / G.d + 0.5f * (float)G.ny - 0.5f) + 1;
            k0 = max(k0, 0);
            k1 = min(k1, G.ny - 1);
            for (int iy = k0; iy <= k1; ++iy) {
                const float w = sp_weight(G, c, s, ix, iy, ub);
                if (w > 0.0f) acc += (double)(img[(int64_t)ix * G.ny + iy] * w);
            }
        }
    }
    float r = G.out_scale * (float)acc;
    if (accumulate) r += sino[t];
    sino[t] = r;
}

__global__ __launch_bounds__(256) void k_sp_fan_bp(SpGeo G, const float* __restrict__ cosv,
                                                  const float* __restrict__ sinv, const float* __restrict__ sino,
                                                  float* __restrict__ img, int accumulate) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nx * G.ny) return;
    const int ix = (int)(t / G.ny), iy = (int)(t % G.ny);
    const float x = sp_pix(G.nx, G.d, ix), y = sp_pix(G.ny, G.d, iy);
    double acc = 0.0;
    for (int a = 0; a < G.na; ++a) {
        const float c = cosv[a], s = sinv[a];
        const float xp = c * x + s * y, yp = -s * x + c * y;
        const float den = G.DSO - yp;
        const float u0 = G.DSD * xp / den, h = G.DSD * G.d / den;
        int b0 = (int)floorf((u0 - h) / G.du + 0.5f * (float)G.nu) - 1;
        int b1 = (int)ceilf((u0 + h) / G.du + 0.5f * (float)G.nu) + 1;
        b0 = max(b0, 0);
        b1 = min(b1, G.nu - 1);
        const float* __restrict__ row = sino + (int64_t)a * G.nu;
        for (int b = b0; b <= b1; ++b) {
            const float w = sp_weight(G, c, s, ix, iy, sp_pix(G.nu, G.du, b));
            if (w > 0.0f) acc += (double)(row[b] * w);
        }
    }
    float r = G.out_scale * (float)acc;
    if (accumulate) r += img[t];
    img[t] = r;
}

// ---- the separable toy model ----

struct SepGeo {
    int nx, ny, nz, nu, nv;
    double dx, dy, dz, su, sv, Ds, Dd;
};

__device__ __forceinline__ double sep_c(int n, double d, int i) { return ((double)i - 0.5 * (double)n + 0.5) * d; }
__device__ __forceinline__ double fp1(double r) { return fmax(1.0 - fabs(r), 0.0); }
__device__ __forceinline__ double sep_wx(const SepGeo& G, int iu, int ix) {
    return fp1(sep_c(G.nu, G.su, iu) / G.Dd - sep_c(G.nx, G.dx, ix) / G.Ds);
}
__device__ __forceinline__ double sep_wy(const SepGeo& G, int iv, int iy) {
    return fp1(sep_c(G.nv, G.sv, iv) / G.Dd - sep_c(G.ny, G.dy, iy) / G.Ds);
}
__device__ __forceinline__ double sep_wz(const SepGeo& G, int iz) { return fp1(sep_c(G.nz, G.dz, iz) / G.Ds); }

// M[x, y] = sum_z wz vol[x, y, z]
__global__ __launch_bounds__(256) void k_sep_z(SepGeo G, const float* __restrict__ vol, double* __restrict__ M) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nx * G.ny) return;
    double s = 0.0;
    for (int iz = 0; iz < G.nz; ++iz) s += sep_wz(G, iz) * (double)vol[t * G.nz + iz];
    M[t] = s;
}

// T[v, x] = sum_y wy(v, y) M[x, y]
__global__ __launch_bounds__(256) void k_sep_y(SepGeo G, const double* __restrict__ M, double* __restrict__ T) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nv * G.nx) return;
    const int iv = (int)(t / G.nx), ix = (int)(t % G.nx);
    double s = 0.0;
    for (int iy = 0; iy < G.ny; ++iy) {
        const double w = sep_wy(G, iv, iy);
        if (w != 0.0) s += w * M[(int64_t)ix * G.ny + iy];
    }
    T[t] = s;
}

// proj[v, u] = sum_x wx(u, x) T[v, x]
__global__ __launch_bounds__(256) void k_sep_x(SepGeo G, const double* __restrict__ T, float* __restrict__ proj) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nv * G.nu) return;
    const int iv = (int)(t / G.nu), iu = (int)(t % G.nu);
    double s = 0.0;
    for (int ix = 0; ix < G.nx; ++ix) {
        const double w = sep_wx(G, iu, ix);
        if (w != 0.0) s += w * T[(int64_t)iv * G.nx + ix];
    }
    proj[t] = (float)s;
}

// S[v, x] = sum_u wx(u, x) proj[v, u]
__global__ __launch_bounds__(256) void k_sep_xt(SepGeo G, const float* __restrict__ proj, double* __restrict__ S) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nv * G.nx) return;
    const int iv = (int)(t / G.nx), ix = (int)(t % G.nx);
    double s = 0.0;
    for (int iu = 0; iu < G.nu; ++iu) {
        const double w = sep_wx(G, iu, ix);
        if (w != 0.0) s += w * (double)proj[(int64_t)iv * G.nu + iu];
    }
    S[t] = s;
}

// vol[x, y, z] = wz(z) sum_v wy(v, y) S[v, x]
__global__ __launch_bounds__(256) void k_sep_yzt(SepGeo G, const double* __restrict__ S, float* __restrict__ vol) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nx * G.ny) return;
    const int ix = (int)(t / G.ny), iy = (int)(t % G.ny);
    double s = 0.0;
    for (int iv = 0; iv < G.nv; ++iv) {
        const double w = sep_wy(G, iv, iy);
        if (w != 0.0) s += w * S[(int64_t)iv * G.nx + ix];
    }
    for (int iz = 0; iz < G.nz; ++iz) vol[t * G.nz + iz] = (float)(sep_wz(G, iz) * s);
}

static int sp_setup(const vir_sp_geom* g, SpGeo& G) {
    if (!g) return fail_arg("geometry is NULL");
    G.nx = g->nx; G.ny = g->ny; G.nu = g->nu; G.na = g->na;
    G.d = g->d_pix; G.du = g->du; G.DSO = g->DSO; G.DSD = g->DSD;
    if (G.nx < 1 || G.ny < 1 || G.nu < 1 || G.na < 0) return fail_arg("bad sizes");
    if (!(G.d > 0.0f) || !(G.du > 0.0f) || !(G.DSO > 0.0f) || !(G.DSD > 0.0f))
        return fail_arg("d_pix, du, DSO and DSD must be > 0");
    // every pixel in front of the source: its corner distance below DSO
    const double hx = 0.5 * G.nx * (double)G.d, hy = 0.5 * G.ny * (double)G.d;
    if (!(std::sqrt(hx * hx + hy * hy) < 0.999 * (double)G.DSO)) return fail_arg("grid reaches the source");
    return VIR_OK;
}

static int sep_setup(const vir_sep_geom* g, SepGeo& G) {
    if (!g) return fail_arg("geometry is NULL");
    G.nx = g->nx; G.ny = g->ny; G.nz = g->nz; G.nu = g->nu; G.nv = g->nv;
    G.dx = g->dx; G.dy = g->dy; G.dz = g->dz; G.su = g->su; G.sv = g->sv; G.Ds = g->Ds; G.Dd = g->Dd;
    if (G.nx < 1 || G.ny < 1 || G.nz < 1 || G.nu < 1 || G.nv < 1) return fail_arg("bad sizes");
    if (!(G.Ds != 0.0) || !(G.Dd != 0.0) || !std::isfinite(G.Ds) || !std::isfinite(G.Dd))
        return fail_arg("Ds and Dd must be finite and non-zero");
    return VIR_OK;
}

static inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace vir

using namespace vir;

extern "C" {

int vir_sp_fan_fp(const vir_sp_geom* geom, const float* cosv, const float* sinv, const float* img, float* sino,
                  float out_scale, uint32_t flags, void* stream) {
    SpGeo G;
    int rc = sp_setup(geom, G);
    if (rc) return rc;
    if (G.na == 0) return VIR_OK;
    if (!cosv || !sinv || !img || !sino) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    hipLaunchKernelGGL(k_sp_fan_fp, dim3(nb256((int64_t)G.na * G.nu)), dim3(256), 0, as_stream(stream), G, cosv,
                       sinv, img, sino, (flags & VIR_ACCUMULATE) ? 1 : 0);
    return check_hip(hipGetLastError(), "vir_sp_fan_fp launch");
}

int vir_sp_fan_bp(const vir_sp_geom* geom, const float* cosv, const float* sinv, const float* sino, float* img,
                  float out_scale, uint32_t flags, void* stream) {
    SpGeo G;
    int rc = sp_setup(geom, G);
    if (rc) return rc;
    if (!img || (G.na > 0 && (!cosv || !sinv || !sino))) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    hipLaunchKernelGGL(k_sp_fan_bp, dim3(nb256((int64_t)G.nx * G.ny)), dim3(256), 0, as_stream(stream), G, cosv,
                       sinv, sino, img, (flags & VIR_ACCUMULATE) ? 1 : 0);
    return check_hip(hipGetLastError(), "vir_sp_fan_bp launch");
}

int64_t vir_sep_workspace(const vir_sep_geom* geom) {
    SepGeo G;
    if (sep_setup(geom, G)) return -1;
    const int64_t a = (int64_t)G.nx * G.ny, b = (int64_t)G.nv * G.nx;
    return (int64_t)sizeof(double) * (a > b ? a : b) + (int64_t)sizeof(double) * b;
}

int vir_sep_fp(const vir_sep_geom* geom, const float* vol, float* proj, void* work, void* stream) {
    SepGeo G;
    int rc = sep_setup(geom, G);
    if (rc) return rc;
    if (!vol || !proj || !work) return fail_arg("NULL pointer argument");
    const int64_t a = (int64_t)G.nx * G.ny, b = (int64_t)G.nv * G.nx;
    double* M = reinterpret_cast<double*>(work);
    double* T = M + (a > b ? a : b);
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_sep_z, dim3(nb256(a)), dim3(256), 0, st, G, vol, M);
    hipLaunchKernelGGL(k_sep_y, dim3(nb256(b)), dim3(256), 0, st, G, M, T);
    hipLaunchKernelGGL(k_sep_x, dim3(nb256((int64_t)G.nv * G.nu)), dim3(256), 0, st, G, T, proj);
    return check_hip(hipGetLastError(), "vir_sep_fp launch");
}

int vir_sep_bp(const vir_sep_geom* geom, const float* proj, float* vol, void* work, void* stream) {
    SepGeo G;
    int rc = sep_setup(geom, G);
    if (rc) return rc;
    if (!vol || !proj || !work) return fail_arg("NULL pointer argument");
    const int64_t a = (int64_t)G.nx * G.ny, b = (int64_t)G.nv * G.nx;
    double* S = reinterpret_cast<double*>(work);
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_sep_xt, dim3(nb256(b)), dim3(256), 0, st, G, proj, S);
    hipLaunchKernelGGL(k_sep_yzt, dim3(nb256(a)), dim3(256), 0, st, G, S, vol);
    return check_hip(hipGetLastError(), "vir_sep_bp launch");
}

}  // extern "C"
