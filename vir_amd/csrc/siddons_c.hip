// The single-ray float32 Siddon of sys_mat/siddons.c (siddons :290-341,
// siddons_proj :345-396), batched one ray per thread.  Its results are not
// those of sys_mat/siddon.py: voxel indices are raveled x fastest
// (nx ny z + nx y + x, :253), the entry/exit range ignores the z slab
// (max4/min4 look at 3 of their 4 values, :109-134), plane index bounds are
// floor/ceil without snapping (:183-188), and the grid lines are
// accumulated by repeated float additions (calc_grid_lines :137-148, done by
// the host).  The steps, each as siddons.c evaluates it in float32:
//
//   dST = trg - src; per axis a0 = (plane[0] - src)/dST, aN = (plane[N] -
//   src)/dST (both 0 when dST == 0); m_min/m_max = their order, or (0, 1)
//   when equal (:151-173); alpha_min = max(0, m_min_x, m_min_y), alpha_max =
//   min(1, m_max_x, m_max_y); no entries unless alpha_max > alpha_min.
//   Per axis with dST != 0, planes i from
//     i_min = floor(N - (plane[N] - a_lo dST - src)/d)
//     i_max = ceil((src + a_hi dST - plane[0])/d - 1)
//   ((a_lo, a_hi) = (alpha_min, alpha_max), swapped when dST < 0), alpha_i =
//   (plane[i] - src)/dST in increasing alpha order (:176-203).  The three
//   lists are merged by the rule of merge() (:206-231): take x if x < y and
//   x < z, else y if y < z, else z, while some head is <= 1; then
//   alpha_max.  Consecutive pairs with positive difference give an entry:
//   mid = 0.5 (a + b), index (int)((src + mid dST - plane[0])/d) per axis
//   (truncation), length = |dST| (a - b) with |dST| = sqrt of the float sum
//   of squares (:235-261); siddons_proj sums vol[idx] * |dST| * (b - a) in
//   float32 in entry order (:265-287).
//
// Where siddons.c reads outside its arrays its result is undefined: a plane
// index of -1 (its floor puts a ray entering exactly at plane 0 a hair
// before it, which happens on about half of all rays) or beyond N (its range
// ignores the z slab), and flat voxel indices outside the volume in
// siddons_proj.  Here only the planes in [0, N] are merged and entries
// outside the volume add nothing to a projection and are counted.
#include <cmath>

#include "vir_common.h"

namespace vir {

struct ScGrid {
    int n[3];
    float d[3];
};

struct ScStream {
    const float* plane;
    float src, dst;
    int i, step, end;  // next plane index, +1/-1, one past the last
    float head;        // current alpha, 2 when exhausted
};

__device__ __forceinline__ void sc_next(ScStream& S) {
    if (S.i == S.end) {
        S.head = 2.0f;
    } else {
        S.head = (S.plane[S.i] - S.src) / S.dst;
        S.i += S.step;
    }
}

__device__ __forceinline__ void sc_range(const float* plane, int N, float src, float dst, float& mn, float& mx) {
    float a0, aN;
    if (dst != 0.0f) {
        a0 = (plane[0] - src) / dst;
        aN = (plane[N] - src) / dst;
    } else {
        a0 = 0.0f;
        aN = 0.0f;
    }
    if (a0 < aN) { mn = a0; mx = aN; }
    else if (a0 > aN) { mn = aN; mx = a0; }
    else { mn = 0.0f; mx = 1.0f; }
}

__device__ __forceinline__ void sc_stream(ScStream& S, const float* plane, int N, float d, float src, float dst,
                                          float amin, float amax) {
    S.plane = plane;
    S.src = src;
    S.dst = dst;
    S.i = 0; S.step = 1; S.end = 0;
    if (dst != 0.0f) {
        float lo = dst > 0.0f ? amin : amax, hi = dst > 0.0f ? amax : amin;
        const float fmin = floorf((float)N - ((plane[N] - lo * dst) - src) / d);
        const float fmax = ceilf(((src + hi * dst) - plane[0]) / d - 1.0f);
        // only the planes that exist: [fmin, fmax] intersected with [0, N]
        const int imin = fmin < 0.0f ? 0 : (int)fminf(fmin, (float)(N + 1));
        const int imax = fmax > (float)N ? N : (int)fmaxf(fmax, -1.0f);
        if (imin <= imax) {
            if (dst > 0.0f) { S.i = imin; S.step = 1; S.end = imax + 1; }
            else { S.i = imax; S.step = -1; S.end = imin - 1; }
        }
    }
    sc_next(S);
}

// Walk the entries of one ray: f(idx, ix, iy, iz, length, delta).
template <typename F>
__device__ __forceinline__ void sc_walk(const ScGrid& g, const float* xl, const float* yl, const float* zl,
                                        const float* s, const float* t, F&& f) {
    const float sx = s[0], sy = s[1], sz = s[2];
    const float dx = t[0] - sx, dy = t[1] - sy, dz = t[2] - sz;
    float mnx, mxx, mny, mxy, mnz, mxz;
    sc_range(xl, g.n[0], sx, dx, mnx, mxx);
    sc_range(yl, g.n[1], sy, dy, mny, mxy);
    sc_range(zl, g.n[2], sz, dz, mnz, mxz);
    // max4/min4 compare entries 0..2 only: the z bounds never take part
    float amin = 0.0f, amax = 1.0f;
    if (mnx > amin) amin = mnx;
    if (mny > amin) amin = mny;
    if (mxx < amax) amax = mxx;
    if (mxy < amax) amax = mxy;
    if (amax <= amin) return;
    ScStream X, Y, Z;
    sc_stream(X, xl, g.n[0], g.d[0], sx, dx, amin, amax);
    sc_stream(Y, yl, g.n[1], g.d[1], sy, dy, amin, amax);
    sc_stream(Z, zl, g.n[2], g.d[2], sz, dz, amin, amax);
    const float dist = __builtin_sqrtf((dx * dx + dy * dy) + dz * dz);
    float prev = amin;
    auto emit = [&](float next) {
        const float da = next - prev;
        if (da > 0.0f) {
            const float mid = 0.5f * (next + prev);
            const int ix = (int)(((sx + mid * dx) - xl[0]) / g.d[0]);
            const int iy = (int)(((sy + mid * dy) - yl[0]) / g.d[1]);
            const int iz = (int)(((sz + mid * dz) - zl[0]) / g.d[2]);
            f(ix, iy, iz, dist, da);
        }
        prev = next;
    };
    while (X.head <= 1.0f || Y.head <= 1.0f || Z.head <= 1.0f) {
        if (X.head < Y.head && X.head < Z.head) { emit(X.head); sc_next(X); }
        else if (Y.head < Z.head) { emit(Y.head); sc_next(Y); }
        else { emit(Z.head); sc_next(Z); }
    }
    emit(amax);
}

__global__ __launch_bounds__(256) void k_sc_count(ScGrid g, const float* __restrict__ xl, const float* __restrict__ yl,
                                                 const float* __restrict__ zl, const float* __restrict__ src,
                                                 const float* __restrict__ trg, int64_t nrays,
                                                 int32_t* __restrict__ counts) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrays) return;
    int c = 0;
    sc_walk(g, xl, yl, zl, src + 3 * r, trg + 3 * r, [&](int, int, int, float, float) { ++c; });
    counts[r] = c;
}

__global__ __launch_bounds__(256) void k_sc_fill(ScGrid g, const float* __restrict__ xl, const float* __restrict__ yl,
                                                const float* __restrict__ zl, const float* __restrict__ src,
                                                const float* __restrict__ trg, int64_t nrays,
                                                const int64_t* __restrict__ off, int32_t* __restrict__ idx,
                                                float* __restrict__ len) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrays) return;
    int64_t e = off[r];
    const int64_t e1 = off[r + 1];
    const int nxy = g.n[0] * g.n[1];
    sc_walk(g, xl, yl, zl, src + 3 * r, trg + 3 * r, [&](int ix, int iy, int iz, float dist, float da) {
        if (e >= e1) return;
        idx[e] = nxy * iz + g.n[0] * iy + ix;
        len[e] = dist * da;
        ++e;
    });
}

__global__ __launch_bounds__(256) void k_sc_proj(ScGrid g, const float* __restrict__ xl, const float* __restrict__ yl,
                                                const float* __restrict__ zl, const float* __restrict__ src,
                                                const float* __restrict__ trg, int64_t nrays,
                                                const float* __restrict__ vol, float* __restrict__ sum,
                                                int32_t* __restrict__ err) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrays) return;
    float acc = 0.0f;
    int bad = 0;
    const int64_t nxy = (int64_t)g.n[0] * g.n[1], nvox = nxy * g.n[2];
    sc_walk(g, xl, yl, zl, src + 3 * r, trg + 3 * r, [&](int ix, int iy, int iz, float dist, float da) {
        // siddons.c reads iArray at the flat index even when one axis index
        // is off the grid (x = -1 or nx lands in a neighbouring row)
        const int64_t fi = nxy * iz + (int64_t)g.n[0] * iy + ix;
        if (fi >= 0 && fi < nvox) acc += (vol[fi] * dist) * da;
        else ++bad;
    });
    sum[r] = acc;
    if (bad && err) atomicAdd(err, bad);
}

static int sc_setup(const vir_sc_grid* gp, const float* xl, const float* yl, const float* zl, ScGrid& g) {
    if (!gp) return fail_arg("grid is NULL");
    g.n[0] = gp->nx; g.n[1] = gp->ny; g.n[2] = gp->nz;
    g.d[0] = gp->dx; g.d[1] = gp->dy; g.d[2] = gp->dz;
    for (int a = 0; a < 3; ++a) {
        if (g.n[a] < 1 || g.n[a] > (1 << 22)) return fail_arg("grid sizes out of range");
        if (!(g.d[a] != 0.0f) || !std::isfinite(g.d[a])) return fail_arg("voxel sizes must be finite, non-zero");
    }
    if ((int64_t)g.n[0] * g.n[1] * g.n[2] >= ((int64_t)1 << 31)) return fail_arg("grid must have < 2^31 voxels");
    if (!xl || !yl || !zl) return fail_arg("grid lines are NULL");
    return VIR_OK;
}

static inline unsigned sc_blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace vir

using namespace vir;

extern "C" {

int vir_siddons_c_count(const float* src, const float* trg, int64_t nrays, const vir_sc_grid* grid,
                        const float* x_lines, const float* y_lines, const float* z_lines, int32_t* counts,
                        void* stream) {
    ScGrid g;
    int rc = sc_setup(grid, x_lines, y_lines, z_lines, g);
    if (rc) return rc;
    if (nrays < 0) return fail_arg("nrays < 0");
    if (nrays == 0) return VIR_OK;
    if (!src || !trg || !counts) return fail_arg("NULL pointer argument");
    hipLaunchKernelGGL(k_sc_count, dim3(sc_blocks(nrays)), dim3(256), 0, as_stream(stream), g, x_lines, y_lines,
                       z_lines, src, trg, nrays, counts);
    return check_hip(hipGetLastError(), "vir_siddons_c_count launch");
}

int vir_siddons_c_fill(const float* src, const float* trg, int64_t nrays, const vir_sc_grid* grid,
                       const float* x_lines, const float* y_lines, const float* z_lines, const int64_t* offsets,
                       int32_t* idx, float* len, void* stream) {
    ScGrid g;
    int rc = sc_setup(grid, x_lines, y_lines, z_lines, g);
    if (rc) return rc;
    if (nrays < 0) return fail_arg("nrays < 0");
    if (nrays == 0) return VIR_OK;
    if (!src || !trg || !offsets || !idx || !len) return fail_arg("NULL pointer argument");
    hipLaunchKernelGGL(k_sc_fill, dim3(sc_blocks(nrays)), dim3(256), 0, as_stream(stream), g, x_lines, y_lines,
                       z_lines, src, trg, nrays, offsets, idx, len);
    return check_hip(hipGetLastError(), "vir_siddons_c_fill launch");
}

int vir_siddons_c_proj(const float* src, const float* trg, int64_t nrays, const vir_sc_grid* grid,
                       const float* x_lines, const float* y_lines, const float* z_lines, const float* vol,
                       float* ray_sum, int32_t* err_count, void* stream) {
    ScGrid g;
    int rc = sc_setup(grid, x_lines, y_lines, z_lines, g);
    if (rc) return rc;
    if (nrays < 0) return fail_arg("nrays < 0");
    if (nrays == 0) return VIR_OK;
    if (!src || !trg || !vol || !ray_sum) return fail_arg("NULL pointer argument");
    hipLaunchKernelGGL(k_sc_proj, dim3(sc_blocks(nrays)), dim3(256), 0, as_stream(stream), g, x_lines, y_lines,
                       z_lines, src, trg, nrays, vol, ray_sum, err_count);
    return check_hip(hipGetLastError(), "vir_siddons_c_proj launch");
}

}  // extern "C"
