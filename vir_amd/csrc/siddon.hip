// Siddon projector kernels for gfx950: trace-to-CSR, on-the-fly forward
// projection and the brick-gather exact-adjoint backprojection.
//
// All kernels share the ray arithmetic of siddon_ray.h, so the (voxel,
// length) entries seen by the forward projector, the backprojector and the
// stored-matrix builder are the same bits.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "vir_common.h"

namespace vir {

constexpr int kRayThreads = 256;

// ---------------------------------------------------------------- trace

template <bool P2, bool ISO>
__global__ __launch_bounds__(kRayThreads) void k_siddon_count(GridK g, const float* __restrict__ src,
                                                             const float* __restrict__ trg,
                                                             int64_t nrays, int32_t* __restrict__ counts,
                                                             uint8_t* __restrict__ flags) {
    int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
    int64_t r = blk * kRayThreads + threadIdx.x;
    if (r >= nrays) return;
    Ray R;
    ray_setup(g, src + 3 * r, trg + 3 * r, R);
    int cnt = 0;
    uint8_t fl = 0;
    if (R.hit) {
        fl = 2;
        Walker W;
        walker_begin(g, R, W);
        walk_dispatch(g, R, [&](auto fast) {
            walk_entries<P2, ISO, false, decltype(fast)::value>(
                g, R, W, VIR_INF, [&](int ix, int iy, int iz, float) {
                    ++cnt;
                    if ((unsigned)ix >= (unsigned)g.n[0] || (unsigned)iy >= (unsigned)g.n[1] ||
                        (unsigned)iz >= (unsigned)g.n[2])
                        fl |= 1;
                });
        });
    }
    counts[r] = cnt;
    flags[r] = fl;
}

template <bool P2, bool ISO>
__global__ __launch_bounds__(kRayThreads) void k_siddon_fill(GridK g, const float* __restrict__ src,
                                                            const float* __restrict__ trg, int64_t nrays,
                                                            const int64_t* __restrict__ off,
                                                            int32_t* __restrict__ rav, int32_t* __restrict__ ox,
                                                            int32_t* __restrict__ oy, int32_t* __restrict__ oz,
                                                            float* __restrict__ olen) {
    int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
    int64_t r = blk * kRayThreads + threadIdx.x;
    if (r >= nrays) return;
    Ray R;
    ray_setup(g, src + 3 * r, trg + 3 * r, R);
    if (!R.hit) return;
    int64_t e = off[r];
    const int64_t e_end = off[r + 1];
    Walker W;
    walker_begin(g, R, W);
    walk_dispatch(g, R, [&](auto fast) {
        walk_entries<P2, ISO, true, decltype(fast)::value>(
            g, R, W, VIR_INF, [&](int ix, int iy, int iz, float len) {
                if (e >= e_end) return;  // count and fill walk identically; never past the row
                if (rav) {
                    bool in = (unsigned)ix < (unsigned)g.n[0] && (unsigned)iy < (unsigned)g.n[1] &&
                              (unsigned)iz < (unsigned)g.n[2];
                    rav[e] = in ? (ix * g.n[1] + iy) * g.n[2] + iz : -1;
                }
                if (ox) ox[e] = ix;
                if (oy) oy[e] = iy;
                if (oz) oz[e] = iz;
                if (olen) olen[e] = len;
                ++e;
            });
    });
}

// Rays whose entry/exit divisions need div_rn's range check
// (entry_unguarded false), rays that meet the grid and need it for a plane
// division or take the float64 list (planes_unguarded false), and, on g.unit
// grids (whose unguarded kernels walk in voxel units), rays for which that
// walk is not exact (unit_exact false); one atomic per wave.
__global__ __launch_bounds__(kRayThreads) void k_siddon_count_guarded(GridK g, const float* __restrict__ src,
                                                                     const float* __restrict__ trg,
                                                                     int64_t nrays, int32_t* __restrict__ count) {
    int64_t r = (int64_t)blockIdx.x * kRayThreads + threadIdx.x;
    bool need = false;
    if (r < nrays) {
        Ray R;
        ray_setup_core(g, src + 3 * r, trg + 3 * r, R);
        need = !entry_unguarded(g, R) || (R.hit && !planes_unguarded(g, R)) ||
               (R.hit && g.unit && !unit_exact(g, R));
    }
    unsigned long long m = __ballot(need);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (int)__popcll(m));
}

// Per-ray records for the backprojector: ray_setup (the forward projector's
// per-ray setup and exact slices) stored as ray_store_record lays it out.
__global__ __launch_bounds__(kRayThreads) void k_siddon_ray_records(GridK g, const float* __restrict__ src,
                                                                   const float* __restrict__ trg,
                                                                   int64_t nrays, float* __restrict__ rec) {
    int64_t r = (int64_t)blockIdx.x * kRayThreads + threadIdx.x;
    if (r >= nrays) return;
    Ray R;
    ray_setup_core(g, src + 3 * r, trg + 3 * r, R);
    if (R.hit) ray_setup_slices(g, R);
    else { R.k0 = i3v{0, 0, 0}; R.k1 = i3v{0, 0, 0}; }
    ray_store_record(R, rec + kRayRec * r);
}

// ------------------------------------------------------------ forward

// NB volumes that share the rays (volume j at vol + j*vstride, its
// projection at sino + j*nrays): one trace per ray serves all of them, and
// each volume's sums are formed exactly as for NB = 1.  Unguarded rays on
// power-of-two grids are walked in voxel units (gu = unit_grid(g); the
// host establishes unit_exact for every ray with the unguarded precondition).
template <bool P2, bool ISO, bool UNG, int NB = 1>
__global__ __launch_bounds__(kRayThreads) __attribute__((amdgpu_waves_per_eu(NB == 1 ? 7 : 6))) void k_siddon_fp(
    GridK g, const float* __restrict__ src, const float* __restrict__ trg, const float* __restrict__ rec,
    int64_t nrays, const float* __restrict__ vol, float* __restrict__ sino, int accumulate,
    int32_t* __restrict__ err, SirtEpi E, int64_t vstride, GridK gu) {
    constexpr bool UNIT = P2 && UNG;
    int64_t blk = xcd_remap(blockIdx.x, gridDim.x);
    int64_t r = blk * kRayThreads + threadIdx.x;
    if (r >= nrays) return;
    Ray R;
    // batched calls always come with records
    if (NB > 1 || rec) ray_load_record(rec + (size_t)kRayRec * r, R);  // the same setup, precomputed
    else ray_setup<UNG>(g, src + 3 * r, trg + 3 * r, R);
    // float64 accumulation of the float32 products (the accumulator of
    // projection/proj_numba.py:17), one-deep software pipeline: the gather
    // issued for entry j lands in pv while entry j+1 is being traced, and is
    // consumed before the gather for j+1 overwrites it.  Every pair issues a
    // gather (of voxel 0 for dropped and out-of-grid pairs, weighted 0), so
    // the body has no branch around the load.
    double sum[NB];
    float pv[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        sum[j] = 0.0;
        pv[j] = 0.0f;
    }
    int bad = 0;
    float pl = 0.0f;
    bool pin = false;
    if (R.hit) {
        if (UNIT) unit_ray(g, R);
        const GridK& gw = UNIT ? gu : g;
        Walker W;
        walker_begin(gw, R, W);
        const unsigned nyz = (unsigned)(g.n[1] * g.n[2]);
        auto visit = [&](int ix, int iy, int iz, float len, bool keep) {
                    bool in = (unsigned)ix < (unsigned)g.n[0] && (unsigned)iy < (unsigned)g.n[1] &&
                              (unsigned)iz < (unsigned)g.n[2];
                    bad += (keep && !in) ? 1 : 0;
                    bool use = keep && in;
                    // iy, n2 <= 2^22 and iy n2 < n1 n2: a 24-bit (full-rate)
                    // multiply is exact for in-grid entries
                    unsigned vi = (unsigned)ix * nyz + __umul24((unsigned)iy, (unsigned)g.n[2]) + (unsigned)iz;
                    const unsigned gi = use ? vi : 0u;
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        sum[j] += (double)((pin ? pv[j] : 0.0f) * pl);
                        pv[j] = vol[j * vstride + gi];
                    }
                    pin = use;
                    pl = use ? len : 0.0f;
        };
        walk_dispatch<UNG>(gw, R, [&](auto fast) {
            constexpr bool F = decltype(fast)::value;
            if constexpr (F && UNG) {
                // the general walk to a few planes past the entry, the walk
                // without the special values through the ray's middle (same
                // entries; window_mid holds there), the general walk to the end
                float gmax = 0.0f;
#pragma unroll
                for (int a = 0; a < 3; ++a) gmax = fmaxf(gmax, fabsf(R.dST[a]) * gw.inv_d[a]);
                const float a1 = R.amin + 3.0f / gmax, a2 = R.amax - 4.0f / gmax;
                walk_entries<P2, ISO, false, true, true, UNIT>(gw, R, W, a1, visit);
                if (W.consumed == 2 && window_mid(gw, R, W.prev, a2))
                    walk_mid<P2, ISO, true, true, UNIT>(gw, R, W, a2, visit);
                walk_entries<P2, ISO, false, true, true, UNIT>(gw, R, W, VIR_INF, visit);
            } else {
                walk_entries<P2, ISO, false, F, true>(gw, R, W, VIR_INF, visit);
            }
        });
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        sum[j] += (double)((pin ? pv[j] : 0.0f) * pl);
        float res = (float)sum[j];
        float* sj = sino + j * nrays;
        if (accumulate) res += sj[r];
        sirt_store(sirt_item(E, j * nrays), sj, r, res);
    }
    if (bad && err) atomicAdd(err, bad);
}

// ------------------------------------------------------------ backward

// Brick geometry shared by all threads of a workgroup for one view.
struct ViewFoot {
    int iu0, iu1, iv0, iv1;  // inclusive ray-index ranges of the guard box footprint
    int su, sv;              // colour strides (rays of one colour never share a voxel)
};

constexpr float kGuard = 0.05f;  // guard band around a brick, in voxels

// Project world point p with view matrix P (row-major 3x4).
__device__ __forceinline__ void project(const float* P, float x, float y, float z, float& u, float& v,
                                        float& w) {
    u = P[0] * x + P[1] * y + P[2] * z + P[3];
    v = P[4] * x + P[5] * y + P[6] * z + P[7];
    w = P[8] * x + P[9] * y + P[10] * z + P[11];
}

// Footprint of the guard box [gb_lo, gb_hi] on the detector in a view P,
// computed by the whole wave (every lane must call it).
//
// Candidate rays: a ray with an entry in the sub-brick meets the guard
// box, so its index lies in the box's projected range (the convex hull of
// the projected corners, all in front of the source) widened by the
// matrix fit error (hosts accept matrices within 0.05 index units).
//
// Colour strides: for two points Y, Z of one voxel, with rows a (u),
// c (w) of the matrix and u = a.Y~ / c.Y~,
//   u(Z) - u(Y) = sum_i (Z_i - Y_i) (a_i - u(Y) c_i) / c.Z~,
// so every voxel of the guard box projects to a u-extent of at most
//   fu = sum_i d_i max(|a_i - umin c_i|, |a_i - umax c_i|) / wmin
// (|a_i - u c_i| is convex in u, so its maximum over the box's range
// [umin, umax] is at an end; c.X~ is affine, so its minimum wmin is at
// a corner).  Two rays through one voxel are at most fu + 2 * 0.05 apart
// in u, so rays whose u-indices differ by su > 1.05 fu + 0.15 never
// share a voxel; likewise in v.  The 5 % cover float rounding and
// entries whose computed index is one voxel off their true midpoint by
// ~1e-3 voxel.
//
// The NS views va0 .. va0 + NS - 1 of a pass are done together: corner lanes
// 8 s .. 8 s + 7 project the guard box for view va0 + s and reduce within
// their group, and each lane takes the results of its own slot's group and
// view (views past the last one, nva, repeat it; their lanes are idle).  Per
// view the operations and their order are those of one view alone.
template <int NS>
__device__ __forceinline__ ViewFoot view_foot(const GridK& g, const float* __restrict__ proj, int va0, int nva,
                                              const float* gb_lo, const float* gb_hi, int nu, int nv, int lane,
                                              int slot) {
    static_assert(8 * NS <= 64, "eight corner lanes per view");
    // element i of slot s's matrix: wave-uniform loads, picked per lane
    auto pick = [&](int s, int i) -> float {
        float x = proj[12 * va0 + i];
#pragma unroll
        for (int t = 1; t < NS; ++t) x = s == t ? proj[12 * (va0 + min(t, nva - 1)) + i] : x;
        return x;
    };
    const int cs = lane >> 3;
    const bool corner = cs < NS;
    float Pc[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) Pc[i] = pick(cs, i);
    float gu = 0.f, gv = 0.f, gw = 0.f;
    project(Pc, (lane & 4) ? gb_hi[0] : gb_lo[0], (lane & 2) ? gb_hi[1] : gb_lo[1],
            (lane & 1) ? gb_hi[2] : gb_lo[2], gu, gv, gw);
    int gok = !corner || (gw > 0.0f);
    float guu = corner ? gu / gw : VIR_INF, gvv = corner ? gv / gw : VIR_INF;
    float gumin = guu, gvmin = gvv;
    float gumax = corner ? guu : -VIR_INF, gvmax = corner ? gvv : -VIR_INF;
    float gwmin = corner ? gw : VIR_INF;
#pragma unroll
    for (int m = 1; m < 8; m <<= 1) {
        gumin = fminf(gumin, __shfl_xor(gumin, m));
        gvmin = fminf(gvmin, __shfl_xor(gvmin, m));
        gumax = fmaxf(gumax, __shfl_xor(gumax, m));
        gvmax = fmaxf(gvmax, __shfl_xor(gvmax, m));
        gwmin = fminf(gwmin, __shfl_xor(gwmin, m));
        gok &= __shfl_xor(gok, m);
    }
    // the first lane of each group holds its view's results
    const int src = 8 * slot;
    gumin = __shfl(gumin, src); gumax = __shfl(gumax, src);
    gvmin = __shfl(gvmin, src); gvmax = __shfl(gvmax, src);
    gwmin = __shfl(gwmin, src);
    gok = __shfl(gok, src);
    float P[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) P[i] = pick(slot, i);
    float ku = 0.0f, kv = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float c = P[8 + i];
        ku = ku + g.d[i] * fmaxf(fabsf(P[i] - gumin * c), fabsf(P[i] - gumax * c));
        kv = kv + g.d[i] * fmaxf(fabsf(P[4 + i] - gvmin * c), fabsf(P[4 + i] - gvmax * c));
    }
    const float fu = ku / gwmin, fv = kv / gwmin;

    ViewFoot f;
    const float range_margin = 0.075f, stride_margin = 0.15f;
    if (!gok || !(fu < 1.0e6f) || !(fv < 1.0e6f)) {
        // degenerate projection (hosts validate geometry, so not taken)
        f.iu0 = 0; f.iu1 = nu - 1; f.iv0 = 0; f.iv1 = nv - 1;
        f.su = nu; f.sv = nv;
    } else {
        float a0 = ceilf(gumin - range_margin), a1 = floorf(gumax + range_margin);
        float c0 = ceilf(gvmin - range_margin), c1 = floorf(gvmax + range_margin);
        f.iu0 = (int)fmaxf(a0, 0.0f);
        f.iu1 = (int)fminf(a1, (float)(nu - 1));
        f.iv0 = (int)fmaxf(c0, 0.0f);
        f.iv1 = (int)fminf(c1, (float)(nv - 1));
        if (a1 < 0.0f || a0 > (float)(nu - 1)) f.iu1 = f.iu0 - 1;
        if (c1 < 0.0f || c0 > (float)(nv - 1)) f.iv1 = f.iv0 - 1;
        f.su = (int)floorf(fu * 1.05f + stride_margin) + 1;
        f.sv = nv == 1 ? 1 : (int)floorf(fv * 1.05f + stride_margin) + 1;
        // a stride wider than the footprint already puts every column
        // (row) in its own phase
        f.su = min(f.su, max(f.iu1 - f.iu0 + 1, 1));
        f.sv = min(f.sv, max(f.iv1 - f.iv0 + 1, 1));
    }
    return f;
}

// Each wave owns one sub-brick of SBX x SBY x SBZ voxels with a private
// on-chip accumulator; a workgroup of four waves covers a 2x2 block of
// sub-bricks (neighbouring sub-bricks share rays through L1/L2).  Per view a
// wave projects its sub-brick's guard box onto the detector, then walks the
// footprint rays colour by colour: rays of one colour are at least one voxel
// footprint apart, so the lanes of a wave never update the same voxel in the
// same colour phase, and consecutive phases of one wave are ordered by the
// in-order LDS pipeline -- no atomics and no workgroup barriers.
//
// NB sinograms that share the rays (sinogram j at sino + j*sstride, volume j
// at out + j*nvox): one trace per candidate ray serves all of them, each
// into its own accumulator set, and every volume receives the additions of
// an NB = 1 run in the same order (a ray whose weight is zero for one
// sinogram adds a zero product to it, which leaves every sum's bits as they
// were).
// (NB accumulator sets take NB times the LDS, so NB > 1 runs at two waves
// per SIMD.)
template <int SBX, int SBY, int SBZ, int WX, int WY, int WPE, int NS, bool P2, bool ISO, bool UNG, int NB = 1>
__global__ __launch_bounds__(64 * WX * WY) __attribute__((amdgpu_waves_per_eu(NB == 1 ? WPE : 2))) void k_siddon_bp_wave(
    GridK g, const float* __restrict__ src, const float* __restrict__ trg, const float* __restrict__ rec,
    int nviews, int nu, int nv, const float* __restrict__ proj, const float* __restrict__ sino,
    float* __restrict__ out, int views_per_chunk, int nbx, int nby, int nbz, int64_t out_stride, SirtEpi E,
    int64_t sstride, int64_t nvox, GridK gu) {
    constexpr int kVox = SBX * SBY * SBZ;
    // unguarded rays on power-of-two grids are walked in voxel units (gu =
    // unit_grid(g); unit_exact holds for every ray with the precondition)
    constexpr bool UNIT = P2 && UNG;
    // NS views at a time per wave: lanes [slot*LPQ, (slot+1)*LPQ) work on view
    // va0 + slot and accumulate into accumulator copy `slot` of the
    // sub-brick (copies summed in slot order at the end), so rays of
    // different views never meet in one accumulator
    constexpr int LPQ = 64 / NS;
    __shared__ float acc_all[WX * WY * NB * NS * kVox];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int slot = lane / LPQ, lq = lane % LPQ;
    // accumulator copy `slot` of batch item j at acc_w + (j NS + slot) kVox
    float* acc_w = acc_all + wave * NB * NS * kVox;
    float* acc = acc_w + slot * kVox;

    // nbx/nby/nbz count workgroup regions (WX*SBX x WY*SBY x SBZ)
    const int region = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int rz = region % nbz;
    const int ry = (region / nbz) % nby;
    const int rx = region / (nbz * nby);
    const int b0[3] = {rx * WX * SBX + (wave % WX) * SBX, ry * WY * SBY + (wave / WX) * SBY, rz * SBZ};
    const int b1[3] = {min(b0[0] + SBX, g.n[0]), min(b0[1] + SBY, g.n[1]), min(b0[2] + SBZ, g.n[2])};
    const bool live = b0[0] < g.n[0] && b0[1] < g.n[1];

    for (int i = lane; i < NB * NS * kVox; i += 64) acc_w[i] = 0.0f;

    // guard box: the sub-brick grown by kGuard voxels.  An entry whose
    // computed index falls in the sub-brick has its true midpoint within
    // ~1e-3 voxel of it (f32 rounding of the midpoint position), so every
    // such entry belongs to a segment that meets the guard box.
    float gb_lo[3], gb_hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        gb_lo[a] = plane_coord(g, a, b0[a]) - kGuard * g.d[a];
        gb_hi[a] = plane_coord(g, a, b1[a]) + kGuard * g.d[a];
    }

    const int a_begin = blockIdx.y * views_per_chunk;
    const int a_end = live ? min(nviews, a_begin + views_per_chunk) : a_begin;

    for (int va0 = a_begin; va0 < a_end; va0 += NS) {
        // each lane keeps the footprint of its slot's view; slots past the
        // chunk's last view get an empty one
        const int nva = min(NS, a_end - va0);
        ViewFoot f = vir::view_foot<NS>(g, proj, va0, nva, gb_lo, gb_hi, nu, nv, lane, slot);
        if (slot >= nva) {
            f.iu0 = 0; f.iu1 = -1; f.iv0 = 0; f.iv1 = -1; f.su = 1; f.sv = 1;
        }
        const int va = min(va0 + slot, a_end - 1);
        // this view's rays: ray (va, iu, iv) at in-view offset iu*nv + iv
        // (the host checks nu nv < 2^30, so 24-bit factors suffice and
        // 3 (iu nv + iv) fits 32 bits)
        const int64_t vbase = (int64_t)va * nu * nv;
        const float* __restrict__ sino_v = sino + vbase;
        const float* __restrict__ src_v = rec ? nullptr : src + 3 * vbase;
        const float* __restrict__ trg_v = rec ? nullptr : trg + 3 * vbase;
        const float* __restrict__ rec_v = rec ? rec + kRayRec * vbase : nullptr;
        const int nU = f.iu1 - f.iu0 + 1, nV = f.iv1 - f.iv0 + 1;
        if (nU <= 0 || nV <= 0) continue;
        // phase (cu, cv) holds the columns iu0 + du + k su, du = (cu - iu0 mod su)
        // mod su, of which there are mU = qU + [du < rU] (nU = qU su + rU), and
        // likewise the rows: per-pass constants, so a phase needs no division
        const int ru = f.iu0 % f.su, qU = nU / f.su, rU = nU % f.su;
        const int rv = f.iv0 % f.sv, qV = nV / f.sv, rV = nV % f.sv;
        auto phase = [&](int cu, int cv, int& us, int& vs, int& mU, int& mV) {
            int xu = cu - ru + f.su, xv = cv - rv + f.sv;
            xu = xu >= f.su ? xu - f.su : xu;
            xv = xv >= f.sv ? xv - f.sv : xv;
            us = f.iu0 + xu;
            vs = f.iv0 + xv;
            mU = qU + (xu < rU ? 1 : 0);
            mV = qV + (xv < rV ? 1 : 0);
        };
        // in-view offset of candidate q of a phase.  With fq (q < 2^21),
        // q / mV is floor((q + 0.5) r) with r the hardware reciprocal of mV
        // (within 1 ulp): the quotient's fraction lies in [0.5/mV, 1 - 0.5/mV]
        // and the roundings move it by less than (q + 0.5)/mV 2^-22 < 0.5/mV
        auto ray_at = [&](int us, int vs, int mV, float inv_mV, bool fq, int q) -> unsigned {
            int qu, qv;
            if (fq) {
                qu = (int)(((float)q + 0.5f) * inv_mV);
                qv = q - (int)__umul24((unsigned)qu, (unsigned)mV);
            } else {
                qu = q / mV;
                qv = q % mV;
            }
            const int iu = us + qu * f.su;
            const int iv = vs + qv * f.sv;
            return __umul24((unsigned)iu, (unsigned)nv) + (unsigned)iv;
        };
        for (int cu = 0; cu < f.su; ++cu) {
            for (int cv = 0; cv < f.sv; ++cv) {
                int us, vs, mU, mV;
                phase(cu, cv, us, vs, mU, mV);
                const int M = mU * mV;
                const bool fq = M < (1 << 21);
                const float inv_mV = __builtin_amdgcn_rcpf((float)mV);
                auto ray_of = [&](int q) -> unsigned { return ray_at(us, vs, mV, inv_mV, fq, q); };
                // the lane's next candidate and its weights, loaded one
                // candidate ahead so the load overlaps the current walk
                unsigned ro_next = lq < M ? ray_of(lq) : 0u;
                float w_next[NB];
#pragma unroll
                for (int j = 0; j < NB; ++j) w_next[j] = lq < M ? sino_v[j * sstride + ro_next] : 0.0f;
                for (int q = lq; q < M; q += LPQ) {
                    const unsigned ro = ro_next;
                    // zero-weight rays (they miss the object) skip the
                    // endpoint loads altogether
                    float wgt[NB];
                    bool live_w = false;
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        wgt[j] = w_next[j];
                        live_w |= wgt[j] != 0.0f;
                    }
                    if (q + LPQ < M) {
                        ro_next = ray_of(q + LPQ);
#pragma unroll
                        for (int j = 0; j < NB; ++j) w_next[j] = sino_v[j * sstride + ro_next];
                    }
                    if (!live_w) continue;
                    Ray R;
                    if (NB > 1 || rec) {
                        // precomputed setup and exact slices (uniform branch;
                        // batched calls always come with records)
                        ray_load_record(rec_v + (size_t)kRayRec * ro, R);
                    } else {
                        const float* sr = src_v + 3 * ro;
                        const float* tr = trg_v + 3 * ro;
                        f3v sp = {sr[0], sr[1], sr[2]};
                        f3v tp = {tr[0], tr[1], tr[2]};
                        ray_setup_core_v<UNG>(g, sp, tp, R);
                    }
                    if (!R.hit) continue;
                    // alpha interval of the ray inside the guard box
                    float alo = -VIR_INF, ahi = VIR_INF;
                    bool miss = false;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        if (R.dST[a] != 0.0f) {
                            float t1 = (gb_lo[a] - R.s[a]) * R.y[a];
                            float t2 = (gb_hi[a] - R.s[a]) * R.y[a];
                            alo = fmaxf(alo, fminf(t1, t2));
                            ahi = fminf(ahi, fmaxf(t1, t2));
                        } else if (R.s[a] < gb_lo[a] || R.s[a] > gb_hi[a]) {
                            miss = true;
                        }
                    }
                    if (miss || alo > ahi) continue;
                    // exact plane slices only matter near the ray's ends
                    if (NB == 1 && !rec && !window_interior(g, R, alo, ahi)) ray_setup_slices(g, R);
                    const bool mid = UNG && window_mid(g, R, alo, ahi);
                    if (UNIT) unit_ray(g, R);
                    const GridK& gw = UNIT ? gu : g;
                    walk_dispatch<UNG>(gw, R, [&](auto fast) {
                        constexpr bool F = decltype(fast)::value;
                        Walker W;
                        walker_seek<F>(gw, R, W, alo);
                        auto visit = [&](int ix, int iy, int iz, float len) {
                            int lx = ix - b0[0], ly = iy - b0[1], lz = iz - b0[2];
                            if ((unsigned)lx < (unsigned)(b1[0] - b0[0]) &&
                                (unsigned)ly < (unsigned)(b1[1] - b0[1]) &&
                                (unsigned)lz < (unsigned)(b1[2] - b0[2])) {
                                const int li = (lx * SBY + ly) * SBZ + lz;
#pragma unroll
                                for (int j = 0; j < NB; ++j) acc[j * NS * kVox + li] += wgt[j] * len;
                            }
                        };
                        // windows away from the rays' ends: the walk without
                        // the special values (same entries) when every
                        // walking lane's window qualifies (in the kernels
                        // for unguarded ray sets; the guarded ones keep their
                        // registers for the float64 and range-checked paths)
                        if constexpr (F && UNG) {
                            if (__ballot(!mid) == 0) {
                                walk_mid<P2, ISO, false, false, UNIT>(gw, R, W, ahi, visit);
                                return;
                            }
                        }
                        walk_entries<P2, ISO, false, F, false, UNIT && F>(gw, R, W, ahi, visit);
                    });
                }
                // next colour phase: keep the compiler from moving LDS accesses
                // across it (the hardware keeps one wave's LDS accesses in order)
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        }
    }

    // write the sub-brick (or this view chunk's partial sub-brick)
    if (!live) return;
    float* dst = out + (int64_t)blockIdx.y * out_stride;
    const int ex = b1[0] - b0[0], ey = b1[1] - b0[1], ez = b1[2] - b0[2];
    for (int i = lane; i < ex * ey * ez; i += 64) {
        int lz = i % ez, ly = (i / ez) % ey, lx = i / (ez * ey);
        int64_t gi = ((int64_t)(b0[0] + lx) * g.n[1] + (b0[1] + ly)) * g.n[2] + (b0[2] + lz);
        const int li = (lx * SBY + ly) * SBZ + lz;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            const float* aj = acc_w + j * NS * kVox;
            float sum = aj[li];
#pragma unroll
            for (int s = 1; s < NS; ++s) sum += aj[s * kVox + li];
            sirt_store(sirt_item(E, j * nvox), dst + j * nvox, gi, sum);
        }
    }
}

// part: nchunks partial volumes of n voxels, chunk k at part + k*cstride.
__global__ void k_sum_partials(const float* __restrict__ part, int nchunks, int64_t n, int64_t cstride,
                               float* __restrict__ vol, int accumulate, SirtEpi E) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = accumulate ? vol[i] : 0.0f;
    float c = 0.0f;
    for (int k = 0; k < nchunks; ++k) {
        float t = part[(int64_t)k * cstride + i];
        float y = s + t;
        if (fabsf(s) >= fabsf(t)) c += (s - y) + t;
        else c += (t - y) + s;
        s = y;
    }
    sirt_store(E, vol, i, s + c);
}

// Sub-brick shapes (one per wave, four waves per workgroup): 3D grids use
// 16x8x8 voxels with two accumulator copies (two views per pass, 8 KiB of
// f32 accumulators per wave), flat (nZ == 1) grids 32x16x1 with four (four
// views per pass, 8 KiB).  Views are split into chunks until there are
// enough workgroups to fill the chip; chunk partial volumes are summed in a
// fixed order.
// Sub-brick shapes of the backprojector: per-wave brick SBX x SBY x SBZ,
// WX x WY waves per workgroup, WPE waves per SIMD the register budget allows,
// NS views per wave pass.  For the 256^3 cone-beam case 16 x 8 x 8 with two
// views per pass and 2 x 2 waves was the fastest measured (one view per
// pass: 8^3 to 32 x 32 x 8; two or four views: 8^3 to 32 x 16 x 2 within
// 8 KiB per wave; 1 x 4 and 4 x 1 waves); for 2-D grids 32 x 16 with four
// views per pass (16 x 16 to 64 x 16, two to eight views, 1024^2 parallel).
template <int SBX, int SBY, int SBZ, int WX, int WY, int WPE, int NS>
struct BpShape {
    static constexpr int sbx = SBX, sby = SBY, sbz = SBZ, wx = WX, wy = WY;
};
// 2-D sub-bricks see few rays of one view per colour phase, so a wave works
// on four views at once, each into its own accumulator copy
using BpShape3 = BpShape<16, 8, 8, 2, 2, 4, 2>;
using BpShape2 = BpShape<32, 16, 1, 2, 2, 4, 4>;

struct BpPlan {
    int kind;  // 0: 3D sub-bricks, 1: 2D sub-bricks
    int sbx, sby, sbz, wx, wy;
    int nbx, nby, nbz;  // workgroup regions (wx*sbx x wy*sby x sbz)
    int nchunks, views_per_chunk;
};

static BpPlan plan_bp(const GridK& g, int nviews) {
    BpPlan p;
    if (g.n[2] == 1) {
        p.kind = 1; p.sbx = BpShape2::sbx; p.sby = BpShape2::sby; p.sbz = BpShape2::sbz;
        p.wx = BpShape2::wx; p.wy = BpShape2::wy;
    } else {
        p.kind = 0; p.sbx = BpShape3::sbx; p.sby = BpShape3::sby; p.sbz = BpShape3::sbz;
        p.wx = BpShape3::wx; p.wy = BpShape3::wy;
    }
    p.nbx = (g.n[0] + p.wx * p.sbx - 1) / (p.wx * p.sbx);
    p.nby = (g.n[1] + p.wy * p.sby - 1) / (p.wy * p.sby);
    p.nbz = (g.n[2] + p.sbz - 1) / p.sbz;
    int64_t nb = (int64_t)p.nbx * p.nby * p.nbz;
    const int target = 1024;
    int nch = (int)std::min<int64_t>(nviews, std::max<int64_t>(1, (target + nb - 1) / nb));
    p.views_per_chunk = (nviews + nch - 1) / nch;
    p.nchunks = (nviews + p.views_per_chunk - 1) / p.views_per_chunk;
    return p;
}

}  // namespace vir

using namespace vir;

extern "C" {

static bool all_pow2(const GridK& g) { return g.pow2[0] && g.pow2[1] && g.pow2[2]; }
static bool isotropic(const GridK& g) {
    return g.n[0] == g.n[1] && g.n[1] == g.n[2] && g.d[0] == g.d[1] && g.d[1] == g.d[2] &&
           g.o[0] == g.o[1] && g.o[1] == g.o[2];
}

// Launch KERNEL<P2, ISO> with the template flags chosen from the grid.
#define VIR_DISPATCH_RAY_KERNEL(KERNEL, G, NB, ...)                                                       \
    do {                                                                                                    \
        bool p2_ = all_pow2(G), iso_ = isotropic(G);                                                       \
        if (p2_ && iso_) hipLaunchKernelGGL((KERNEL<true, true>), __VA_ARGS__);                             \
        else if (p2_) hipLaunchKernelGGL((KERNEL<true, false>), __VA_ARGS__);                               \
        else if (iso_) hipLaunchKernelGGL((KERNEL<false, true>), __VA_ARGS__);                              \
        else hipLaunchKernelGGL((KERNEL<false, false>), __VA_ARGS__);                                       \
    } while (0)

static int check_rays(const float* src, const float* trg, int64_t nrays) {
    if (nrays < 0) return fail_arg("nrays < 0");
    if (nrays > 0 && (!src || !trg)) return fail_arg("src/trg is NULL");
    return VIR_OK;
}

int vir_siddon_count(const float* src, const float* trg, int64_t nrays, const vir_grid* grid,
                     int32_t* counts, uint8_t* flags, void* stream) {
    GridK g;
    int rc = make_grid(grid, g);
    if (rc) return rc;
    if ((rc = check_rays(src, trg, nrays))) return rc;
    if (nrays == 0) return VIR_OK;
    if (!counts || !flags) return fail_arg("counts/flags is NULL");
    int64_t nb = (nrays + kRayThreads - 1) / kRayThreads;
    VIR_DISPATCH_RAY_KERNEL(k_siddon_count, g, nb, dim3((unsigned)nb), dim3(kRayThreads), 0, as_stream(stream),
                            g, src, trg, nrays, counts, flags);
    return check_hip(hipGetLastError(), "vir_siddon_count launch");
}

int vir_siddon_fill(const float* src, const float* trg, int64_t nrays, const vir_grid* grid,
                    const int64_t* offsets, int32_t* ravel_idx, int32_t* ix, int32_t* iy, int32_t* iz,
                    float* len, void* stream) {
    GridK g;
    int rc = make_grid(grid, g);
    if (rc) return rc;
    if ((rc = check_rays(src, trg, nrays))) return rc;
    if (nrays == 0) return VIR_OK;
    if (!offsets) return fail_arg("offsets is NULL");
    int64_t nb = (nrays + kRayThreads - 1) / kRayThreads;
    VIR_DISPATCH_RAY_KERNEL(k_siddon_fill, g, nb, dim3((unsigned)nb), dim3(kRayThreads), 0, as_stream(stream), g,
                            src, trg, nrays, offsets, ravel_idx, ix, iy, iz, len);
    return check_hip(hipGetLastError(), "vir_siddon_fill launch");
}

// The forward projector from endpoints (rec == NULL) or from per-ray
// records (src == trg == NULL).
// Largest number of volumes one FP launch traces together; larger batches
// run as groups of at most this many.
constexpr int kFpBatch = 4;

extern "C++" {
template <int NB>
static void siddon_fp_launch(const GridK& g, const float* src, const float* trg, const float* rec, int64_t nrays,
                             const float* vol, float* sino, int acc, int32_t* err_count, hipStream_t st,
                             const SirtEpi& E, int64_t nvox, bool p2, bool iso, bool ung) {
    const int64_t nb = (nrays + kRayThreads - 1) / kRayThreads;
#define VIR_FP_LAUNCH(P2_, ISO_, UNG_)                                                                       \
    hipLaunchKernelGGL((k_siddon_fp<P2_, ISO_, UNG_, NB>), dim3((unsigned)nb), dim3(kRayThreads), 0, st, g,  \
                       src, trg, rec, nrays, vol, sino, acc, err_count, E, nvox, unit_grid(g))
    if (ung) {
        if (p2 && iso) VIR_FP_LAUNCH(true, true, true);
        else if (p2) VIR_FP_LAUNCH(true, false, true);
        else if (iso) VIR_FP_LAUNCH(false, true, true);
        else VIR_FP_LAUNCH(false, false, true);
    } else {
        if (p2 && iso) VIR_FP_LAUNCH(true, true, false);
        else if (p2) VIR_FP_LAUNCH(true, false, false);
        else if (iso) VIR_FP_LAUNCH(false, true, false);
        else VIR_FP_LAUNCH(false, false, false);
    }
#undef VIR_FP_LAUNCH
}
}  // extern "C++"

// nbatch volumes (volume j at vol + j*nvox, projection at sino + j*nrays;
// records only when nbatch > 1), traced in groups of kFpBatch, 2 and 1.
static int siddon_fp(const float* src, const float* trg, const float* rec, int64_t nrays, const vir_grid* grid,
                     const float* vol, float* sino, uint32_t flags, int32_t* err_count, void* stream,
                     const SirtEpi& E = SirtEpi{}, int32_t nbatch = 1) {
    GridK g;
    int rc = make_grid(grid, g);
    if (rc) return rc;
    if (nrays < 0) return fail_arg("nrays < 0");
    if (nbatch < 1) return fail_arg("nbatch must be >= 1");
    if (nrays > 0 && !(rec || (src && trg))) return fail_arg("src/trg is NULL");
    if (nbatch > 1 && nrays > 0 && !rec) return fail_arg("batched projection needs ray records");
    const int64_t nvox = (int64_t)g.n[0] * g.n[1] * g.n[2];
    if (nvox >= ((int64_t)1 << 32)) return fail_arg("volume must have fewer than 2^32 voxels");
    if (nrays == 0) return VIR_OK;
    if (!vol || !sino) return fail_arg("vol/sino is NULL");
    int acc = (flags & VIR_ACCUMULATE) ? 1 : 0;
    const bool ung = (flags & VIR_UNGUARDED) != 0, iso = isotropic(g);
    // unguarded power-of-two kernels walk in voxel units: g.unit grids only
    const bool p2 = all_pow2(g) && (!ung || g.unit);
    hipStream_t st = as_stream(stream);
    for (int32_t j0 = 0; j0 < nbatch;) {
        const int32_t left = nbatch - j0;
        const int nb = left >= kFpBatch ? kFpBatch : (left >= 2 ? 2 : 1);
        const float* vj = vol + j0 * nvox;
        float* sj = sino + j0 * nrays;
        const SirtEpi Ej = sirt_item(E, j0 * nrays);
        if (nb == kFpBatch) siddon_fp_launch<kFpBatch>(g, src, trg, rec, nrays, vj, sj, acc, err_count, st, Ej, nvox, p2, iso, ung);
        else if (nb == 2) siddon_fp_launch<2>(g, src, trg, rec, nrays, vj, sj, acc, err_count, st, Ej, nvox, p2, iso, ung);
        else siddon_fp_launch<1>(g, src, trg, rec, nrays, vj, sj, acc, err_count, st, Ej, nvox, p2, iso, ung);
        rc = check_hip(hipGetLastError(), "vir_siddon_fp launch");
        if (rc) return rc;
        // out-of-grid entries are a property of the rays: count them once
        err_count = nullptr;
        j0 += nb;
    }
    return VIR_OK;
}

int vir_siddon_fp(const float* src, const float* trg, int64_t nrays, const vir_grid* grid, const float* vol,
                  float* sino, uint32_t flags, int32_t* err_count, void* stream) {
    if (nrays > 0 && (!src || !trg)) return fail_arg("src/trg is NULL");
    return siddon_fp(src, trg, nullptr, nrays, grid, vol, sino, flags, err_count, stream);
}

int vir_siddon_fp_rec(const float* rec, int64_t nrays, const vir_grid* grid, const float* vol, float* sino,
                      uint32_t flags, int32_t* err_count, void* stream) {
    if (nrays > 0 && !rec) return fail_arg("rec is NULL");
    if (reinterpret_cast<uintptr_t>(rec) % 16) return fail_arg("rec must be 16-byte aligned");
    return siddon_fp(nullptr, nullptr, rec, nrays, grid, vol, sino, flags, err_count, stream);
}

int vir_siddon_count_guarded(const float* src, const float* trg, int64_t nrays, const vir_grid* grid,
                             int32_t* count, void* stream) {
    GridK g;
    int rc = make_grid(grid, g);
    if (rc) return rc;
    if ((rc = check_rays(src, trg, nrays))) return rc;
    if (!count) return fail_arg("count is NULL");
    if (nrays == 0) return VIR_OK;
    int64_t nb = (nrays + kRayThreads - 1) / kRayThreads;
    hipLaunchKernelGGL(k_siddon_count_guarded, dim3((unsigned)nb), dim3(kRayThreads), 0, as_stream(stream), g, src,
                       trg, nrays, count);
    return check_hip(hipGetLastError(), "vir_siddon_count_guarded launch");
}

int64_t vir_siddon_bp_workspace(const vir_grid* grid, int32_t nviews) {
    GridK g;
    if (make_grid(grid, g) || nviews < 1) return -1;
    BpPlan p = plan_bp(g, nviews);
    if (p.nchunks <= 1) return 0;
    int64_t nvox = (int64_t)g.n[0] * g.n[1] * g.n[2];
    return (int64_t)p.nchunks * nvox * (int64_t)sizeof(float);
}

// Largest number of sinograms one BP launch backprojects together (each
// adds a set of sub-brick accumulators to the workgroup's LDS); larger
// batches run as groups.
constexpr int kBpBatch = 2;

// The brick-gather backprojector from endpoints (rec == NULL) or from
// per-ray records (src == trg == NULL), for nbatch sinograms (sinogram j at
// sino + j*nviews*nu*nv, volume j at vol + j*nvox; records only when
// nbatch > 1), backprojected in groups of kBpBatch and 1.  The workspace
// holds one group's staging.
static int siddon_bp(const float* src, const float* trg, const float* rec, int32_t nviews, int32_t nu,
                     int32_t nv, const float* proj, const vir_grid* grid, const float* sino, float* vol,
                     uint32_t flags, void* workspace, void* stream, const SirtEpi& E = SirtEpi{},
                     int32_t nbatch = 1) {
    GridK g;
    int rc = make_grid(grid, g);
    if (rc) return rc;
    if (nviews < 1 || nu < 1 || nv < 1) return fail_arg("nviews/nu/nv must be >= 1");
    if ((int64_t)nu * nv >= ((int64_t)1 << 30)) return fail_arg("nu*nv must be < 2^30");
    if (nbatch < 1) return fail_arg("nbatch must be >= 1");
    if (!(rec || (src && trg)) || !proj || !sino || !vol) return fail_arg("NULL pointer argument");
    if (nbatch > 1 && !rec) return fail_arg("batched backprojection needs ray records");
    BpPlan p = plan_bp(g, nviews);
    const int64_t nvox = (int64_t)g.n[0] * g.n[1] * g.n[2];
    const int64_t nrays = (int64_t)nviews * nu * nv;
    hipStream_t st = as_stream(stream);
    const bool direct = (p.nchunks == 1) && !(flags & VIR_ACCUMULATE);
    if (!direct) {
        if (p.nchunks > 1 && !workspace) return fail_arg("workspace required (vir_siddon_bp_workspace)");
        // accumulate with a single chunk: stage through the workspace
        if (p.nchunks == 1 && !workspace) return fail_arg("workspace required for VIR_ACCUMULATE");
    }
    const int nb = p.nbx * p.nby * p.nbz;
    const dim3 grd((unsigned)nb, (unsigned)p.nchunks);
    const bool ung = (flags & VIR_UNGUARDED) != 0, iso = isotropic(g);
    // unguarded power-of-two kernels walk in voxel units: g.unit grids only
    const bool p2 = all_pow2(g) && (!ung || g.unit);
    for (int32_t j0 = 0; j0 < nbatch;) {
        const int NBg = nbatch - j0 >= kBpBatch ? kBpBatch : 1;
        const float* sino_g = sino + j0 * nrays;
        float* vol_g = vol + j0 * nvox;
        const SirtEpi Eg = sirt_item(E, j0 * nvox);
        // staging: chunk k of item j at out + (k NBg + j) nvox
        float* out = direct ? vol_g : reinterpret_cast<float*>(workspace);
        const int64_t out_stride = (int64_t)NBg * nvox;
        // the fused SIRT step goes where the final volume is written
        const SirtEpi Ek = direct ? Eg : SirtEpi{};
#define VIR_BP_LAUNCH1(SHAPE, UNG_, NB_)                                                                      \
    do {                                                                                                        \
        dim3 blk(64 * (SHAPE##_WX) * (SHAPE##_WY));                                                             \
        if (p2 && iso) hipLaunchKernelGGL((k_siddon_bp_wave<SHAPE##_ARGS, true, true, UNG_, NB_>), grd, blk, 0,  \
                                          st, VIR_BP_ARGS);                                                     \
        else if (p2) hipLaunchKernelGGL((k_siddon_bp_wave<SHAPE##_ARGS, true, false, UNG_, NB_>), grd, blk, 0,   \
                                        st, VIR_BP_ARGS);                                                       \
        else hipLaunchKernelGGL((k_siddon_bp_wave<SHAPE##_ARGS, false, false, UNG_, NB_>), grd, blk, 0, st,      \
                                VIR_BP_ARGS);                                                                   \
    } while (0)
#define VIR_BP_LAUNCH(SHAPE)                                                                                  \
    do {                                                                                                        \
        if (NBg == kBpBatch) {                                                                                  \
            if (ung) VIR_BP_LAUNCH1(SHAPE, true, kBpBatch);                                                     \
            else VIR_BP_LAUNCH1(SHAPE, false, kBpBatch);                                                        \
        } else {                                                                                                \
            if (ung) VIR_BP_LAUNCH1(SHAPE, true, 1);                                                            \
            else VIR_BP_LAUNCH1(SHAPE, false, 1);                                                               \
        }                                                                                                       \
    } while (0)
#define VIR_BP_ARGS                                                                                                   \
    g, src, trg, rec, nviews, nu, nv, proj, sino_g, out, p.views_per_chunk, p.nbx, p.nby, p.nbz, out_stride, Ek, nrays, \
        nvox, unit_grid(g)
#define BP3_ARGS 16, 8, 8, 2, 2, 4, 2
#define BP3_WX BpShape3::wx
#define BP3_WY BpShape3::wy
#define BP2_ARGS 32, 16, 1, 2, 2, 4, 4
#define BP2_WX 2
#define BP2_WY 2
        if (p.kind == 0) VIR_BP_LAUNCH(BP3);
        else VIR_BP_LAUNCH(BP2);
#undef BP3_ARGS
#undef BP3_WX
#undef BP3_WY
#undef BP2_ARGS
#undef BP2_WX
#undef BP2_WY
#undef VIR_BP_ARGS
#undef VIR_BP_LAUNCH
#undef VIR_BP_LAUNCH1
        rc = check_hip(hipGetLastError(), "vir_siddon_bp launch");
        if (rc) return rc;
        if (!direct) {
            const int64_t nbk = (nvox + 255) / 256;
            for (int j = 0; j < NBg; ++j) {
                hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)nbk), dim3(256), 0, st, out + j * nvox, p.nchunks,
                                   nvox, out_stride, vol_g + j * nvox, (flags & VIR_ACCUMULATE) ? 1 : 0,
                                   sirt_item(Eg, j * nvox));
                rc = check_hip(hipGetLastError(), "vir_siddon_bp reduce launch");
                if (rc) return rc;
            }
        }
        j0 += NBg;
    }
    return VIR_OK;
}

int vir_siddon_bp(const float* src, const float* trg, int32_t nviews, int32_t nu, int32_t nv,
                  const float* proj, const vir_grid* grid, const float* sino, float* vol, uint32_t flags,
                  void* workspace, void* stream) {
    if (!src || !trg) return fail_arg("NULL pointer argument");
    return siddon_bp(src, trg, nullptr, nviews, nu, nv, proj, grid, sino, vol, flags, workspace, stream);
}

int vir_siddon_bp_rec(const float* rec, int32_t nviews, int32_t nu, int32_t nv, const float* proj,
                      const vir_grid* grid, const float* sino, float* vol, uint32_t flags, void* workspace,
                      void* stream) {
    if (!rec) return fail_arg("NULL pointer argument");
    if (reinterpret_cast<uintptr_t>(rec) % 16) return fail_arg("rec must be 16-byte aligned");
    return siddon_bp(nullptr, nullptr, rec, nviews, nu, nv, proj, grid, sino, vol, flags, workspace, stream);
}

int vir_siddon_fp_rec_res(const float* rec, int64_t nrays, const vir_grid* grid, const float* vol, const float* b,
                          const float* w_sino, float* r, uint32_t flags, void* stream) {
    if (nrays > 0 && (!rec || !b || !w_sino)) return fail_arg("NULL pointer argument");
    if (reinterpret_cast<uintptr_t>(rec) % 16) return fail_arg("rec must be 16-byte aligned");
    if (flags & VIR_ACCUMULATE) return fail_arg("VIR_ACCUMULATE does not apply to the fused residual");
    return siddon_fp(nullptr, nullptr, rec, nrays, grid, vol, r, flags, nullptr, stream,
                     SirtEpi{b, w_sino, nullptr, 0.0f});
}

int vir_siddon_bp_rec_upd(const float* rec, int32_t nviews, int32_t nu, int32_t nv, const float* proj,
                          const vir_grid* grid, const float* r, const float* w_img, float gamma, float* x,
                          uint32_t flags, void* workspace, void* stream) {
    if (!rec || !w_img || !x) return fail_arg("NULL pointer argument");
    if (reinterpret_cast<uintptr_t>(rec) % 16) return fail_arg("rec must be 16-byte aligned");
    if (flags & VIR_ACCUMULATE) return fail_arg("VIR_ACCUMULATE does not apply to the fused update");
    return siddon_bp(nullptr, nullptr, rec, nviews, nu, nv, proj, grid, r, x, flags, workspace, stream,
                     SirtEpi{nullptr, w_img, x, gamma});
}

// ---------------------------------------------- batches sharing the rays

int vir_siddon_fp_rec_batch(const float* rec, int64_t nrays, const vir_grid* grid, int32_t nbatch,
                            const float* vol, float* sino, uint32_t flags, int32_t* err_count, void* stream) {
    if (nrays > 0 && !rec) return fail_arg("rec is NULL");
    if (reinterpret_cast<uintptr_t>(rec) % 16) return fail_arg("rec must be 16-byte aligned");
    return siddon_fp(nullptr, nullptr, rec, nrays, grid, vol, sino, flags, err_count, stream, SirtEpi{}, nbatch);
}

int64_t vir_siddon_bp_batch_workspace(const vir_grid* grid, int32_t nviews, int32_t nbatch) {
    if (nbatch < 1) return -1;
    const int64_t one = vir_siddon_bp_workspace(grid, nviews);
    return one < 0 ? one : one * (nbatch < kBpBatch ? nbatch : kBpBatch);
}

int vir_siddon_bp_rec_batch(const float* rec, int32_t nviews, int32_t nu, int32_t nv, const float* proj,
                            const vir_grid* grid, int32_t nbatch, const float* sino, float* vol, uint32_t flags,
                            void* workspace, void* stream) {
    if (!rec) return fail_arg("NULL pointer argument");
    if (reinterpret_cast<uintptr_t>(rec) % 16) return fail_arg("rec must be 16-byte aligned");
    return siddon_bp(nullptr, nullptr, rec, nviews, nu, nv, proj, grid, sino, vol, flags, workspace, stream,
                     SirtEpi{}, nbatch);
}

int vir_siddon_fp_rec_res_batch(const float* rec, int64_t nrays, const vir_grid* grid, int32_t nbatch,
                                const float* x, const float* b, const float* w_sino, float* r, uint32_t flags,
                                void* stream) {
    if (nrays > 0 && (!rec || !b || !w_sino)) return fail_arg("NULL pointer argument");
    if (reinterpret_cast<uintptr_t>(rec) % 16) return fail_arg("rec must be 16-byte aligned");
    if (flags & VIR_ACCUMULATE) return fail_arg("VIR_ACCUMULATE does not apply to the fused residual");
    return siddon_fp(nullptr, nullptr, rec, nrays, grid, x, r, flags, nullptr, stream,
                     SirtEpi{b, w_sino, nullptr, 0.0f}, nbatch);
}

int vir_siddon_bp_rec_upd_batch(const float* rec, int32_t nviews, int32_t nu, int32_t nv, const float* proj,
                                const vir_grid* grid, int32_t nbatch, const float* r, const float* w_img,
                                float gamma, float* x, uint32_t flags, void* workspace, void* stream) {
    if (!rec || !w_img || !x) return fail_arg("NULL pointer argument");
    if (reinterpret_cast<uintptr_t>(rec) % 16) return fail_arg("rec must be 16-byte aligned");
    if (flags & VIR_ACCUMULATE) return fail_arg("VIR_ACCUMULATE does not apply to the fused update");
    return siddon_bp(nullptr, nullptr, rec, nviews, nu, nv, proj, grid, r, x, flags, workspace, stream,
                     SirtEpi{nullptr, w_img, x, gamma}, nbatch);
}

}  // extern "C"

// Staging memory of vir_siddon_bp_rec_noalloc: the view chunks' partial
// volumes (or the volume staged for VIR_ACCUMULATE), one array per GPU for
// the life of the process.  Host code must not use this name as a device
// address; it uses the address hipGetSymbolAddress returns for the current
// device (bp_scratch_ptr).
static __device__ __align__(256) float scratch[VIR_SIDDON_BP_SCRATCH_BYTES / sizeof(float)];

namespace vir {

static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // per device ordinal, resolved once

// The current device's copy of `scratch`, resolved on the first call on
// that device (with it current).
static int bp_scratch_ptr(void** out) {
    int dev = 0;
    int rc = check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        rc = check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (rc) return rc;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return VIR_OK;
}

}  // namespace vir

extern "C" {

int vir_siddon_bp_rec_noalloc(const float* rec, int32_t nviews, int32_t nu, int32_t nv, const float* proj,
                              const vir_grid* grid, const float* sino, float* vol, uint32_t flags,
                              void* stream) {
    GridK g;
    int rc = make_grid(grid, g);
    if (rc) return rc;
    if (nviews < 1) return fail_arg("nviews must be >= 1");
    // the staging this call needs, as vir_siddon_bp_rec would take it from
    // its workspace argument
    const BpPlan p = plan_bp(g, nviews);
    const int64_t nvox = (int64_t)g.n[0] * g.n[1] * g.n[2];
    int64_t need = 0;
    if (p.nchunks > 1) need = (int64_t)p.nchunks * nvox * (int64_t)sizeof(float);
    else if (flags & VIR_ACCUMULATE) need = nvox * (int64_t)sizeof(float);
    if (need > (int64_t)VIR_SIDDON_BP_SCRATCH_BYTES)
        return fail_arg("vir_siddon_bp_rec_noalloc: needs " + std::to_string(need) +
                        " bytes of staging, more than its fixed " +
                        std::to_string((int64_t)VIR_SIDDON_BP_SCRATCH_BYTES) +
                        " (use vir_siddon_bp_rec with a workspace)");
    void* ws = nullptr;
    if (need > 0 && (rc = bp_scratch_ptr(&ws))) return rc;
    return vir_siddon_bp_rec(rec, nviews, nu, nv, proj, grid, sino, vol, flags, ws, stream);
}

int vir_siddon_ray_records(const float* src, const float* trg, int64_t nrays, const vir_grid* grid, float* rec,
                           void* stream) {
    GridK g;
    int rc = make_grid(grid, g);
    if (rc) return rc;
    if ((rc = check_rays(src, trg, nrays))) return rc;
    if (nrays == 0) return VIR_OK;
    if (!rec) return fail_arg("rec is NULL");
    if (reinterpret_cast<uintptr_t>(rec) % 16) return fail_arg("rec must be 16-byte aligned");
    int64_t nb = (nrays + kRayThreads - 1) / kRayThreads;
    hipLaunchKernelGGL(k_siddon_ray_records, dim3((unsigned)nb), dim3(kRayThreads), 0, as_stream(stream), g, src,
                       trg, nrays, rec);
    return check_hip(hipGetLastError(), "vir_siddon_ray_records launch");
}

}  // extern "C"
