// Pixel-driven separable-footprint projectors for gfx950 (sys_mat/pd.py):
// 2-D parallel and fan, 3-D cone, and a batched 3-D parallel variant whose
// detector rows run along z.
//
// Footprint model: each pixel's four corners are projected onto the detector
// and the pixel deposits over [min, max] of them with amplitude
//   parallel : 1/(|sin| + |cos|)                           (pd.py:181)
//   fan      : 1/(|cos(a+g)| + |sin(a+g)|), g = atan(u_c/DSD) (pd.py:292-298)
//   cone     : 1/(|sin| + |cos|), u from the four parallel-projected
//              corners times the centre's fan magnification, v from the
//              slab bounds times magnification/cos(g)   (pd.py:423-451)
// weight per bin = overlap/du (* overlap_v/dv); the forward output is
// scaled by d_pix (pd.py:234, 350, 485).
//
// Forward: one thread per (batch, view, bin[, row]) gathers the pixels whose
// footprints meet the bin (centre projection inverted along the driving axis,
// widened by the largest footprint half-width).  Backward: one thread per
// (batch, pixel[, slab]) visits the bins its footprint meets.  Both evaluate
// every weight with the same functions: exact transposes, no atomics.
//
// VIR_COMPAT (fan forward only): the reference skips a zero pixel without
// advancing its o boundary (pd.py:321-325), so the next nonzero pixel's
// amplitude takes its lower y boundary from just after the previous nonzero
// pixel of the column (the first boundary if none); footprints are unchanged.
#include <cmath>
#include <type_traits>

#include "vir_common.h"

namespace vir {

enum SFMode { SF_PAR2 = 0, SF_FAN2 = 1, SF_CONE3 = 2, SF_PAR3 = 3 };

struct SFGeo {
    int nx, ny, nz, nu, nv, na, nb;
    float d, du, dv, su, sv, DSO, DSD;
    float out_scale;
    float wmax;  // largest footprint half-width along u (detector units)
    int unit_amp;  // parallel modes: amplitude 1 (VIR_SF_UNIT_AMPLITUDE)
    int compat;    // fan forward: stale o boundary after zero pixels (VIR_COMPAT)
};

__device__ __forceinline__ float sf_bnd(int n, float d, int k) { return d * ((float)k - 0.5f * (float)n); }
__device__ __forceinline__ float sf_det(int n, float du, float su, int j) {
    return du * (((float)j - 0.5f * (float)n) + su);
}

struct SFFoot {
    float u0, u1;   // footprint along u
    float amp;      // amplitude
    float cm;       // v magnification (cone) or 1
};

// Per-view constants.
struct SFView {
    float c, s, ang;
    float amp;  // 1/(|sin| + |cos|)
    float h;    // half the projected width of a pixel in parallel beam
};

__device__ __forceinline__ SFView sf_view(const SFGeo& G, float c, float s, float ang) {
    SFView V;
    V.c = c;
    V.s = s;
    V.ang = ang;
    const float w = fabsf(s) + fabsf(c);
    V.amp = G.unit_amp ? 1.0f : 1.0f / w;
    V.h = 0.5f * G.d * w;
    return V;
}

// Fan amplitude 1/(|cos(a+g)| + |sin(a+g)|) of pixel (ix, iy), g the angle
// of the ray through its centre, whose o coordinate takes the pixel's lower
// y boundary at boundary ky (ky = iy: the pixel's own).
__device__ __forceinline__ float sf_fan_amp(const SFGeo& G, const SFView& V, int ix, int iy, int ky) {
    const float c = V.c, s = V.s;
    const float xl = sf_bnd(G.nx, G.d, ix), xr = sf_bnd(G.nx, G.d, ix + 1);
    const float yl = sf_bnd(G.ny, G.d, iy), yr = sf_bnd(G.ny, G.d, iy + 1);
    const float pxc = (-s * xl + -s * xr) * 0.5f, oxc = (c * xl + c * xr) * 0.5f;
    const float pyc = (c * yl + c * yr) * 0.5f, oyc = (s * sf_bnd(G.ny, G.d, ky) + s * yr) * 0.5f;
    const float uc = G.DSD * (pxc + pyc) / (G.DSO - (oxc + oyc));
    const float gam = atanf(uc / G.DSD);
    const float rx = cosf(V.ang + gam), ry = sinf(V.ang + gam);
    return 1.0f / (fabsf(rx) + fabsf(ry));
}

// Footprint of pixel (ix, iy) in view V.  In parallel beam the four
// projected corners -s*x + c*y span exactly [p_c - h, p_c + h] around the
// centre's projection p_c, h = d(|sin| + |cos|)/2.
template <int MODE>
__device__ __forceinline__ SFFoot sf_foot(const SFGeo& G, const SFView& V, int ix, int iy) {
    const float c = V.c, s = V.s;
    SFFoot f;
    f.cm = 1.0f;
    f.amp = V.amp;
    if (MODE == SF_PAR2 || MODE == SF_PAR3) {
        const float xc = G.d * (((float)ix + 0.5f) - 0.5f * (float)G.nx);
        const float yc = G.d * (((float)iy + 0.5f) - 0.5f * (float)G.ny);
        const float pc = -s * xc + c * yc;
        f.u0 = pc - V.h;
        f.u1 = pc + V.h;
        return f;
    }
    const float xl = sf_bnd(G.nx, G.d, ix), xr = sf_bnd(G.nx, G.d, ix + 1);
    const float yl = sf_bnd(G.ny, G.d, iy), yr = sf_bnd(G.ny, G.d, iy + 1);
    const float pxl = -s * xl, pxr = -s * xr, pyl = c * yl, pyr = c * yr;
    const float oxl = c * xl, oxr = c * xr, oyl = s * yl, oyr = s * yr;
    if (MODE == SF_FAN2) {
        float p0 = G.DSD * (pxl + pyl) / (G.DSO - (oxl + oyl));
        float p1 = G.DSD * (pxr + pyl) / (G.DSO - (oxr + oyl));
        float p2 = G.DSD * (pxl + pyr) / (G.DSO - (oxl + oyr));
        float p3 = G.DSD * (pxr + pyr) / (G.DSO - (oxr + oyr));
        f.u0 = fminf(fminf(p0, p1), fminf(p2, p3));
        f.u1 = fmaxf(fmaxf(p0, p1), fmaxf(p2, p3));
        f.amp = sf_fan_amp(G, V, ix, iy, iy);
        return f;
    }
    // cone
    const float pxc = (pxl + pxr) * 0.5f, oxc = (oxl + oxr) * 0.5f;
    const float pyc = (pyl + pyr) * 0.5f, oyc = (oyl + oyr) * 0.5f;
    const float den_c = G.DSO - (oxc + oyc);
    const float uc = G.DSD * (pxc + pyc) / den_c;
    const float gam = atanf(uc / G.DSD);
    const float fm = G.DSD / den_c;
    float u0 = fm * (pxl + pyl), u1 = fm * (pxr + pyl), u2 = fm * (pxl + pyr), u3 = fm * (pxr + pyr);
    f.u0 = fminf(fminf(u0, u1), fminf(u2, u3));
    f.u1 = fmaxf(fmaxf(u0, u1), fmaxf(u2, u3));
    f.cm = fm / cosf(gam);
    return f;
}

// Rows (FP: slabs) a slab (row) of the parallel 3-D geometry meets, with
// their weights overlap/dv: independent of the pixel and the view, so
// computed once per thread.  Up to kSlabs entries -- enough for rows no
// taller than a voxel, rounding included; more (taller rows) fall back to
// the general loop.
constexpr int kSlabs = 5;
struct SlabWeights {
    int first, count;
    float w[kSlabs];
};

__device__ __forceinline__ float ovl(float a0, float a1, float b0, float b1) {
    return fminf(a1, b1) - fmaxf(a0, b0);
}

// Slabs that detector row iv of the parallel 3-D geometry meets, with their
// weights overlap/dv (zero-padded to kSlabs).  Returns false when the row
// meets more than kSlabs slabs (rows taller than a voxel).
__device__ __forceinline__ bool par3_row_slabs(const SFGeo& G, int iv, SlabWeights& SW) {
    const float vl = sf_det(G.nv, G.dv, G.sv, iv), vr = sf_det(G.nv, G.dv, G.sv, iv + 1);
    const float inv_d = 1.0f / G.d, inv_dv = 1.0f / G.dv;
    float z0f = fminf(vl, vr) * inv_d + 0.5f * (float)G.nz;
    float z1f = fmaxf(vl, vr) * inv_d + 0.5f * (float)G.nz;
    int z0 = (int)fmaxf(floorf(z0f) - 1.0f, 0.0f);
    int z1 = (int)fminf(floorf(z1f) + 1.0f, (float)(G.nz - 1));
    SW.first = z0;
    SW.count = z1 - z0 + 1;
    if (z1 - z0 + 1 > kSlabs) {
#pragma unroll
        for (int k = 0; k < kSlabs; ++k) SW.w[k] = 0.0f;
        return false;
    }
#pragma unroll
    for (int k = 0; k < kSlabs; ++k) {
        const int z = z0 + k;
        float ov = z <= z1 ? ovl(sf_bnd(G.nz, G.d, z), sf_bnd(G.nz, G.d, z + 1), vl, vr) : 0.0f;
        SW.w[k] = ov > 0.0f ? ov * inv_dv : 0.0f;
    }
    return true;
}

// Detector rows that slab iz of the parallel 3-D geometry meets, with their
// weights overlap/dv (zero-padded to kSlabs).  Returns false when the slab
// meets more than kSlabs rows (rows much shorter than a voxel).
__device__ __forceinline__ bool par3_slab_rows(const SFGeo& G, int iz, SlabWeights& SW) {
    const float zl = sf_bnd(G.nz, G.d, iz), zr = sf_bnd(G.nz, G.d, iz + 1);
    const float v0 = sf_det(G.nv, G.dv, G.sv, 0), inv_dv = 1.0f / G.dv;
    const float lo = fminf(zl, zr), hi = fmaxf(zl, zr);
    int m0 = (int)fmaxf(floorf((lo - v0) * inv_dv) - 1.0f, 0.0f);
    int m1 = (int)fminf(floorf((hi - v0) * inv_dv) + 1.0f, (float)(G.nv - 1));
    SW.first = m0;
    SW.count = m1 - m0 + 1;
    if (m1 - m0 + 1 > kSlabs) {
#pragma unroll
        for (int k = 0; k < kSlabs; ++k) SW.w[k] = 0.0f;
        return false;
    }
#pragma unroll
    for (int k = 0; k < kSlabs; ++k) {
        const int m = m0 + k;
        float ov = m <= m1 ? ovl(zl, zr, sf_det(G.nv, G.dv, G.sv, m), sf_det(G.nv, G.dv, G.sv, m + 1)) : 0.0f;
        SW.w[k] = ov > 0.0f ? ov * inv_dv : 0.0f;
    }
    return true;
}

// centre projection along the driving axis, inverted: the drive coordinate
// whose pixel centre projects to u on orthogonal line coordinate oc
template <int MODE>
__device__ __forceinline__ float sf_invert(const SFGeo& G, bool xd, float c, float s, float u, float oc) {
    // p = pd*drive + po*orth, o = od*drive + oo*orth (o toward the source)
    float pd = xd ? -s : c, po = xd ? c : -s;
    float od = xd ? c : s, oo = xd ? s : c;
    if (MODE == SF_PAR2 || MODE == SF_PAR3) return (u - po * oc) / pd;
    float t = u / G.DSD;
    return (t * (G.DSO - oo * oc) - po * oc) / (pd + t * od);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_sf_fp(SFGeo G, const float* __restrict__ ang,
                                              const float* __restrict__ cs, const float* __restrict__ sn,
                                              const float* __restrict__ img, float* __restrict__ sino,
                                              int accumulate) {
    const bool is3 = MODE == SF_CONE3 || MODE == SF_PAR3;
    const int nv = is3 ? G.nv : 1;
    const int64_t per_b = (int64_t)G.na * G.nu * nv;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per_b * G.nb) return;
    const int b = (int)(t / per_b);
    int64_t r = t % per_b;
    const int iv = (int)(r % nv);
    const int iu = (int)((r / nv) % G.nu);
    const int ia = (int)(r / ((int64_t)nv * G.nu));
    const SFView V = sf_view(G, cs[ia], sn[ia], ang[ia]);
    const float c = V.c, s = V.s;
    const bool xd = fabsf(s) >= fabsf(c);
    const int nd = xd ? G.nx : G.ny, no = xd ? G.ny : G.nx;
    const float ul = sf_det(G.nu, G.du, G.su, iu), ur = sf_det(G.nu, G.du, G.su, iu + 1);
    const float vl = sf_det(nv, G.dv, G.sv, iv), vr = sf_det(nv, G.dv, G.sv, iv + 1);
    const float inv_d = 1.0f / G.d, inv_du = 1.0f / G.du, inv_dv = 1.0f / G.dv;
    const int nz = is3 ? G.nz : 1;
    const float* vol = img + (int64_t)b * G.nx * G.ny * nz;
    // parallel 3-D: the slabs of this detector row and their weights
    SlabWeights SW;
    SW.first = 0;
    SW.count = 0;
    bool slab_general = false;
    if (MODE == SF_PAR3) slab_general = !par3_row_slabs(G, iv, SW);
    double acc = 0.0;
    for (int io = 0; io < no; ++io) {
        const float oc = (sf_bnd(no, G.d, io) + sf_bnd(no, G.d, io + 1)) * 0.5f;
        float x0 = sf_invert<MODE>(G, xd, c, s, ul - G.wmax, oc);
        float x1 = sf_invert<MODE>(G, xd, c, s, ur + G.wmax, oc);
        float k0f = fminf(x0, x1) * inv_d + 0.5f * (float)nd - 0.5f;
        float k1f = fmaxf(x0, x1) * inv_d + 0.5f * (float)nd - 0.5f;
        if (!(k0f < (float)nd + 1.0f) || !(k1f > -2.0f)) continue;
        int k0 = (int)fmaxf(floorf(k0f) - 1.0f, 0.0f);
        int k1 = (int)fminf(floorf(k1f) + 2.0f, (float)(nd - 1));
        for (int k = k0; k <= k1; ++k) {
            const int ix = xd ? k : io, iy = xd ? io : k;
            const SFFoot f = sf_foot<MODE>(G, V, ix, iy);
            const float ou = ovl(f.u0, f.u1, ul, ur);
            if (!(ou > 0.0f)) continue;
            float wu = (ou * inv_du) * f.amp;
            const float* col = vol + ((int64_t)ix * G.ny + iy) * nz;
            if (MODE == SF_FAN2 && G.compat && col[0] != 0.0f) {
                // the last nonzero pixel before iy in column ix
                int j = iy - 1;
                while (j >= 0 && vol[(int64_t)ix * G.ny + j] == 0.0f) --j;
                if (j + 1 != iy) wu = (ou * inv_du) * sf_fan_amp(G, V, ix, iy, j + 1);
            }
            if (!is3) {
                acc += (double)(col[0] * wu);
            } else if (MODE == SF_PAR3 && !slab_general) {
                float sv_ = 0.0f;
#pragma unroll
                for (int q = 0; q < kSlabs; ++q)
                    if (SW.w[q] > 0.0f) sv_ += col[SW.first + q] * SW.w[q];
                acc += (double)(wu * sv_);
            } else {
                // slabs whose magnified z interval meets [vl, vr]
                float z0f = fminf(vl, vr) / f.cm * inv_d + 0.5f * (float)G.nz;
                float z1f = fmaxf(vl, vr) / f.cm * inv_d + 0.5f * (float)G.nz;
                int z0 = (int)fmaxf(floorf(z0f) - 1.0f, 0.0f);
                int z1 = (int)fminf(floorf(z1f) + 1.0f, (float)(G.nz - 1));
                for (int z = z0; z <= z1; ++z) {
                    float ov = ovl(f.cm * sf_bnd(G.nz, G.d, z), f.cm * sf_bnd(G.nz, G.d, z + 1), vl, vr);
                    if (ov > 0.0f) acc += (double)(col[z] * (wu * (ov * inv_dv)));
                }
            }
        }
    }
    float res = (float)acc * G.out_scale;
    if (accumulate) res += sino[t];
    sino[t] = res;
}

// ------------------------------------------ tiled parallel 3-D forward
//
// One workgroup computes a tile of kTUB bins x kTV detector rows of kVG
// neighbouring views of one batch item, one view per wave; lane m of every
// wave owns detector row m_lo + m.  For every orthogonal line, over the
// pixels whose footprints can meet the tile in any of the views (the union
// of the per-bin ranges of k_sf_fp):
//   P[k][m] = sum over the row's slabs q of w_q(m) * vox[k][first(m) + q]
// read straight from the volume (lane m reads z = first(m) + q: consecutive
// lanes, consecutive z) into LDS -- P depends on the line and the pixel
// only, so the views share it -- while each wave lists, for its view's
// bins, the pixels whose footprints meet the bin and their weights; then
//   acc[u][m] += wu(k, u) * P[k][m].
// Views whose driving axis differs from the first view's are done in a
// second pass over the lines.  Weights, pixel order and arithmetic are those
// of k_sf_fp, so both give the same bits.
constexpr int kTUB = 16;         // detector bins per tile and view
constexpr int kVG = 4;           // views per workgroup (one per wave)
constexpr int kTU = kTUB * kVG;  // footprint lists per workgroup
constexpr int kTV = 64;          // detector rows per tile (one per lane)
constexpr int kKMax = 80;        // pixels of one line handled at once (LDS: 5-6 workgroups per CU)
constexpr int kWMax = 12;        // pixels whose footprint meets one bin (checked on the host)
// WPER candidate pixels per lane, 4 lanes per bin: 4 WPER <= kWMax pixels per
// bin and line (the host picks the smallest WPER its bound allows)
constexpr int kZW = 66;          // slabs a tile's 64 rows can meet (rows no taller than a voxel: 65)

template <int WPER>
__global__ __launch_bounds__(64 * kVG) __attribute__((amdgpu_waves_per_eu(6))) void k_sf_fp_par3_tile(SFGeo G, const float* __restrict__ ang,
                                                        const float* __restrict__ cs,
                                                        const float* __restrict__ sn,
                                                        const float* __restrict__ img, float* __restrict__ sino,
                                                        int accumulate, int nut, int nmt, int nvg) {
    // row kk of the chunk: the voxels vox[k][zb + z] of pixel k = kc + kk,
    // then, in place, its slab sums P[k][m] for the tile's rows m
    __shared__ float R[kKMax * kZW];
    // footprint list per bin: (LDS offset (k - kc) * kTV of the j-th pixel,
    // its weight overlap/du * amplitude as bits), read as one 64-bit word
    constexpr int kWSlot = 4 * WPER;  // footprint list stride per bin
    __shared__ int2 wlist[kTU * kWSlot];
    __shared__ int wcnt[kTU];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: scalar
    int blk = blockIdx.x;
    const int mt = blk % nmt;
    blk /= nmt;
    const int ut = blk % nut;
    blk /= nut;
    const int vg = blk % nvg;
    const int b = blk / nvg;
    const int u_lo = ut * kTUB, u_n = min(kTUB, G.nu - u_lo);
    const int m_lo = mt * kTV, m_n = min(kTV, G.nv - m_lo);
    const int a0 = vg * kVG, nva = min(kVG, G.na - a0);

    // this wave's view
    const int ia = a0 + wave;
    const bool vwave = wave < nva;
    const int iav = vwave ? ia : a0;
    const SFView V = sf_view(G, cs[iav], sn[iav], ang[iav]);
    const float c = V.c, s = V.s;
    const bool xd = fabsf(s) >= fabsf(c);
    const float inv_d = 1.0f / G.d, inv_du = 1.0f / G.du;
    const float* vol = img + (int64_t)b * G.nx * G.ny * G.nz;

    // this lane's detector row: its slabs with non-zero weight, in slab
    // order (the k_sf_fp sum skips the zero weights)
    const bool mrow = lane < m_n;
    SlabWeights SW;
    par3_row_slabs(G, m_lo + (mrow ? lane : 0), SW);
    int nzs = 0, sz[kSlabs];
    float swt[kSlabs];
#pragma unroll
    for (int q = 0; q < kSlabs; ++q) {
        sz[q] = 0;
        swt[q] = 0.0f;
    }
#pragma unroll
    for (int q = 0; q < kSlabs; ++q) {
        if (SW.w[q] > 0.0f) {
#pragma unroll
            for (int r = 0; r < kSlabs; ++r) {  // slot nzs, without dynamic register indexing
                if (r == nzs) {
                    sz[r] = SW.first + q;
                    swt[r] = SW.w[q];
                }
            }
            ++nzs;
        }
    }
    if (!mrow) nzs = 0;
    // slabs any lane of the wave needs (wave-uniform): the staging loads
    // are bounded by it, not masked per lane
    int nzs_max = nzs;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) nzs_max = max(nzs_max, __shfl_xor(nzs_max, m));
    const int nzs_u = __builtin_amdgcn_readfirstlane(nzs_max);
    // the slabs of all the tile's rows: zb .. zb + span - 1 (the same in every wave)
    int zlo = 0x7fffffff, zhi = -1;
#pragma unroll
    for (int q = 0; q < kSlabs; ++q)
        if (q < nzs) {
            zlo = min(zlo, sz[q]);
            zhi = max(zhi, sz[q]);
        }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        zlo = min(zlo, __shfl_xor(zlo, m));
        zhi = max(zhi, __shfl_xor(zhi, m));
    }
    const int zb = __builtin_amdgcn_readfirstlane(zhi >= 0 ? zlo : 0);
    const int span = min(__builtin_amdgcn_readfirstlane(zhi >= 0 ? zhi - zlo + 1 : 0), kZW);
    // this lane's slabs as offsets into an LDS row
    int zo[kSlabs];
#pragma unroll
    for (int q = 0; q < kSlabs; ++q) zo[q] = min(max(sz[q] - zb, 0), kZW - 1);

    // footprint-list lanes: bin u_lo + (lane & 15) of this wave's view (list
    // slot tu_w), candidate group grp
    const int tu_w = wave * kTUB + (lane & 15), grp = lane >> 4;
    const int iu_w = u_lo + (lane & 15);
    const bool ubin = vwave && (lane & 15) < u_n;
    const float ul = sf_det(G.nu, G.du, G.su, iu_w), ur = sf_det(G.nu, G.du, G.su, iu_w + 1);

    // the tile's bin edges (footprint margin included)
    const float ut0 = sf_det(G.nu, G.du, G.su, u_lo) - G.wmax;
    const float ut1 = sf_det(G.nu, G.du, G.su, u_lo + u_n) + G.wmax;
    // sf_invert<SF_PAR3>(u, oc) = (u - po oc) / pd as u/pd - oc po/pd: one
    // FMA per line.  The inverse only bounds pixel ranges that are widened
    // by a pixel or more (|pd| >= 1/sqrt 2, rounding moves it by ~1e-4
    // pixel), and pixels of zero overlap add nothing, so the sums are those
    // of k_sf_fp.
    const float pd = xd ? -s : c, po = xd ? c : -s;
    const float inv_pd = 1.0f / pd, lin_b = po * inv_pd;
    const float al = (ul - G.wmax) * inv_pd, ar = (ur + G.wmax) * inv_pd;
    // every view of the group (wave-uniform): driving axis and the tile's
    // band as at0 - oc lin_b .. at1 - oc lin_b
    bool gxd[kVG];
    float glin[kVG], gat0[kVG], gat1[kVG];
    bool mixed = false;
#pragma unroll
    for (int w = 0; w < kVG; ++w) {
        const int aw = a0 + min(w, nva - 1);
        const float cw = cs[aw], sw = sn[aw];
        gxd[w] = fabsf(sw) >= fabsf(cw);
        const float pdw = gxd[w] ? -sw : cw, pow_ = gxd[w] ? cw : -sw;
        const float ipd = 1.0f / pdw;
        glin[w] = pow_ * ipd;
        gat0[w] = ut0 * ipd;
        gat1[w] = ut1 * ipd;
        mixed |= w < nva && gxd[w] != gxd[0];
    }

    double acc[kTUB];
#pragma unroll
    for (int i = 0; i < kTUB; ++i) acc[i] = 0.0;

    // views whose driving axis is that of the first view, then the others
    for (int pass = 0; pass < (mixed ? 2 : 1); ++pass) {
        const bool pxd = pass == 0 ? gxd[0] : !gxd[0];
        const bool active = vwave && xd == pxd;
        const int nd = pxd ? G.nx : G.ny, no = pxd ? G.ny : G.nx;
        for (int io = 0; io < no; ++io) {
            const float oc = (sf_bnd(no, G.d, io) + sf_bnd(no, G.d, io + 1)) * 0.5f;
            // the union of the pass's views' pixel bands on this line
            int k0t = nd, k1t = -1;
#pragma unroll
            for (int w = 0; w < kVG; ++w) {
                if (w >= nva || gxd[w] != pxd) continue;
                const float x0 = __builtin_fmaf(-oc, glin[w], gat0[w]);
                const float x1 = __builtin_fmaf(-oc, glin[w], gat1[w]);
                const float k0f = fminf(x0, x1) * inv_d + 0.5f * (float)nd - 0.5f;
                const float k1f = fmaxf(x0, x1) * inv_d + 0.5f * (float)nd - 0.5f;
                if (!(k0f < (float)nd + 1.0f) || !(k1f > -2.0f)) continue;
                k0t = min(k0t, (int)fmaxf(floorf(k0f) - 1.0f, 0.0f));
                k1t = max(k1t, (int)fminf(floorf(k1f) + 2.0f, (float)(nd - 1)));
            }
            if (k0t > k1t) continue;
            // this lane's bin: its candidate pixel range (as k_sf_fp)
            int kb0 = 0, kb1 = -1;
            if (ubin && active) {
                const float y0 = __builtin_fmaf(-oc, lin_b, al);
                const float y1 = __builtin_fmaf(-oc, lin_b, ar);
                const float b0f = fminf(y0, y1) * inv_d + 0.5f * (float)nd - 0.5f;
                const float b1f = fmaxf(y0, y1) * inv_d + 0.5f * (float)nd - 0.5f;
                if (b0f < (float)nd + 1.0f && b1f > -2.0f) {
                    kb0 = (int)fmaxf(floorf(b0f) - 1.0f, 0.0f);
                    kb1 = (int)fminf(floorf(b1f) + 2.0f, (float)(nd - 1));
                }
            }
            for (int kc = k0t; kc <= k1t; kc += kKMax) {
                const int nk = min(kKMax, k1t - kc + 1);
                __syncthreads();  // the previous chunk's reads are done
                // the chunk's voxel rows, straight from the volume into LDS
                // (consecutive lanes, consecutive z; no registers held): all of
                // a wave's rows are in flight together, while the footprint
                // lists below are built
                for (int kk = wave; kk < nk; kk += kVG) {
                    const int k = kc + kk;
                    const int ix = pxd ? k : io, iy = pxd ? io : k;
                    const float* col = vol + ((int64_t)ix * G.ny + iy) * G.nz + zb;
                    float* row = R + kk * kZW;
                    if (lane < span)
                        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(col + lane),
                                                         (__attribute__((address_space(3))) void*)row, 4, 0, 0);
                    if (lane < span - 64)
                        __builtin_amdgcn_global_load_lds(
                            (const __attribute__((address_space(1))) void*)(col + 64 + lane),
                            (__attribute__((address_space(3))) void*)(row + 64), 4, 0, 0);
                }
                // footprint lists of this wave's bins over the chunk, in pixel
                // order: lane group grp tests candidates ka + WPER grp .. ka + WPER grp + WPER - 1
                {
                    const int ka = max(kb0, kc), kz = min(kb1, kc + nk - 1);
                    float wv[WPER];
                    int nzc = 0;
#pragma unroll
                    for (int j = 0; j < WPER; ++j) {
                        const int k = ka + grp * WPER + j;
                        float w = 0.0f;
                        if (k <= kz) {
                            const int ix = xd ? k : io, iy = xd ? io : k;
                            const SFFoot f = sf_foot<SF_PAR3>(G, V, ix, iy);
                            const float ou = ovl(f.u0, f.u1, ul, ur);
                            if (ou > 0.0f) w = (ou * inv_du) * f.amp;
                        }
                        wv[j] = w;
                        nzc += (k <= kz && w != 0.0f) ? 1 : 0;
                    }
                    // exclusive prefix of the group counts of this bin (lanes l, l+16, l+32, l+48)
                    const int c1 = __shfl_up(nzc, 16), c2 = __shfl_up(nzc, 32), c3 = __shfl_up(nzc, 48);
                    int pos = (grp >= 1 ? c1 : 0) + (grp >= 2 ? c2 : 0) + (grp >= 3 ? c3 : 0);
                    const int total = __shfl(pos + nzc, (lane & 15) + 48);
#pragma unroll
                    for (int j = 0; j < WPER; ++j) {
                        const int k = ka + grp * WPER + j;
                        if (k <= kz && wv[j] != 0.0f) {
                            wlist[tu_w * kWSlot + pos] = make_int2((k - kc) * kZW, __float_as_int(wv[j]));
                            ++pos;
                        }
                    }
                    if (grp == 0) wcnt[tu_w] = ubin ? total : 0;
                }
                __syncthreads();  // the voxel rows have landed
                // slab sums of this wave's rows, in place: each row is read
                // and rewritten by one wave, reads before writes
                for (int kk = wave; kk < nk; kk += kVG) {
                    float x[kSlabs];
#pragma unroll
                    for (int q = 0; q < kSlabs; ++q) x[q] = q < nzs_u ? R[kk * kZW + zo[q]] : 0.0f;
                    float sv_ = 0.0f;
#pragma unroll
                    for (int q = 0; q < kSlabs; ++q)
                        if (q < nzs) sv_ += x[q] * swt[q];
                    R[kk * kZW + lane] = sv_;
                }
                __syncthreads();
                if (mrow && active) {
#pragma unroll
                    for (int i = 0; i < kTUB; ++i) {
                        const int tu = wave * kTUB + i;
                        const int cnt = wcnt[tu];  // wave-uniform
                        const int2* wl = wlist + tu * kWSlot;
                        int j = 0;
                        // offsets (k - kc) * kZW with 0 <= k - kc < nk <= kKMax
                        for (; j + 1 < cnt; j += 2) {
                            const int2 e0 = wl[j], e1 = wl[j + 1];
                            const float p0 = R[e0.x + lane];
                            const float p1 = R[e1.x + lane];
                            acc[i] += (double)(__int_as_float(e0.y) * p0);
                            acc[i] += (double)(__int_as_float(e1.y) * p1);
                        }
                        if (j < cnt) {
                            const int2 e0 = wl[j];
                            acc[i] += (double)(__int_as_float(e0.y) * R[e0.x + lane]);
                        }
                    }
                }
            }
        }
    }
    if (!mrow || !vwave) return;
    const int64_t base = (((int64_t)b * G.na + ia) * G.nu) * G.nv + m_lo + lane;
#pragma unroll
    for (int i = 0; i < kTUB; ++i) {
        if (i >= u_n) continue;
        const int64_t o = base + (int64_t)(u_lo + i) * G.nv;
        float res = (float)acc[i] * G.out_scale;
        if (accumulate) res += sino[o];
        sino[o] = res;
    }
}

// ------------------------------------------ parallel 3-D backward by columns
//
// One wave per (batch item, pixel, 64 slabs): lane z accumulates voxel z of
// the column over all views.  The footprint of the pixel -- bins with
// positive overlap and their weights -- is computed for 64 views at once
// (lane = view) and broadcast with readlane in the view loop, which then only
// reads the sinogram rows each slab meets (consecutive lanes, consecutive
// rows) and accumulates.  Weights, bin and view order and arithmetic are
// those of k_sf_bp, so both give the same bits.  The kernel waits on memory:
// the rows of a view's first kJPre bins are loaded together, and 64 VGPRs
// keep eight waves per SIMD.
constexpr int kJMax = 8;  // bins a pixel footprint meets (checked on the host)
constexpr int kJPre = 2;  // bins whose sinogram rows are loaded together

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_sf_bp_par3_col(SFGeo G, const float* __restrict__ ang,
                                                       const float* __restrict__ cs,
                                                       const float* __restrict__ sn,
                                                       const float* __restrict__ sino, float* __restrict__ img,
                                                       int accumulate, int nzt) {
    const int lane = threadIdx.x & 63;
    const int64_t wid = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wid >= (int64_t)G.nb * G.nx * G.ny * nzt) return;  // whole wave
    const int zt = (int)(wid % nzt);
    int64_t r = wid / nzt;
    const int iy = (int)(r % G.ny);
    r /= G.ny;
    const int ix = (int)(r % G.nx);
    const int b = (int)(r / G.nx);
    const int iz = zt * 64 + lane;
    const bool zin = iz < G.nz;

    // this lane's slab: rows with non-zero weight, in row order (k_sf_bp skips
    // the zero weights, and the whole view when the slab meets no row)
    SlabWeights SW;
    par3_slab_rows(G, zin ? iz : 0, SW);
    int nzr = 0, rq[kSlabs];
    float rw[kSlabs];
#pragma unroll
    for (int q = 0; q < kSlabs; ++q) {
        rq[q] = 0;
        rw[q] = 0.0f;
    }
#pragma unroll
    for (int q = 0; q < kSlabs; ++q) {
        if (SW.w[q] > 0.0f) {
#pragma unroll
            for (int s = 0; s < kSlabs; ++s) {  // slot nzr, without dynamic register indexing
                if (s == nzr) {
                    rq[s] = SW.first + q;
                    rw[s] = SW.w[q];
                }
            }
            ++nzr;
        }
    }
    if (!zin || SW.count <= 0) nzr = 0;

    // rows any lane of the wave needs per bin (wave-uniform)
    int nzr_max = nzr;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) nzr_max = max(nzr_max, __shfl_xor(nzr_max, m));
    const int nzr_u = __builtin_amdgcn_readfirstlane(nzr_max);
    const bool pre = nzr_u <= 2;

    const float u0 = sf_det(G.nu, G.du, G.su, 0);
    const float inv_du = 1.0f / G.du;
    const float* sb = sino + (int64_t)b * G.na * G.nu * G.nv;
    double acc = 0.0;
    // the view loop, for NR preloaded rows per bin: one where no slab of the
    // wave meets a second row (rows aligned with the slabs), else two
    auto views = [&](auto nr_tag) {
        constexpr int NR = decltype(nr_tag)::value;
        for (int va = 0; va < G.na; va += 64) {
            // lane = view va + lane: bins of this pixel's footprint with positive overlap
            int cnt = 0, jl[kJMax];
            float wl[kJMax];
#pragma unroll
            for (int s = 0; s < kJMax; ++s) {
                jl[s] = 0;
                wl[s] = 0.0f;
            }
            const int ial = va + lane;
            if (ial < G.na) {
                const SFView V = sf_view(G, cs[ial], sn[ial], ang[ial]);
                const SFFoot f = sf_foot<SF_PAR3>(G, V, ix, iy);
                if (f.u1 > u0 - G.du && (f.u0 - u0) * inv_du < (float)G.nu + 1.0f) {
                    const int j0 = (int)fmaxf(floorf((f.u0 - u0) * inv_du) - 1.0f, 0.0f);
                    const int j1 = (int)fminf(floorf((f.u1 - u0) * inv_du) + 1.0f, (float)(G.nu - 1));
#pragma unroll
                    for (int t = 0; t < kJMax; ++t) {
                        const int j = j0 + t;
                        if (j > j1) break;
                        const float ou = ovl(f.u0, f.u1, sf_det(G.nu, G.du, G.su, j), sf_det(G.nu, G.du, G.su, j + 1));
                        if (!(ou > 0.0f)) continue;
                        const float wu = (ou * inv_du) * f.amp;
#pragma unroll
                        for (int s = 0; s < kJMax; ++s) {
                            if (s == cnt) {
                                jl[s] = j;
                                wl[s] = wu;
                            }
                        }
                        ++cnt;
                    }
                }
            }
            const int nvw = min(64, G.na - va);
            for (int v = 0; v < nvw; ++v) {
                const int c = __builtin_amdgcn_readlane(cnt, v);
                const float* plane = sb + (int64_t)(va + v) * G.nu * G.nv;
                // the rows of the first kJPre bins, two per slab, are loaded
                // together before any is used, by every lane and for every one
                // of these bins: bins past the footprint's count read bin jl = 0,
                // rows past the slab's count row rq = 0, both valid, and the sums
                // below skip them
                float xv[kJPre][2];
                if (pre) {
#pragma unroll
                    for (int s = 0; s < kJPre; ++s) {
                        const float* row = plane + (int64_t)__builtin_amdgcn_readlane(jl[s], v) * G.nv;
                        xv[s][0] = row[rq[0]];
                        xv[s][1] = NR > 1 ? row[rq[1]] : 0.0f;
                    }
                }
#pragma unroll
                for (int s = 0; s < kJMax; ++s) {
                    if (s >= c) break;
                    const float wu = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wl[s]), v));
                    float sv_ = 0.0f;
                    if (s < kJPre && pre) {
                        const int sp = s < kJPre ? s : 0;
                        sv_ = 0 < nzr ? sv_ + xv[sp][0] * rw[0] : sv_;
                        if (NR > 1) sv_ = 1 < nzr ? sv_ + xv[sp][1] * rw[1] : sv_;
                    } else {
                        const int j = __builtin_amdgcn_readlane(jl[s], v);
                        const float* row = plane + (int64_t)j * G.nv;
#pragma unroll
                        for (int q = 0; q < kSlabs; ++q)
                            if (q < nzr) sv_ += row[rq[q]] * rw[q];
                    }
                    acc += (double)(wu * sv_);
                }
            }
        }
    };
    if (nzr_u <= 1)
        views(std::integral_constant<int, 1>{});
    else
        views(std::integral_constant<int, 2>{});
    if (!zin) return;
    const int64_t o = (((int64_t)b * G.nx + ix) * G.ny + iy) * G.nz + iz;
    float res = (float)acc * G.out_scale;
    if (accumulate) res += img[o];
    img[o] = res;
}

// True when the column backprojector applies: a slab meets at most kSlabs
// detector rows (rows at least half a voxel tall) and a pixel footprint at
// most kJMax bins.
static bool sf_par3_bp_col_ok(const SFGeo& G) {
    if (!((double)G.d / (double)G.dv <= 1.99)) return false;
    return 1.4143 * (double)G.d / (double)G.du + 4.0 <= (double)kJMax;
}

// Bound on the candidate pixels of one bin along a line: a bin's pixel range
// is floor(b1) - floor(b0) + 4 long, b1 - b0 = (du + 2 wmax) / (|pd| d) with
// |pd| >= 1/sqrt 2, so it is at most floor(X) + 5 for X that ratio's bound;
// per_bin = X + 5 keeps a margin for rounding.
static double sf_par3_per_bin(const SFGeo& G) {
    return ((double)G.du + 2.0 * (double)G.wmax) / (0.70710678 * (double)G.d) + 5.0;
}

// True when the tiled parallel 3-D forward kernel applies: detector rows no
// taller than a voxel (each row meets at most kSlabs slabs) and at most
// kWMax - 1 pixels per bin along a line.
static bool sf_par3_tiled_ok(const SFGeo& G) {
    if (!(G.dv <= G.d)) return false;
    return sf_par3_per_bin(G) <= (double)(kWMax - 1);
}

template <int MODE>
__global__ __launch_bounds__(256) void k_sf_bp(SFGeo G, const float* __restrict__ ang,
                                              const float* __restrict__ cs, const float* __restrict__ sn,
                                              const float* __restrict__ sino, float* __restrict__ img,
                                              int accumulate) {
    const bool is3 = MODE == SF_CONE3 || MODE == SF_PAR3;
    const int nz = is3 ? G.nz : 1, nv = is3 ? G.nv : 1;
    const int64_t per_b = (int64_t)G.nx * G.ny * nz;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= per_b * G.nb) return;
    const int b = (int)(t / per_b);
    int64_t r = t % per_b;
    const int iz = (int)(r % nz);
    const int iy = (int)((r / nz) % G.ny);
    const int ix = (int)(r / ((int64_t)nz * G.ny));
    const float u0 = sf_det(G.nu, G.du, G.su, 0), v0 = sf_det(nv, G.dv, G.sv, 0);
    const float inv_du = 1.0f / G.du, inv_dv = 1.0f / G.dv;
    const float zl = sf_bnd(G.nz, G.d, iz), zr = sf_bnd(G.nz, G.d, iz + 1);
    const float* sb = sino + (int64_t)b * G.na * G.nu * nv;
    // parallel 3-D: the detector rows this slab meets and their weights
    SlabWeights SW;
    SW.first = 0;
    SW.count = 0;
    bool row_general = false;
    if (MODE == SF_PAR3) row_general = !par3_slab_rows(G, iz, SW);
    double acc = 0.0;
    for (int ia = 0; ia < G.na; ++ia) {
        const SFView V = sf_view(G, cs[ia], sn[ia], ang[ia]);
        const SFFoot f = sf_foot<MODE>(G, V, ix, iy);
        if (!(f.u1 > u0 - G.du) || !((f.u0 - u0) * inv_du < (float)G.nu + 1.0f)) continue;
        int j0 = (int)fmaxf(floorf((f.u0 - u0) * inv_du) - 1.0f, 0.0f);
        int j1 = (int)fminf(floorf((f.u1 - u0) * inv_du) + 1.0f, (float)(G.nu - 1));
        const float* plane = sb + (int64_t)ia * G.nu * nv;
        if (!is3) {
            for (int j = j0; j <= j1; ++j) {
                const float ou = ovl(f.u0, f.u1, sf_det(G.nu, G.du, G.su, j), sf_det(G.nu, G.du, G.su, j + 1));
                if (ou > 0.0f) acc += (double)(plane[j] * ((ou * inv_du) * f.amp));
            }
            continue;
        }
        if (MODE == SF_PAR3 && !row_general) {
            if (SW.count <= 0) continue;
            for (int j = j0; j <= j1; ++j) {
                const float ou = ovl(f.u0, f.u1, sf_det(G.nu, G.du, G.su, j), sf_det(G.nu, G.du, G.su, j + 1));
                if (!(ou > 0.0f)) continue;
                const float wu = (ou * inv_du) * f.amp;
                const float* rowp = plane + (int64_t)j * nv + SW.first;
                float sv_ = 0.0f;
#pragma unroll
                for (int q = 0; q < kSlabs; ++q)
                    if (SW.w[q] > 0.0f) sv_ += rowp[q] * SW.w[q];
                acc += (double)(wu * sv_);
            }
            continue;
        }
        float qa = f.cm * zl, qb = f.cm * zr;
        float lo = fminf(qa, qb), hi = fmaxf(qa, qb);
        if (!(hi > v0 - G.dv) || !((lo - v0) * inv_dv < (float)nv + 1.0f)) continue;
        int m0 = (int)fmaxf(floorf((lo - v0) * inv_dv) - 1.0f, 0.0f);
        int m1 = (int)fminf(floorf((hi - v0) * inv_dv) + 1.0f, (float)(nv - 1));
        for (int j = j0; j <= j1; ++j) {
            const float ou = ovl(f.u0, f.u1, sf_det(G.nu, G.du, G.su, j), sf_det(G.nu, G.du, G.su, j + 1));
            if (!(ou > 0.0f)) continue;
            const float wu = (ou * inv_du) * f.amp;
            for (int m = m0; m <= m1; ++m) {
                float ov = ovl(qa, qb, sf_det(nv, G.dv, G.sv, m), sf_det(nv, G.dv, G.sv, m + 1));
                if (ov > 0.0f) acc += (double)(plane[(int64_t)j * nv + m] * (wu * (ov * inv_dv)));
            }
        }
    }
    float res = (float)acc * G.out_scale;
    if (accumulate) res += img[t];
    img[t] = res;
}

static int sf_setup(const vir_sf_geom* g, SFGeo& G) {
    if (!g) return fail_arg("geometry is NULL");
    if (g->mode < 0 || g->mode > 3) return fail_arg("mode must be 0..3");
    bool is3 = g->mode == SF_CONE3 || g->mode == SF_PAR3;
    G.nx = g->nx; G.ny = g->ny; G.nz = is3 ? g->nz : 1;
    G.nu = g->nu; G.nv = is3 ? g->nv : 1; G.na = g->na; G.nb = g->nbatch;
    G.d = g->d_pix; G.du = g->du; G.dv = is3 ? g->dv : 1.0f;
    G.su = g->su; G.sv = is3 ? g->sv : 0.0f; G.DSO = g->DSO; G.DSD = g->DSD;
    G.out_scale = 1.0f;
    G.unit_amp = 0;
    G.compat = 0;
    if (G.nx < 1 || G.ny < 1 || G.nz < 1 || G.nu < 1 || G.nv < 1 || G.na < 0 || G.nb < 1)
        return fail_arg("bad sizes");
    if (!(G.d > 0.0f) || !(G.du > 0.0f) || !(G.dv > 0.0f)) return fail_arg("spacings must be > 0");
    // Largest distance between a pixel's centre projection and any point of
    // its footprint (the footprint contains the centre projection):
    //   parallel: half the projected diagonal, d/sqrt(2);
    //   fan/cone: a corner offset (dp, do) from the centre moves the
    //   projection by DSD*(dp + t*do)/(DSO - o_corner) with t = u_c/DSD, so
    //   the distance is at most mag * (d/sqrt(2)) * sqrt(1 + t^2), where
    //   mag = DSD/(DSO - R) and |t| <= R/sqrt(DSO^2 - R^2) over the disc of
    //   radius R (the grid's half-diagonal) that holds every pixel.
    const double hx = 0.5 * G.nx * (double)G.d, hy = 0.5 * G.ny * (double)G.d;
    const double R = std::sqrt(hx * hx + hy * hy);
    double w = (double)G.d * 0.70710678118654752;
    if (g->mode == SF_FAN2 || g->mode == SF_CONE3) {
        if (!(G.DSO > 0.0f) || !(G.DSD > 0.0f)) return fail_arg("DSO and DSD must be > 0");
        if (!(R < 0.999 * G.DSO)) return fail_arg("grid reaches the source (DSO too small)");
        const double D = G.DSO;
        w *= (double)G.DSD / (D - R) * (D / std::sqrt(D * D - R * R));
    }
    G.wmax = (float)(1.02 * w + 1e-3 * (double)G.du);
    if (g->mode == SF_FAN2 || g->mode == SF_CONE3) {
        // the forward gather inverts the centre projection along the driving
        // axis; that inverse is monotone while |u| < DSD (fan half-angle
        // below 45 degrees), so the detector plus margin must stay inside
        const double umax = (double)G.du * std::fmax(std::fabs(-0.5 * G.nu + G.su),
                                                     std::fabs(0.5 * G.nu + G.su)) + G.wmax;
        if (!(umax < 0.999 * (double)G.DSD))
            return fail_arg("detector reaches a 45 degree fan half-angle (|u| >= DSD)");
    }
    return VIR_OK;
}

}  // namespace vir

using namespace vir;

extern "C" {

int vir_sf_fp(const vir_sf_geom* geom, const float* ang, const float* cosv, const float* sinv, const float* img,
              float* sino, float out_scale, uint32_t flags, void* stream) {
    SFGeo G;
    int rc = sf_setup(geom, G);
    if (rc) return rc;
    if (G.na == 0) return VIR_OK;
    if (!ang || !cosv || !sinv || !img || !sino) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    G.unit_amp = (flags & VIR_SF_UNIT_AMPLITUDE) ? 1 : 0;
    G.compat = (flags & VIR_COMPAT) ? 1 : 0;
    int64_t n = (int64_t)G.nb * G.na * G.nu * G.nv;
    dim3 grd((unsigned)((n + 255) / 256));
    hipStream_t st = as_stream(stream);
    int acc = (flags & VIR_ACCUMULATE) ? 1 : 0;
    switch (geom->mode) {
        case SF_PAR2: hipLaunchKernelGGL(k_sf_fp<SF_PAR2>, grd, dim3(256), 0, st, G, ang, cosv, sinv, img, sino, acc); break;
        case SF_FAN2: hipLaunchKernelGGL(k_sf_fp<SF_FAN2>, grd, dim3(256), 0, st, G, ang, cosv, sinv, img, sino, acc); break;
        case SF_CONE3: hipLaunchKernelGGL(k_sf_fp<SF_CONE3>, grd, dim3(256), 0, st, G, ang, cosv, sinv, img, sino, acc); break;
        default:
            if (sf_par3_tiled_ok(G) && !(flags & VIR_PER_OUTPUT)) {
                const int nut = (G.nu + kTUB - 1) / kTUB, nmt = (G.nv + kTV - 1) / kTV;
                const int nvg = (G.na + kVG - 1) / kVG;
                const int64_t nblk = (int64_t)G.nb * nvg * nut * nmt;
                // two candidates per lane when a bin's range is at most 8 pixels
                // (X <= 3.9 < 4), else three (X <= 6: at most 11 of 12)
                if (sf_par3_per_bin(G) <= 8.9)
                    hipLaunchKernelGGL(k_sf_fp_par3_tile<2>, dim3((unsigned)nblk), dim3(64 * kVG), 0, st, G, ang, cosv,
                                       sinv, img, sino, acc, nut, nmt, nvg);
                else
                    hipLaunchKernelGGL(k_sf_fp_par3_tile<3>, dim3((unsigned)nblk), dim3(64 * kVG), 0, st, G, ang, cosv,
                                       sinv, img, sino, acc, nut, nmt, nvg);
            } else {
                hipLaunchKernelGGL(k_sf_fp<SF_PAR3>, grd, dim3(256), 0, st, G, ang, cosv, sinv, img, sino, acc);
            }
            break;
    }
    return check_hip(hipGetLastError(), "vir_sf_fp launch");
}

int vir_sf_bp(const vir_sf_geom* geom, const float* ang, const float* cosv, const float* sinv, const float* sino,
              float* img, float out_scale, uint32_t flags, void* stream) {
    SFGeo G;
    int rc = sf_setup(geom, G);
    if (rc) return rc;
    if ((G.na > 0 && (!ang || !cosv || !sinv || !sino)) || !img) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    G.unit_amp = (flags & VIR_SF_UNIT_AMPLITUDE) ? 1 : 0;
    int64_t n = (int64_t)G.nb * G.nx * G.ny * G.nz;
    dim3 grd((unsigned)((n + 255) / 256));
    hipStream_t st = as_stream(stream);
    int acc = (flags & VIR_ACCUMULATE) ? 1 : 0;
    switch (geom->mode) {
        case SF_PAR2: hipLaunchKernelGGL(k_sf_bp<SF_PAR2>, grd, dim3(256), 0, st, G, ang, cosv, sinv, sino, img, acc); break;
        case SF_FAN2: hipLaunchKernelGGL(k_sf_bp<SF_FAN2>, grd, dim3(256), 0, st, G, ang, cosv, sinv, sino, img, acc); break;
        case SF_CONE3: hipLaunchKernelGGL(k_sf_bp<SF_CONE3>, grd, dim3(256), 0, st, G, ang, cosv, sinv, sino, img, acc); break;
        default:
            if (sf_par3_bp_col_ok(G) && !(flags & VIR_PER_OUTPUT)) {
                const int nzt = (G.nz + 63) / 64;
                const int64_t nw = (int64_t)G.nb * G.nx * G.ny * nzt;
                hipLaunchKernelGGL(k_sf_bp_par3_col, dim3((unsigned)((nw + 3) / 4)), dim3(256), 0, st, G, ang, cosv,
                                   sinv, sino, img, acc, nzt);
            } else {
                hipLaunchKernelGGL(k_sf_bp<SF_PAR3>, grd, dim3(256), 0, st, G, ang, cosv, sinv, sino, img, acc);
            }
            break;
    }
    return check_hip(hipGetLastError(), "vir_sf_bp launch");
}

}  // extern "C"
