// Distance-driven projectors for gfx950: 2-D parallel, 2-D fan (flat or
// equiangular detector) and 3-D cone, forward and exact-adjoint backward.
//
// Model (sys_mat/dd.py): per view the driving axis is x when |sin| >= |cos|
// (else y); each pixel row along the driving axis, taken at its orthogonal
// centre, is projected onto the detector boundary by boundary; a pixel
// contributes img * overlap(projected pixel interval, bin interval) * scale,
// scale = 1/|sin| (parallel) or r / ((DSO - o) |sin|) (fan/cone, r the
// in-plane source distance of the row centre).  The cone variant multiplies
// by the overlap of the slab's magnified z interval with the detector row.
//
// Both directions are gathers (no atomics): the forward kernel runs one
// thread per detector bin and finds the pixels whose shadows meet it by
// inverting the projection; the backward kernel runs one thread per pixel
// (voxel) and visits the bins its shadow meets.  Both evaluate every weight
// through the same functions, so they are exact transposes.
#include <algorithm>
#include <cmath>

#include "vir_common.h"

namespace vir {

enum DDMode { DD_PAR = 0, DD_FAN_FLAT = 1, DD_FAN_ARC = 2 };

struct DDGeo {
    int nx, ny, nz;
    int nu, nv, na;
    float d;            // voxel size
    float du, dv;       // detector pitch (arc length for the equiangular fan)
    float su, sv;       // detector offsets in bins
    float DSO, DSD;
    float out_scale;    // applied to every output element
};

struct DDView {
    bool xd;            // x is the driving axis
    float pd, po;       // p = pd * drive + po * orth
    float od, oo;       // o = od * drive + oo * orth (distance toward the source)
    float rs;           // |sin| or |cos| of the driving direction
};

__device__ __forceinline__ DDView dd_view(float c, float s) {
    DDView v;
    v.xd = fabsf(s) >= fabsf(c);
    if (v.xd) { v.pd = -s; v.po = c; v.od = c; v.oo = s; v.rs = fabsf(s); }
    else { v.pd = c; v.po = -s; v.od = s; v.oo = c; v.rs = fabsf(c); }
    return v;
}

// boundary k of n cells of width d centred at 0: d * (k - n/2) in f32
__device__ __forceinline__ float cell_bnd(int n, float d, int k) {
    return d * ((float)k - 0.5f * (float)n);
}

__device__ __forceinline__ float cell_cen(int n, float d, int k) {
    return (cell_bnd(n, d, k) + cell_bnd(n, d, k + 1)) * 0.5f;
}

__device__ __forceinline__ float det_bnd(int n, float du, float su, int j) {
    return du * (((float)j - 0.5f * (float)n) + su);
}

// projected position of the driving-axis boundary k of orthogonal line io
template <int MODE>
__device__ __forceinline__ float dd_bound(const DDGeo& G, const DDView& V, float drv_b, float orth_c) {
    float p = V.pd * drv_b + V.po * orth_c;
    if (MODE == DD_PAR) return p;
    float o = V.od * drv_b + V.oo * orth_c;
    if (MODE == DD_FAN_FLAT) return (G.DSD * p) / (G.DSO - o);
    return G.DSD * atan2f(p, G.DSO - o);
}

// ray scale of orthogonal line io
template <int MODE>
__device__ __forceinline__ float dd_scale(const DDGeo& G, const DDView& V, float orth_c) {
    if (MODE == DD_PAR) return 1.0f / V.rs;
    float po = V.po * orth_c, oo = V.oo * orth_c;
    float den = G.DSO - oo;
    return sqrtf(den * den + po * po) / (den * V.rs);
}

// detector coordinate u as the slope t = p / (DSO - o) of its ray (fan), or
// u itself (parallel): the per-bin part of dd_invert, hoisted out of loops
template <int MODE>
__device__ __forceinline__ float dd_slope(const DDGeo& G, float u) {
    if (MODE == DD_PAR) return u;
    return MODE == DD_FAN_FLAT ? u / G.DSD : tanf(u / G.DSD);
}

// inverse of dd_bound along the driving axis: the drive coordinate whose
// boundary projects to the detector coordinate of slope t (rounding only;
// callers widen the index range by one)
template <int MODE>
__device__ __forceinline__ float dd_invert(const DDGeo& G, const DDView& V, float t, float orth_c) {
    float po = V.po * orth_c, oo = V.oo * orth_c;
    if (MODE == DD_PAR) return (t - po) / V.pd;
    return (t * (G.DSO - oo) - po) / (V.pd + t * V.od);
}

// dd_invert as x = A - oc B for a fixed slope t (the tiled forward's range).
template <int MODE>
__device__ __forceinline__ void dd_invert_lin(const DDGeo& G, const DDView& V, float t, float& A, float& B) {
    if (MODE == DD_PAR) {
        const float r = 1.0f / V.pd;
        A = t * r;
        B = V.po * r;
    } else {
        const float r = 1.0f / (V.pd + t * V.od);
        A = t * G.DSO * r;
        B = (t * V.oo + V.po) * r;
    }
}

__device__ __forceinline__ float overlap(float a0, float a1, float b0, float b1) {
    float lo = a0 < a1 ? a0 : a1, hi = a0 < a1 ? a1 : a0;
    return fminf(hi, b1) - fmaxf(lo, b0);
}

// ---------------------------------------------------------------- 2-D

// One thread per (view, bin).  img is (nx, ny) C-order.
template <int MODE>
__global__ __launch_bounds__(256) void k_dd2d_fp(DDGeo G, const float* __restrict__ cs,
                                                const float* __restrict__ sn, const float* __restrict__ img,
                                                float* __restrict__ sino, int accumulate, SirtEpi E) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.na * G.nu) return;
    const int ia = (int)(t / G.nu), iu = (int)(t % G.nu);
    const DDView V = dd_view(cs[ia], sn[ia]);
    const int nd = V.xd ? G.nx : G.ny, no = V.xd ? G.ny : G.nx;
    const float ul = det_bnd(G.nu, G.du, G.su, iu), ur = det_bnd(G.nu, G.du, G.su, iu + 1);
    const float tl = dd_slope<MODE>(G, ul), tr = dd_slope<MODE>(G, ur);
    const float inv_d = 1.0f / G.d;
    double acc = 0.0;
    for (int io = 0; io < no; ++io) {
        const float oc = cell_cen(no, G.d, io);
        float x0 = dd_invert<MODE>(G, V, tl, oc), x1 = dd_invert<MODE>(G, V, tr, oc);
        float k0f = fminf(x0, x1) * inv_d + 0.5f * (float)nd;
        float k1f = fmaxf(x0, x1) * inv_d + 0.5f * (float)nd;
        int k0 = (int)fmaxf(floorf(k0f) - 1.0f, 0.0f);
        int k1 = (int)fminf(floorf(k1f) + 1.0f, (float)(nd - 1));
        if (!(k0f < (float)nd + 2.0f) || !(k1f > -2.0f)) continue;
        const float sc = dd_scale<MODE>(G, V, oc);
        float pb = dd_bound<MODE>(G, V, cell_bnd(nd, G.d, k0), oc);
        for (int k = k0; k <= k1; ++k) {
            float pn = dd_bound<MODE>(G, V, cell_bnd(nd, G.d, k + 1), oc);
            float ov = overlap(pb, pn, ul, ur);
            if (ov > 0.0f) {
                int ix = V.xd ? k : io, iy = V.xd ? io : k;
                acc += (double)(img[(int64_t)ix * G.ny + iy] * (ov * sc));
            }
            pb = pn;
        }
    }
    float r = (float)acc * G.out_scale;
    if (accumulate) r += sino[t];
    sirt_store(E, sino, t, r);
}

// Tiled forward: one workgroup per (view, 256 bins).  For every orthogonal
// line the projected boundaries of the pixels that can meet the tile's bins
// (the union of the per-bin ranges of k_dd2d_fp, widened by one) are
// evaluated once into LDS with the pixel values, and each bin merges its own
// pixels from there -- on the equiangular detector that is one atan2 per
// boundary instead of one per (bin, boundary).  The per-line ray scale and
// the tile's pixel band are computed once per workgroup, and two lines share
// each pair of barriers.  Same functions, pixel order and arithmetic as
// k_dd2d_fp: the same bits.
constexpr int kDDKMax = 256;    // pixels of one line staged at once (per line of a pair)
constexpr int kDDLines = 2048;  // lines whose scale and band are tabulated per workgroup

template <int MODE>
__global__ __launch_bounds__(256) void k_dd2d_fp_tile(DDGeo G, const float* __restrict__ cs,
                                                     const float* __restrict__ sn, const float* __restrict__ img,
                                                     float* __restrict__ sino, int accumulate, int nut, SirtEpi E) {
    __shared__ float bnd[2][kDDKMax + 1];  // projected boundary kc + e of the two lines
    __shared__ float val[2][kDDKMax];      // pixel kc + e
    // per-line tables sized by the launch to max(nx, ny) lines (<= kDDLines):
    // 8 bytes a line, so small images keep more workgroups resident
    extern __shared__ float dyn_lds[];
    float* lsc = dyn_lds;                                                        // ray scale of line io
    short2* lband = reinterpret_cast<short2*>(dyn_lds + max(G.nx, G.ny));  // the tile's pixel band of line io (empty: y < x)
    const int ut = blockIdx.x % nut, ia = blockIdx.x / nut;
    const int u_lo = ut * 256, u_n = min(256, G.nu - u_lo);
    const int iu = u_lo + (int)threadIdx.x;
    const bool ubin = (int)threadIdx.x < u_n;
    const DDView V = dd_view(cs[ia], sn[ia]);
    const int nd = V.xd ? G.nx : G.ny, no = V.xd ? G.ny : G.nx;
    const float ul = det_bnd(G.nu, G.du, G.su, iu), ur = det_bnd(G.nu, G.du, G.su, iu + 1);
    const float tl = dd_slope<MODE>(G, ul), tr = dd_slope<MODE>(G, ur);
    // dd_invert(t, oc) = A(t) - oc B(t): its denominator depends on the bin
    // only.  The inverse only bounds the pixel range, which is widened by a
    // pixel, so this rounding (~1e-5 pixel) leaves the range covering every
    // pixel of positive overlap and the sum unchanged.
    float Al, Bl, Ar, Br;
    dd_invert_lin<MODE>(G, V, tl, Al, Bl);
    dd_invert_lin<MODE>(G, V, tr, Ar, Br);
    const float tt0 = dd_slope<MODE>(G, det_bnd(G.nu, G.du, G.su, u_lo));
    const float tt1 = dd_slope<MODE>(G, det_bnd(G.nu, G.du, G.su, u_lo + u_n));
    const float inv_d = 1.0f / G.d;
    // per-line table (host guarantees no <= kDDLines)
    for (int io = threadIdx.x; io < no; io += 256) {
        const float oc = cell_cen(no, G.d, io);
        const float y0 = dd_invert<MODE>(G, V, tt0, oc), y1 = dd_invert<MODE>(G, V, tt1, oc);
        const float t0f = fminf(y0, y1) * inv_d + 0.5f * (float)nd;
        const float t1f = fmaxf(y0, y1) * inv_d + 0.5f * (float)nd;
        short2 bandv = make_short2(0, -1);  // pixel indices < kDDLines fit 16 bits
        if (t0f < (float)nd + 3.0f && t1f > -3.0f) {
            bandv.x = (short)fmaxf(floorf(t0f) - 2.0f, 0.0f);
            bandv.y = (short)fminf(floorf(t1f) + 2.0f, (float)(nd - 1));
        }
        lband[io] = bandv;
        lsc[io] = dd_scale<MODE>(G, V, oc);
    }
    __syncthreads();
    // stage boundaries and values of pixels kc .. kc + nk - 1 of line io into slot h
    auto stage = [&](int h, int io, int kc, int nk) {
        const float oc = cell_cen(no, G.d, io);
        for (int e = threadIdx.x; e <= nk; e += 256) {
            bnd[h][e] = dd_bound<MODE>(G, V, cell_bnd(nd, G.d, kc + e), oc);
            if (e < nk) {
                const int k = kc + e;
                const int ix = V.xd ? k : io, iy = V.xd ? io : k;
                val[h][e] = img[(int64_t)ix * G.ny + iy];
            }
        }
    };
    double acc = 0.0;
    // this bin's pixels of line io within the staged chunk, as k_dd2d_fp
    auto merge = [&](int h, int io, int kc, int nk) {
        const float oc = cell_cen(no, G.d, io);
        const float x0 = __builtin_fmaf(-oc, Bl, Al), x1 = __builtin_fmaf(-oc, Br, Ar);
        const float k0f = fminf(x0, x1) * inv_d + 0.5f * (float)nd;
        const float k1f = fmaxf(x0, x1) * inv_d + 0.5f * (float)nd;
        if (!(k0f < (float)nd + 2.0f) || !(k1f > -2.0f)) return;
        const int k0 = (int)fmaxf(floorf(k0f) - 1.0f, 0.0f);
        const int k1 = (int)fminf(floorf(k1f) + 1.0f, (float)(nd - 1));
        const float sc = lsc[io];
        const int ka = max(k0, kc), kz = min(k1, kc + nk - 1);
        for (int k = ka; k <= kz; ++k) {
            const float ov = overlap(bnd[h][k - kc], bnd[h][k - kc + 1], ul, ur);
            if (ov > 0.0f) acc += (double)(val[h][k - kc] * (ov * sc));
        }
    };
    for (int io0 = 0; io0 < no; io0 += 2) {
        const short2 s0 = lband[io0];
        const short2 s1 = io0 + 1 < no ? lband[io0 + 1] : make_short2(0, -1);
        const int2 b0 = make_int2(s0.x, s0.y), b1 = make_int2(s1.x, s1.y);
        const int n0 = b0.y - b0.x + 1, n1 = b1.y - b1.x + 1;
        if (n0 <= 0 && n1 <= 0) continue;
        if (n0 <= kDDKMax && n1 <= kDDKMax) {
            // both lines in one chunk each: one pair of barriers for two lines
            __syncthreads();
            if (n0 > 0) stage(0, io0, b0.x, n0);
            if (n1 > 0) stage(1, io0 + 1, b1.x, n1);
            __syncthreads();
            if (ubin && n0 > 0) merge(0, io0, b0.x, n0);
            if (ubin && n1 > 0) merge(1, io0 + 1, b1.x, n1);
            continue;
        }
        // a long band: each line on its own, chunk by chunk
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int2 bb = h ? b1 : b0;
            for (int kc = bb.x; kc <= bb.y; kc += kDDKMax) {
                const int nk = min(kDDKMax, bb.y - kc + 1);
                __syncthreads();
                stage(0, io0 + h, kc, nk);
                __syncthreads();
                if (ubin) merge(0, io0 + h, kc, nk);
            }
        }
    }
    if (!ubin) return;
    const int64_t t = (int64_t)ia * G.nu + iu;
    float r = (float)acc * G.out_scale;
    if (accumulate) r += sino[t];
    sirt_store(E, sino, t, r);
}

// One thread per pixel.
template <int MODE>
__global__ __launch_bounds__(256) void k_dd2d_bp(DDGeo G, const float* __restrict__ cs,
                                                const float* __restrict__ sn, const float* __restrict__ sino,
                                                float* __restrict__ img, int accumulate, SirtEpi E) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nx * G.ny) return;
    const int ix = (int)(t / G.ny), iy = (int)(t % G.ny);
    const float u0 = det_bnd(G.nu, G.du, G.su, 0);
    const float inv_du = 1.0f / G.du;
    double acc = 0.0;
    for (int ia = 0; ia < G.na; ++ia) {
        const DDView V = dd_view(cs[ia], sn[ia]);
        const int nd = V.xd ? G.nx : G.ny, no = V.xd ? G.ny : G.nx;
        const int k = V.xd ? ix : iy, io = V.xd ? iy : ix;
        const float oc = cell_cen(no, G.d, io);
        const float pa = dd_bound<MODE>(G, V, cell_bnd(nd, G.d, k), oc);
        const float pb = dd_bound<MODE>(G, V, cell_bnd(nd, G.d, k + 1), oc);
        const float lo = fminf(pa, pb), hi = fmaxf(pa, pb);
        int j0 = (int)fmaxf(floorf((lo - u0) * inv_du) - 1.0f, 0.0f);
        int j1 = (int)fminf(floorf((hi - u0) * inv_du) + 1.0f, (float)(G.nu - 1));
        if (!(hi > u0 - G.du) || !((lo - u0) * inv_du < (float)G.nu + 1.0f)) continue;
        const float sc = dd_scale<MODE>(G, V, oc);
        const float* row = sino + (int64_t)ia * G.nu;
        for (int j = j0; j <= j1; ++j) {
            float ov = overlap(pa, pb, det_bnd(G.nu, G.du, G.su, j), det_bnd(G.nu, G.du, G.su, j + 1));
            if (ov > 0.0f) acc += (double)(row[j] * (ov * sc));
        }
    }
    float r = (float)acc * G.out_scale;
    if (accumulate) r += img[t];
    sirt_store(E, img, t, r);
}

// ---------------------------------------------------------------- cone

// slab z interval magnified onto the detector at the line's centre
__device__ __forceinline__ float cone_zmag(const DDGeo& G, const DDView& V, float orth_c) {
    return G.DSD / (G.DSO - V.oo * orth_c);
}

// One thread per (view, u, v) bin; sino is (na, nu, nv), vol (nx, ny, nz).
__global__ __launch_bounds__(256) void k_ddcone_fp(DDGeo G, const float* __restrict__ cs,
                                                  const float* __restrict__ sn, const float* __restrict__ vol,
                                                  float* __restrict__ sino, int accumulate) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.na * G.nu * G.nv) return;
    const int iv = (int)(t % G.nv);
    const int iu = (int)((t / G.nv) % G.nu);
    const int ia = (int)(t / ((int64_t)G.nv * G.nu));
    const DDView V = dd_view(cs[ia], sn[ia]);
    const int nd = V.xd ? G.nx : G.ny, no = V.xd ? G.ny : G.nx;
    const float ul = det_bnd(G.nu, G.du, G.su, iu), ur = det_bnd(G.nu, G.du, G.su, iu + 1);
    const float vl = det_bnd(G.nv, G.dv, G.sv, iv), vr = det_bnd(G.nv, G.dv, G.sv, iv + 1);
    const float tl = dd_slope<DD_FAN_FLAT>(G, ul), tr = dd_slope<DD_FAN_FLAT>(G, ur);
    const float inv_d = 1.0f / G.d;
    double acc = 0.0;
    for (int io = 0; io < no; ++io) {
        const float oc = cell_cen(no, G.d, io);
        float x0 = dd_invert<DD_FAN_FLAT>(G, V, tl, oc), x1 = dd_invert<DD_FAN_FLAT>(G, V, tr, oc);
        float k0f = fminf(x0, x1) * inv_d + 0.5f * (float)nd;
        float k1f = fmaxf(x0, x1) * inv_d + 0.5f * (float)nd;
        if (!(k0f < (float)nd + 2.0f) || !(k1f > -2.0f)) continue;
        const float zm = cone_zmag(G, V, oc);
        float z0f = vl / zm * inv_d + 0.5f * (float)G.nz, z1f = vr / zm * inv_d + 0.5f * (float)G.nz;
        if (!(z0f < (float)G.nz + 2.0f) || !(z1f > -2.0f)) continue;
        int k0 = (int)fmaxf(floorf(k0f) - 1.0f, 0.0f);
        int k1 = (int)fminf(floorf(k1f) + 1.0f, (float)(nd - 1));
        int z0 = (int)fmaxf(floorf(z0f) - 1.0f, 0.0f);
        int z1 = (int)fminf(floorf(z1f) + 1.0f, (float)(G.nz - 1));
        const float sc = dd_scale<DD_FAN_FLAT>(G, V, oc);
        float pb = dd_bound<DD_FAN_FLAT>(G, V, cell_bnd(nd, G.d, k0), oc);
        for (int k = k0; k <= k1; ++k) {
            float pn = dd_bound<DD_FAN_FLAT>(G, V, cell_bnd(nd, G.d, k + 1), oc);
            float ovu = overlap(pb, pn, ul, ur);
            pb = pn;
            if (!(ovu > 0.0f)) continue;
            int ix = V.xd ? k : io, iy = V.xd ? io : k;
            const float* col = vol + ((int64_t)ix * G.ny + iy) * G.nz;
            for (int z = z0; z <= z1; ++z) {
                float ovv = overlap(zm * cell_bnd(G.nz, G.d, z), zm * cell_bnd(G.nz, G.d, z + 1), vl, vr);
                if (ovv > 0.0f) acc += (double)(col[z] * ((ovu * ovv) * sc));
            }
        }
    }
    float r = (float)acc * G.out_scale;
    if (accumulate) r += sino[t];
    sino[t] = r;
}

// VIR_COMPAT cone forward: the reference's sweep (dd.py:298-379) with its
// u buffer, one thread per (view, bin) owning sino[ia, iu, :].  For every
// orthogonal line io and slab iz whose magnified z interval meets the
// detector rows, the buffer entry of bin iu is recomputed when iu lies in
// the line's re-zeroed range [iu_min, iu_max) (dd.py:313-317) and otherwise
// keeps the value of the last line that recomputed it; the entry times each
// row's v overlap is added to the sinogram in float32, in that order
// (dd.py:366-379).  Lines wholly beyond the detector (iu_min >= nu) add
// nothing.
__global__ __launch_bounds__(256) void k_ddcone_fp_compat(DDGeo G, const float* __restrict__ cs,
                                                         const float* __restrict__ sn,
                                                         const float* __restrict__ vol, float* __restrict__ sino,
                                                         int accumulate) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.na * G.nu) return;
    const int iu = (int)(t % G.nu);
    const int ia = (int)(t / G.nu);
    const DDView V = dd_view(cs[ia], sn[ia]);
    const int nd = V.xd ? G.nx : G.ny, no = V.xd ? G.ny : G.nx;
    const float u0 = det_bnd(G.nu, G.du, G.su, 0), v0 = det_bnd(G.nv, G.dv, G.sv, 0);
    const float ul = det_bnd(G.nu, G.du, G.su, iu), ur = det_bnd(G.nu, G.du, G.su, iu + 1);
    const float tl = dd_slope<DD_FAN_FLAT>(G, ul), tr = dd_slope<DD_FAN_FLAT>(G, ur);
    const float inv_d = 1.0f / G.d, inv_du = 1.0f / G.du, inv_dv = 1.0f / G.dv;
    float* row = sino + t * G.nv;
    if (!accumulate)
        for (int iv = 0; iv < G.nv; ++iv) row[iv] = 0.0f;
    float buf = 0.0f;
    for (int io = 0; io < no; ++io) {
        const float oc = cell_cen(no, G.d, io);
        const float pa = dd_bound<DD_FAN_FLAT>(G, V, cell_bnd(nd, G.d, 0), oc);
        const float pz = dd_bound<DD_FAN_FLAT>(G, V, cell_bnd(nd, G.d, nd), oc);
        // int() truncation toward zero, arguments clamped to [-1, nu + 1]
        const float fa = fminf(fmaxf((fminf(pa, pz) - u0) * inv_du, -1.0f), (float)G.nu + 1.0f);
        const float fz = fminf(fmaxf((fmaxf(pa, pz) - u0) * inv_du, -1.0f), (float)G.nu + 1.0f);
        const int iu_min = max((int)fa, 0), iu_max = min((int)fz + 1, G.nu);
        if (iu_min >= G.nu) continue;
        const bool refresh = iu >= iu_min && iu < iu_max;
        const float zm = cone_zmag(G, V, oc);
        const float sc = dd_scale<DD_FAN_FLAT>(G, V, oc);
        // pixels of this line whose shadows can meet the bin
        float x0 = dd_invert<DD_FAN_FLAT>(G, V, tl, oc), x1 = dd_invert<DD_FAN_FLAT>(G, V, tr, oc);
        const float k0f = fminf(fmaxf(fminf(x0, x1) * inv_d + 0.5f * (float)nd, -4.0f), (float)nd + 4.0f);
        const float k1f = fminf(fmaxf(fmaxf(x0, x1) * inv_d + 0.5f * (float)nd, -4.0f), (float)nd + 4.0f);
        const int k0 = (int)fmaxf(floorf(k0f) - 1.0f, 0.0f);
        const int k1 = (int)fminf(floorf(k1f) + 1.0f, (float)(nd - 1));
        for (int iz = 0; iz < G.nz; ++iz) {
            const float zl = zm * cell_bnd(G.nz, G.d, iz), zr = zm * cell_bnd(G.nz, G.d, iz + 1);
            const float ga = fminf(fmaxf((zl - v0) * inv_dv, -1.0f), (float)G.nv + 1.0f);
            const float gz = fminf(fmaxf((zr - v0) * inv_dv, -1.0f), (float)G.nv + 1.0f);
            const int iv_min = max((int)ga, 0), iv_max = min((int)gz + 1, G.nv);
            if (iv_max <= 0 || iv_min >= G.nv) continue;
            if (refresh) {
                buf = 0.0f;
                float pb = dd_bound<DD_FAN_FLAT>(G, V, cell_bnd(nd, G.d, k0), oc);
                for (int k = k0; k <= k1; ++k) {
                    const float pn = dd_bound<DD_FAN_FLAT>(G, V, cell_bnd(nd, G.d, k + 1), oc);
                    const float ovu = overlap(pb, pn, ul, ur);
                    pb = pn;
                    if (!(ovu > 0.0f)) continue;
                    const int ix = V.xd ? k : io, iy = V.xd ? io : k;
                    buf += (vol[((int64_t)ix * G.ny + iy) * G.nz + iz] * sc) * ovu;
                }
            }
            if (buf == 0.0f) continue;
            for (int iv = iv_min; iv < iv_max; ++iv) {
                const float ov = fminf(zr, det_bnd(G.nv, G.dv, G.sv, iv + 1)) - fmaxf(zl, det_bnd(G.nv, G.dv, G.sv, iv));
                if (ov > 0.0f) row[iv] += (buf * ov) * G.out_scale;
            }
        }
    }
}

// One thread per voxel.
__global__ __launch_bounds__(256) void k_ddcone_bp(DDGeo G, const float* __restrict__ cs,
                                                  const float* __restrict__ sn, const float* __restrict__ sino,
                                                  float* __restrict__ vol, int accumulate) {
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nx * G.ny * G.nz) return;
    const int iz = (int)(t % G.nz);
    const int iy = (int)((t / G.nz) % G.ny);
    const int ix = (int)(t / ((int64_t)G.nz * G.ny));
    const float u0 = det_bnd(G.nu, G.du, G.su, 0), v0 = det_bnd(G.nv, G.dv, G.sv, 0);
    const float inv_du = 1.0f / G.du, inv_dv = 1.0f / G.dv;
    const float zb0 = cell_bnd(G.nz, G.d, iz), zb1 = cell_bnd(G.nz, G.d, iz + 1);
    double acc = 0.0;
    for (int ia = 0; ia < G.na; ++ia) {
        const DDView V = dd_view(cs[ia], sn[ia]);
        const int nd = V.xd ? G.nx : G.ny, no = V.xd ? G.ny : G.nx;
        const int k = V.xd ? ix : iy, io = V.xd ? iy : ix;
        const float oc = cell_cen(no, G.d, io);
        const float pa = dd_bound<DD_FAN_FLAT>(G, V, cell_bnd(nd, G.d, k), oc);
        const float pb = dd_bound<DD_FAN_FLAT>(G, V, cell_bnd(nd, G.d, k + 1), oc);
        const float zm = cone_zmag(G, V, oc);
        const float qa = zm * zb0, qb = zm * zb1;
        const float lo = fminf(pa, pb), hi = fmaxf(pa, pb);
        const float vlo = fminf(qa, qb), vhi = fmaxf(qa, qb);
        if (!(hi > u0 - G.du) || !((lo - u0) * inv_du < (float)G.nu + 1.0f)) continue;
        if (!(vhi > v0 - G.dv) || !((vlo - v0) * inv_dv < (float)G.nv + 1.0f)) continue;
        int j0 = (int)fmaxf(floorf((lo - u0) * inv_du) - 1.0f, 0.0f);
        int j1 = (int)fminf(floorf((hi - u0) * inv_du) + 1.0f, (float)(G.nu - 1));
        int m0 = (int)fmaxf(floorf((vlo - v0) * inv_dv) - 1.0f, 0.0f);
        int m1 = (int)fminf(floorf((vhi - v0) * inv_dv) + 1.0f, (float)(G.nv - 1));
        const float sc = dd_scale<DD_FAN_FLAT>(G, V, oc);
        const float* plane = sino + (int64_t)ia * G.nu * G.nv;
        for (int j = j0; j <= j1; ++j) {
            float ovu = overlap(pa, pb, det_bnd(G.nu, G.du, G.su, j), det_bnd(G.nu, G.du, G.su, j + 1));
            if (!(ovu > 0.0f)) continue;
            for (int m = m0; m <= m1; ++m) {
                float ovv = overlap(qa, qb, det_bnd(G.nv, G.dv, G.sv, m), det_bnd(G.nv, G.dv, G.sv, m + 1));
                if (ovv > 0.0f) acc += (double)(plane[(int64_t)j * G.nv + m] * ((ovu * ovv) * sc));
            }
        }
    }
    float r = (float)acc * G.out_scale;
    if (accumulate) r += vol[t];
    vol[t] = r;
}

static int dd_check(const vir_dd_geom* g, DDGeo& G, bool cone) {
    if (!g) return fail_arg("geometry is NULL");
    G.nx = g->nx; G.ny = g->ny; G.nz = cone ? g->nz : 1;
    G.nu = g->nu; G.nv = cone ? g->nv : 1; G.na = g->na;
    G.d = g->d_pix; G.du = g->du; G.dv = cone ? g->dv : 1.0f;
    G.su = g->su; G.sv = cone ? g->sv : 0.0f;
    G.DSO = g->DSO; G.DSD = g->DSD;
    if (G.nx < 1 || G.ny < 1 || G.nz < 1 || G.nu < 1 || G.nv < 1 || G.na < 0) return fail_arg("bad sizes");
    if (!(G.d > 0.0f) || !(G.du > 0.0f) || !(G.dv > 0.0f)) return fail_arg("spacings must be > 0");
    if (g->mode != DD_PAR || cone) {
        if (!(G.DSO > 0.0f) || !(G.DSD > 0.0f)) return fail_arg("DSO and DSD must be > 0");
        // every pixel must lie strictly in front of the source (monotone
        // projections): half-diagonal of the grid < DSO
        float hx = 0.5f * G.nx * G.d, hy = 0.5f * G.ny * G.d;
        if (!(sqrtf(hx * hx + hy * hy) < G.DSO)) return fail_arg("grid reaches the source (DSO too small)");
    }
    return VIR_OK;
}

}  // namespace vir

using namespace vir;

static int dd2d_fp(const vir_dd_geom* geom, const float* cosv, const float* sinv, const float* img, float* sino,
                   float out_scale, uint32_t flags, void* stream, const SirtEpi& E) {
    DDGeo G;
    int rc = dd_check(geom, G, false);
    if (rc) return rc;
    if (geom->mode < 0 || geom->mode > 2) return fail_arg("mode must be 0 (par), 1 (fan flat) or 2 (fan arc)");
    if (G.na == 0) return VIR_OK;
    if (!cosv || !sinv || !img || !sino) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    int64_t n = (int64_t)G.na * G.nu;
    dim3 grd((unsigned)((n + 255) / 256));
    int acc = (flags & VIR_ACCUMULATE) ? 1 : 0;
    hipStream_t st = as_stream(stream);
    if (!(flags & VIR_PER_OUTPUT) && std::max(G.nx, G.ny) <= kDDLines) {
        const int nut = (G.nu + 255) / 256;
        dim3 tg((unsigned)((int64_t)G.na * nut));
        const size_t tables = (size_t)std::max(G.nx, G.ny) * (sizeof(float) + sizeof(short2));
        if (geom->mode == DD_PAR)
            hipLaunchKernelGGL(k_dd2d_fp_tile<DD_PAR>, tg, dim3(256), tables, st, G, cosv, sinv, img, sino, acc, nut, E);
        else if (geom->mode == DD_FAN_FLAT)
            hipLaunchKernelGGL(k_dd2d_fp_tile<DD_FAN_FLAT>, tg, dim3(256), tables, st, G, cosv, sinv, img, sino, acc, nut, E);
        else hipLaunchKernelGGL(k_dd2d_fp_tile<DD_FAN_ARC>, tg, dim3(256), tables, st, G, cosv, sinv, img, sino, acc, nut, E);
        return check_hip(hipGetLastError(), "vir_dd2d_fp launch");
    }
    if (geom->mode == DD_PAR) hipLaunchKernelGGL(k_dd2d_fp<DD_PAR>, grd, dim3(256), 0, st, G, cosv, sinv, img, sino, acc, E);
    else if (geom->mode == DD_FAN_FLAT)
        hipLaunchKernelGGL(k_dd2d_fp<DD_FAN_FLAT>, grd, dim3(256), 0, st, G, cosv, sinv, img, sino, acc, E);
    else hipLaunchKernelGGL(k_dd2d_fp<DD_FAN_ARC>, grd, dim3(256), 0, st, G, cosv, sinv, img, sino, acc, E);
    return check_hip(hipGetLastError(), "vir_dd2d_fp launch");
}

static int dd2d_bp(const vir_dd_geom* geom, const float* cosv, const float* sinv, const float* sino, float* img,
                   float out_scale, uint32_t flags, void* stream, const SirtEpi& E) {
    DDGeo G;
    int rc = dd_check(geom, G, false);
    if (rc) return rc;
    if (geom->mode < 0 || geom->mode > 2) return fail_arg("mode must be 0 (par), 1 (fan flat) or 2 (fan arc)");
    if ((G.na > 0 && (!cosv || !sinv || !sino)) || !img) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    int64_t n = (int64_t)G.nx * G.ny;
    dim3 grd((unsigned)((n + 255) / 256));
    int acc = (flags & VIR_ACCUMULATE) ? 1 : 0;
    hipStream_t st = as_stream(stream);
    if (geom->mode == DD_PAR) hipLaunchKernelGGL(k_dd2d_bp<DD_PAR>, grd, dim3(256), 0, st, G, cosv, sinv, sino, img, acc, E);
    else if (geom->mode == DD_FAN_FLAT)
        hipLaunchKernelGGL(k_dd2d_bp<DD_FAN_FLAT>, grd, dim3(256), 0, st, G, cosv, sinv, sino, img, acc, E);
    else hipLaunchKernelGGL(k_dd2d_bp<DD_FAN_ARC>, grd, dim3(256), 0, st, G, cosv, sinv, sino, img, acc, E);
    return check_hip(hipGetLastError(), "vir_dd2d_bp launch");
}

extern "C" {

int vir_dd2d_fp(const vir_dd_geom* geom, const float* cosv, const float* sinv, const float* img, float* sino,
                float out_scale, uint32_t flags, void* stream) {
    return dd2d_fp(geom, cosv, sinv, img, sino, out_scale, flags, stream, SirtEpi{});
}

int vir_dd2d_bp(const vir_dd_geom* geom, const float* cosv, const float* sinv, const float* sino, float* img,
                float out_scale, uint32_t flags, void* stream) {
    return dd2d_bp(geom, cosv, sinv, sino, img, out_scale, flags, stream, SirtEpi{});
}

int vir_dd2d_fp_res(const vir_dd_geom* geom, const float* cosv, const float* sinv, const float* img, const float* b,
                    const float* w_sino, float* r, float out_scale, uint32_t flags, void* stream) {
    if (!b || !w_sino) return fail_arg("NULL pointer argument");
    if (flags & VIR_ACCUMULATE) return fail_arg("VIR_ACCUMULATE does not apply to the fused residual");
    return dd2d_fp(geom, cosv, sinv, img, r, out_scale, flags, stream, SirtEpi{b, w_sino, nullptr, 0.0f});
}

int vir_dd2d_bp_upd(const vir_dd_geom* geom, const float* cosv, const float* sinv, const float* r,
                    const float* w_img, float gamma, float* x, float out_scale, uint32_t flags, void* stream) {
    if (!w_img || !x) return fail_arg("NULL pointer argument");
    if (flags & VIR_ACCUMULATE) return fail_arg("VIR_ACCUMULATE does not apply to the fused update");
    return dd2d_bp(geom, cosv, sinv, r, x, out_scale, flags, stream, SirtEpi{nullptr, w_img, x, gamma});
}

int vir_ddcone_fp(const vir_dd_geom* geom, const float* cosv, const float* sinv, const float* vol, float* sino,
                  float out_scale, uint32_t flags, void* stream) {
    DDGeo G;
    int rc = dd_check(geom, G, true);
    if (rc) return rc;
    if (G.na == 0) return VIR_OK;
    if (!cosv || !sinv || !vol || !sino) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    if (flags & VIR_COMPAT) {
        const int64_t nb = (int64_t)G.na * G.nu;
        hipLaunchKernelGGL(k_ddcone_fp_compat, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, as_stream(stream),
                           G, cosv, sinv, vol, sino, (flags & VIR_ACCUMULATE) ? 1 : 0);
        return check_hip(hipGetLastError(), "vir_ddcone_fp launch");
    }
    int64_t n = (int64_t)G.na * G.nu * G.nv;
    hipLaunchKernelGGL(k_ddcone_fp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), G, cosv,
                       sinv, vol, sino, (flags & VIR_ACCUMULATE) ? 1 : 0);
    return check_hip(hipGetLastError(), "vir_ddcone_fp launch");
}

int vir_ddcone_bp(const vir_dd_geom* geom, const float* cosv, const float* sinv, const float* sino, float* vol,
                  float out_scale, uint32_t flags, void* stream) {
    DDGeo G;
    int rc = dd_check(geom, G, true);
    if (rc) return rc;
    if ((G.na > 0 && (!cosv || !sinv || !sino)) || !vol) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    int64_t n = (int64_t)G.nx * G.ny * G.nz;
    hipLaunchKernelGGL(k_ddcone_bp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(stream), G, cosv,
                       sinv, sino, vol, (flags & VIR_ACCUMULATE) ? 1 : 0);
    return check_hip(hipGetLastError(), "vir_ddcone_bp launch");
}

}  // extern "C"
