// Ray-driven projectors for gfx950 (sys_mat/rd.py): Amanatides-Woo voxel
// traversal (path length per voxel) and Joseph interpolation (samples every
// 0.5 length units, bi/trilinear), 2-D parallel, 2-D fan (flat detector)
// and 3-D cone.
//
// Rays (rd.py:472-790): bin centres u = du*(iu - (nu/2 - 0.5) + su) and, for
// the cone, v = dv*(iv - (nu/2 - 0.5) + sv) (the reference centres v with nu);
// parallel rays start at u*(-sin, cos) along (cos, sin); fan/cone rays run
// from the source DSO*(cos, sin, 0) towards the flat detector at
// -(DSD - DSO)*(cos, sin) + u*(-sin, cos) + v*z, unit direction.  [t0, t1] is
// the ray's intersection with the image box (rd.py:51-64, 595-605).
//
// Forward (one thread per ray):
//   AW      sum over voxels of value * intersection length; plane crossings
//           are computed from the plane index (no accumulated t_next += step)
//           and ties step x before y before z (rd.py:111-247); the cone
//           result is scaled by DSD / sqrt(DSD^2 + u^2 + v^2) (rd.py:784-785).
//   Joseph  samples t0 + k/2 <= t1; 2-D clamps the sample into
//           [0, n-2] and interpolates bilinearly (rd.py:250-289); 3-D
//           truncates toward zero and adds only samples whose cell is inside
//           (rd.py:292-353).
// Backward (one thread per voxel, a gather over the rays whose footprint
// can reach it; no atomics): the reference's backprojections (rd.py:151-175,
// 215-247 path-length splat; 356-445 bilinear / trilinear splat with floor
// indices and per-corner bounds), i.e. the transpose of AW's forward (cone:
// without the ray scale), and Joseph's splat.
//
// Rays parallel to an axis (|direction| <= 1e-6) that lie outside the image
// along that axis miss; the reference clamps their entry index onto the
// border row instead (rd.py:66-73), which VIR_COMPAT reproduces: such a ray
// then hits, its index on that axis clamped, and the backprojection also
// visits, for border voxels, the bins whose rays can be axis-parallel.
#include <cmath>

#include "vir_common.h"

namespace vir {

enum RDMode { RD_PAR2 = 0, RD_FAN2 = 1, RD_CONE3 = 2 };
constexpr float kRdEps = 1e-6f;
constexpr float kRdStep = 0.5f;

struct RDGeo {
    int nx, ny, nz, nu, nv, na;
    float d, du, dv, su, sv, DSO, DSD;
    float ucen;   // nu/2 - 0.5 in float32, the bin-centre offset
    float out_scale;
    float bmin[3], bmax[3], c0[3];
    int compat;   // VIR_COMPAT: axis-parallel rays outside the image hit (clamped)
};

struct RDRay {
    float o[3], h[3];
    float t0, t1;
    float u, v;
    bool hit;
};

__device__ __forceinline__ float rd_ucoord(const RDGeo& G, int iu) {
    return G.du * (((float)iu - G.ucen) + G.su);
}
__device__ __forceinline__ float rd_vcoord(const RDGeo& G, int iv) {
    return G.dv * (((float)iv - G.ucen) + G.sv);
}

template <int MODE>
__device__ __forceinline__ RDRay rd_ray(const RDGeo& G, float c, float s, int iu, int iv) {
    RDRay R;
    const int na = MODE == RD_CONE3 ? 3 : 2;
    R.o[2] = 0.0f;
    R.h[2] = 0.0f;
    R.u = rd_ucoord(G, iu);
    R.v = MODE == RD_CONE3 ? rd_vcoord(G, iv) : 0.0f;
    if (MODE == RD_PAR2) {
        R.o[0] = R.u * -s;
        R.o[1] = R.u * c;
        R.h[0] = c;
        R.h[1] = s;
    } else {
        R.o[0] = G.DSO * c;
        R.o[1] = G.DSO * s;
        const float dx = -(G.DSD - G.DSO) * c + R.u * -s;
        const float dy = -(G.DSD - G.DSO) * s + R.u * c;
        const float vx = dx - R.o[0], vy = dy - R.o[1];
        if (MODE == RD_CONE3) {
            R.o[2] = 0.0f;
            const float vz = R.v;
            const float m = __builtin_sqrtf(vx * vx + vy * vy + vz * vz);
            R.h[0] = vx / m;
            R.h[1] = vy / m;
            R.h[2] = vz / m;
        } else {
            const float m = __builtin_sqrtf(vx * vx + vy * vy);
            R.h[0] = vx / m;
            R.h[1] = vy / m;
        }
    }
    float t0 = -VIR_INF, t1 = VIR_INF;
    bool miss = false;
    for (int a = 0; a < na; ++a) {
        if (fabsf(R.h[a]) > kRdEps) {
            const float ta = (G.bmin[a] - R.o[a]) / R.h[a], tb = (G.bmax[a] - R.o[a]) / R.h[a];
            t0 = fmaxf(t0, fminf(ta, tb));
            t1 = fminf(t1, fmaxf(ta, tb));
        } else if (!(R.o[a] >= G.bmin[a] && R.o[a] < G.bmax[a])) {
            miss = true;
        }
    }
    R.t0 = t0;
    R.t1 = t1;
    R.hit = (!miss || G.compat) && t1 > t0;
    return R;
}

__device__ __forceinline__ int rd_index(const RDGeo& G, int a, float r, int n) {
    r = fminf(fmaxf(r, G.bmin[a] + kRdEps), G.bmax[a] - kRdEps);
    int i = (int)floorf((r - G.bmin[a]) / G.d);
    return min(max(i, 0), n - 1);
}

// ---------------------------------------------------------------- forward

template <int MODE>
__device__ float rd_aw_sum(const RDGeo& G, const RDRay& R, const float* __restrict__ img) {
    const int na = MODE == RD_CONE3 ? 3 : 2;
    const int n[3] = {G.nx, G.ny, MODE == RD_CONE3 ? G.nz : 1};
    int idx[3] = {0, 0, 0}, dir[3] = {0, 0, 0}, nxt[3] = {0, 0, 0};
    float tn[3] = {VIR_INF, VIR_INF, VIR_INF};
    for (int a = 0; a < na; ++a) {
        idx[a] = rd_index(G, a, R.o[a] + R.t0 * R.h[a], n[a]);
        if (fabsf(R.h[a]) > kRdEps) {
            dir[a] = R.h[a] > 0.0f ? 1 : -1;
            nxt[a] = R.h[a] > 0.0f ? idx[a] + 1 : idx[a];
            tn[a] = fmaxf((G.bmin[a] + (float)nxt[a] * G.d - R.o[a]) / R.h[a], R.t0);
        }
    }
    double acc = 0.0;
    float t = R.t0;
    while (t < R.t1) {
        if (idx[0] < 0 || idx[0] >= n[0] || idx[1] < 0 || idx[1] >= n[1] || idx[2] < 0 || idx[2] >= n[2]) break;
        int a;
        if (na == 3) a = (tn[0] <= tn[1] && tn[0] <= tn[2]) ? 0 : (tn[1] <= tn[2] ? 1 : 2);
        else a = tn[0] <= tn[1] ? 0 : 1;
        const float te = fminf(tn[a], R.t1);
        const int64_t vi = na == 3 ? ((int64_t)idx[0] * G.ny + idx[1]) * G.nz + idx[2]
                                   : (int64_t)idx[0] * G.ny + idx[1];
        acc += (double)(img[vi] * (te - t));
        t = te;
        idx[a] += dir[a];
        nxt[a] += dir[a];
        tn[a] = (G.bmin[a] + (float)nxt[a] * G.d - R.o[a]) / R.h[a];
    }
    return (float)acc;
}

template <int MODE>
__device__ float rd_joseph_sum(const RDGeo& G, const RDRay& R, const float* __restrict__ img) {
    double acc = 0.0;
    const float inv_d = 1.0f / G.d;
    for (int k = 0;; ++k) {
        const float t = R.t0 + (float)k * kRdStep;
        if (!(t <= R.t1)) break;
        float fx = (R.o[0] + t * R.h[0] - G.c0[0]) * inv_d;
        float fy = (R.o[1] + t * R.h[1] - G.c0[1]) * inv_d;
        if (MODE != RD_CONE3) {
            fx = fminf(fmaxf(fx, 0.0f), (float)G.nx - 2.0f);
            fy = fminf(fmaxf(fy, 0.0f), (float)G.ny - 2.0f);
            const int ix = (int)fx, iy = (int)fy;
            const float dx = fx - (float)ix, dy = fy - (float)iy;
            const float* p = img + (int64_t)ix * G.ny + iy;
            const float val = p[0] * (1.0f - dx) * (1.0f - dy) + p[G.ny] * dx * (1.0f - dy) +
                              p[1] * (1.0f - dx) * dy + p[G.ny + 1] * dx * dy;
            acc += (double)(val * kRdStep);
        } else {
            const float fz = (R.o[2] + t * R.h[2] - G.c0[2]) * inv_d;
            const int ix = (int)fx, iy = (int)fy, iz = (int)fz;   // truncation, as int() in rd.py
            if (ix >= 0 && ix < G.nx - 1 && iy >= 0 && iy < G.ny - 1 && iz >= 0 && iz < G.nz - 1) {
                const float dx = fx - (float)ix, dy = fy - (float)iy, dz = fz - (float)iz;
                const int64_t sy = G.nz, sx = (int64_t)G.ny * G.nz;
                const float* p = img + ix * sx + iy * sy + iz;
                const float c00 = p[0] + dx * (p[sx] - p[0]);
                const float c10 = p[sy] + dx * (p[sx + sy] - p[sy]);
                const float c01 = p[1] + dx * (p[sx + 1] - p[1]);
                const float c11 = p[sy + 1] + dx * (p[sx + sy + 1] - p[sy + 1]);
                const float c0v = c00 + dy * (c10 - c00);
                const float c1v = c01 + dy * (c11 - c01);
                acc += (double)((c0v + dz * (c1v - c0v)) * kRdStep);
            }
        }
    }
    return (float)acc;
}

template <int MODE, bool JOSEPH>
__global__ __launch_bounds__(256) void k_rd_fp(RDGeo G, const float* __restrict__ cs, const float* __restrict__ sn,
                                              const float* __restrict__ img, float* __restrict__ sino,
                                              int accumulate) {
    const int nv = MODE == RD_CONE3 ? G.nv : 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.na * G.nu * nv) return;
    const int iv = (int)(t % nv);
    const int iu = (int)((t / nv) % G.nu);
    const int ia = (int)(t / ((int64_t)nv * G.nu));
    const RDRay R = rd_ray<MODE>(G, cs[ia], sn[ia], iu, iv);
    float val = 0.0f;
    if (R.hit) {
        val = JOSEPH ? rd_joseph_sum<MODE>(G, R, img) : rd_aw_sum<MODE>(G, R, img);
        if (MODE == RD_CONE3 && !JOSEPH)
            val = val * (G.DSD / __builtin_sqrtf(G.DSD * G.DSD + R.u * R.u + R.v * R.v));
    }
    val *= G.out_scale;
    if (accumulate) val += sino[t];
    sino[t] = val;
}

// --------------------------------------------------------------- backward

// Detector coordinates (u, v) of world point (x, y, z) in view (c, s).
template <int MODE>
__device__ __forceinline__ void rd_project(const RDGeo& G, float c, float s, float x, float y, float z, float& u,
                                           float& v) {
    const float p = -s * x + c * y;
    if (MODE == RD_PAR2) {
        u = p;
        v = 0.0f;
        return;
    }
    const float m = G.DSD / (G.DSO - (c * x + s * y));
    u = m * p;
    v = m * z;
}

// Bin index ranges [lo, hi] whose rays can meet the box [x0,x1]x[y0,y1]x[z0,z1].
template <int MODE>
__device__ __forceinline__ void rd_bins(const RDGeo& G, float c, float s, const float* lo, const float* hi, int& u0,
                                        int& u1, int& v0, int& v1) {
    float umin = VIR_INF, umax = -VIR_INF, vmin = VIR_INF, vmax = -VIR_INF;
    const int nc = MODE == RD_CONE3 ? 8 : 4;
    for (int k = 0; k < nc; ++k) {
        float u, v;
        rd_project<MODE>(G, c, s, (k & 1) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1], (k & 4) ? hi[2] : lo[2], u, v);
        umin = fminf(umin, u);
        umax = fmaxf(umax, u);
        vmin = fminf(vmin, v);
        vmax = fmaxf(vmax, v);
    }
    const float a0 = umin / G.du + G.ucen - G.su, a1 = umax / G.du + G.ucen - G.su;
    u0 = (int)fmaxf(floorf(a0) - 1.0f, 0.0f);
    u1 = (int)fminf(ceilf(a1) + 1.0f, (float)(G.nu - 1));
    if (!(a1 > -2.0f) || !(a0 < (float)G.nu + 1.0f)) u1 = u0 - 1;
    if (MODE == RD_CONE3) {
        const float b0 = vmin / G.dv + G.ucen - G.sv, b1 = vmax / G.dv + G.ucen - G.sv;
        v0 = (int)fmaxf(floorf(b0) - 1.0f, 0.0f);
        v1 = (int)fminf(ceilf(b1) + 1.0f, (float)(G.nv - 1));
        if (!(b1 > -2.0f) || !(b0 < (float)G.nv + 1.0f)) v1 = v0 - 1;
    } else {
        v0 = 0;
        v1 = 0;
    }
}

// Path length of ray R inside voxel (vx, vy, vz), clipped to [t0, t1].
template <int MODE>
__device__ __forceinline__ float rd_aw_len(const RDGeo& G, const RDRay& R, const int* vi) {
    const int na = MODE == RD_CONE3 ? 3 : 2;
    const int n[3] = {G.nx, G.ny, G.nz};
    float lo = R.t0, hi = R.t1;
    for (int a = 0; a < na; ++a) {
        const float pl = G.bmin[a] + (float)vi[a] * G.d, ph = G.bmin[a] + (float)(vi[a] + 1) * G.d;
        if (fabsf(R.h[a]) > kRdEps) {
            const float ta = (pl - R.o[a]) / R.h[a], tb = (ph - R.o[a]) / R.h[a];
            lo = fmaxf(lo, fminf(ta, tb));
            hi = fminf(hi, fmaxf(ta, tb));
        } else if (rd_index(G, a, R.o[a], n[a]) != vi[a]) {
            return 0.0f;
        }
    }
    return hi > lo ? hi - lo : 0.0f;
}

// Joseph splat weight of ray R into voxel vi (sum over its samples).
template <int MODE>
__device__ __forceinline__ float rd_joseph_w(const RDGeo& G, const RDRay& R, const int* vi) {
    const int na = MODE == RD_CONE3 ? 3 : 2;
    const float inv_d = 1.0f / G.d;
    // samples whose floor cell is vi - 1 or vi on every axis lie in this box
    float lo = R.t0, hi = R.t1;
    for (int a = 0; a < na; ++a) {
        const float pl = G.c0[a] + (float)(vi[a] - 1) * G.d, ph = G.c0[a] + (float)(vi[a] + 1) * G.d;
        if (fabsf(R.h[a]) > kRdEps) {
            const float ta = (pl - R.o[a]) / R.h[a], tb = (ph - R.o[a]) / R.h[a];
            lo = fmaxf(lo, fminf(ta, tb));
            hi = fminf(hi, fmaxf(ta, tb));
        }
    }
    if (!(hi >= lo)) return 0.0f;
    int k0 = (int)fmaxf(floorf((lo - R.t0) / kRdStep) - 1.0f, 0.0f);
    int k1 = (int)floorf((hi - R.t0) / kRdStep) + 1;
    float w = 0.0f;
    for (int k = k0; k <= k1; ++k) {
        const float t = R.t0 + (float)k * kRdStep;
        if (!(t <= R.t1)) break;
        float wk = 1.0f;
        for (int a = 0; a < na; ++a) {
            const float f = (R.o[a] + t * R.h[a] - G.c0[a]) * inv_d;
            const float fl = floorf(f);
            const float dd = f - fl;
            const int i = (int)fl;
            wk *= i == vi[a] ? 1.0f - dd : (i == vi[a] - 1 ? dd : 0.0f);
        }
        w += wk;
    }
    return w * kRdStep;
}

// Bins [b0, b1] of view (c, s) whose rays can run parallel to axis a
// (|h_a| <= kRdEps, rounding included); b1 < b0 when none can.  Parallel
// beam: all bins or none.  Fan/cone: h_a = v_a/|v| with v_x = -DSD c - u s,
// v_y = -DSD s + u c, so |v_a| <= W bounds u to an interval.
template <int MODE>
__device__ void rd_parallel_bins(const RDGeo& G, float c, float s, int a, int& b0, int& b1) {
    b0 = 0;
    b1 = -1;
    if (MODE == RD_PAR2) {
        if (fabsf(a == 0 ? c : s) <= kRdEps) b1 = G.nu - 1;
        return;
    }
    const float k = a == 0 ? -s : c, q = a == 0 ? -G.DSD * c : -G.DSD * s;
    const float umax = fmaxf(fabsf(rd_ucoord(G, 0)), fabsf(rd_ucoord(G, G.nu - 1)));
    const float vmax = MODE == RD_CONE3 ? fmaxf(fabsf(rd_vcoord(G, 0)), fabsf(rd_vcoord(G, G.nv - 1))) : 0.0f;
    const float W = 2e-6f * __builtin_sqrtf(G.DSD * G.DSD + umax * umax + vmax * vmax) + 1e-6f * (G.DSD + umax);
    if (k == 0.0f) {
        if (fabsf(q) <= W) b1 = G.nu - 1;
        return;
    }
    const float ua = (-W - q) / k, ub = (W - q) / k;
    const float fa = fminf(ua, ub) / G.du + G.ucen - G.su, fb = fmaxf(ua, ub) / G.du + G.ucen - G.su;
    b0 = (int)fminf(fmaxf(floorf(fa) - 1.0f, 0.0f), (float)G.nu);
    b1 = (int)fminf(fmaxf(ceilf(fb) + 1.0f, -1.0f), (float)(G.nu - 1));
}

// Adds ray (ia, iu, iv)'s weight on voxel vi times its sinogram value.
template <int MODE, bool JOSEPH>
__device__ __forceinline__ void rd_bp_bin(const RDGeo& G, const float* __restrict__ sino, float c, float s, int ia,
                                          int iu, int iv, const int* vi, double& acc) {
    const int nv = MODE == RD_CONE3 ? G.nv : 1;
    const float y = sino[((int64_t)ia * G.nu + iu) * nv + iv];
    if (y == 0.0f) return;
    const RDRay R = rd_ray<MODE>(G, c, s, iu, iv);
    if (!R.hit) return;
    const float w = JOSEPH ? rd_joseph_w<MODE>(G, R, vi) : rd_aw_len<MODE>(G, R, vi);
    if (w != 0.0f) acc += (double)(y * w);
}

template <int MODE, bool JOSEPH>
__global__ __launch_bounds__(256) void k_rd_bp(RDGeo G, const float* __restrict__ cs, const float* __restrict__ sn,
                                              const float* __restrict__ sino, float* __restrict__ img,
                                              int accumulate) {
    const int nz = MODE == RD_CONE3 ? G.nz : 1, nv = MODE == RD_CONE3 ? G.nv : 1;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nx * G.ny * nz) return;
    int vi[3];
    vi[2] = (int)(t % nz);
    vi[1] = (int)((t / nz) % G.ny);
    vi[0] = (int)(t / ((int64_t)nz * G.ny));
    // the box whose rays can reach this voxel
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        if (JOSEPH) {
            lo[a] = G.c0[a] + (float)(vi[a] - 1) * G.d;
            hi[a] = G.c0[a] + (float)(vi[a] + 1) * G.d;
        } else {
            lo[a] = G.bmin[a] + (float)vi[a] * G.d;
            hi[a] = G.bmin[a] + (float)(vi[a] + 1) * G.d;
        }
    }
    if (MODE != RD_CONE3) { lo[2] = 0.0f; hi[2] = 0.0f; }
    // VIR_COMPAT, AW: a ray parallel to axis a outside the image is clamped
    // onto voxel row 0 or n-1 of that axis without passing through it
    const bool edge[2] = {G.compat && !JOSEPH && (vi[0] == 0 || vi[0] == G.nx - 1),
                          G.compat && !JOSEPH && (vi[1] == 0 || vi[1] == G.ny - 1)};
    double acc = 0.0;
    for (int ia = 0; ia < G.na; ++ia) {
        const float c = cs[ia], s = sn[ia];
        int u0, u1, v0, v1;
        rd_bins<MODE>(G, c, s, lo, hi, u0, u1, v0, v1);
        for (int iu = u0; iu <= u1; ++iu)
            for (int iv = v0; iv <= v1; ++iv) rd_bp_bin<MODE, JOSEPH>(G, sino, c, s, ia, iu, iv, vi, acc);
        if (!edge[0] && !edge[1]) continue;
        int e0[2], e1[2];
        for (int a = 0; a < 2; ++a) {
            e0[a] = 0;
            e1[a] = -1;
            if (edge[a]) rd_parallel_bins<MODE>(G, c, s, a, e0[a], e1[a]);
        }
        // every bin of the bands not visited above (nor by the band before)
        for (int a = 0; a < 2; ++a)
            for (int iu = e0[a]; iu <= e1[a]; ++iu) {
                if (a == 1 && iu >= e0[0] && iu <= e1[0]) continue;
                for (int iv = 0; iv < nv; ++iv) {
                    if (iu >= u0 && iu <= u1 && iv >= v0 && iv <= v1) continue;
                    rd_bp_bin<MODE, JOSEPH>(G, sino, c, s, ia, iu, iv, vi, acc);
                }
            }
    }
    float r = (float)acc * G.out_scale;
    if (accumulate) r += img[t];
    img[t] = r;
}

static int rd_setup(const vir_rd_geom* g, RDGeo& G) {
    if (!g) return fail_arg("geometry is NULL");
    if (g->mode < 0 || g->mode > 2) return fail_arg("mode must be 0 (par), 1 (fan) or 2 (cone)");
    const bool is3 = g->mode == RD_CONE3;
    G.nx = g->nx; G.ny = g->ny; G.nz = is3 ? g->nz : 1;
    G.nu = g->nu; G.nv = is3 ? g->nv : 1; G.na = g->na;
    G.d = g->d_pix; G.du = g->du; G.dv = is3 ? g->dv : 1.0f; G.su = g->su; G.sv = is3 ? g->sv : 0.0f;
    G.DSO = g->DSO; G.DSD = g->DSD;
    G.compat = 0;
    if (G.nx < 2 || G.ny < 2 || G.nz < 1 || (is3 && G.nz < 2) || G.nu < 1 || G.nv < 1 || G.na < 0)
        return fail_arg("bad sizes (image axes need >= 2 voxels)");
    if (!(G.d > 0.0f) || !(G.du > 0.0f) || !(G.dv > 0.0f)) return fail_arg("spacings must be > 0");
    G.ucen = (float)(0.5 * (double)G.nu - 0.5);
    const int n[3] = {G.nx, G.ny, G.nz};
    for (int a = 0; a < 3; ++a) {
        G.bmin[a] = -G.d * (float)n[a] / 2.0f;
        G.bmax[a] = G.d * (float)n[a] / 2.0f;
        G.c0[a] = -G.d * ((float)n[a] / 2.0f - 0.5f);
    }
    if (g->mode != RD_PAR2) {
        if (!(G.DSO > 0.0f) || !(G.DSD > G.DSO)) return fail_arg("need 0 < DSO < DSD");
        const double hx = 0.5 * G.nx * (double)G.d, hy = 0.5 * G.ny * (double)G.d;
        if (!(std::sqrt(hx * hx + hy * hy) + 2.0 * G.d < G.DSO)) return fail_arg("grid reaches the source");
    }
    return VIR_OK;
}

}  // namespace vir

using namespace vir;

extern "C" {

int vir_rd_fp(const vir_rd_geom* geom, const float* cosv, const float* sinv, const float* img, float* sino,
              float out_scale, uint32_t flags, void* stream) {
    RDGeo G;
    int rc = rd_setup(geom, G);
    if (rc) return rc;
    if (G.na == 0) return VIR_OK;
    if (!cosv || !sinv || !img || !sino) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    G.compat = (flags & VIR_COMPAT) ? 1 : 0;
    const int64_t n = (int64_t)G.na * G.nu * G.nv;
    dim3 grd((unsigned)((n + 255) / 256));
    hipStream_t st = as_stream(stream);
    const int acc = (flags & VIR_ACCUMULATE) ? 1 : 0;
    const bool jo = geom->joseph != 0;
#define VIR_RD_FP(M)                                                                                        \
    do {                                                                                                      \
        if (jo) hipLaunchKernelGGL((k_rd_fp<M, true>), grd, dim3(256), 0, st, G, cosv, sinv, img, sino, acc);  \
        else hipLaunchKernelGGL((k_rd_fp<M, false>), grd, dim3(256), 0, st, G, cosv, sinv, img, sino, acc);   \
    } while (0)
    if (geom->mode == RD_PAR2) VIR_RD_FP(RD_PAR2);
    else if (geom->mode == RD_FAN2) VIR_RD_FP(RD_FAN2);
    else VIR_RD_FP(RD_CONE3);
#undef VIR_RD_FP
    return check_hip(hipGetLastError(), "vir_rd_fp launch");
}

int vir_rd_bp(const vir_rd_geom* geom, const float* cosv, const float* sinv, const float* sino, float* img,
              float out_scale, uint32_t flags, void* stream) {
    RDGeo G;
    int rc = rd_setup(geom, G);
    if (rc) return rc;
    if ((G.na > 0 && (!cosv || !sinv || !sino)) || !img) return fail_arg("NULL pointer argument");
    G.out_scale = out_scale;
    G.compat = (flags & VIR_COMPAT) ? 1 : 0;
    const int64_t n = (int64_t)G.nx * G.ny * G.nz;
    dim3 grd((unsigned)((n + 255) / 256));
    hipStream_t st = as_stream(stream);
    const int acc = (flags & VIR_ACCUMULATE) ? 1 : 0;
    const bool jo = geom->joseph != 0;
#define VIR_RD_BP(M)                                                                                        \
    do {                                                                                                      \
        if (jo) hipLaunchKernelGGL((k_rd_bp<M, true>), grd, dim3(256), 0, st, G, cosv, sinv, sino, img, acc);  \
        else hipLaunchKernelGGL((k_rd_bp<M, false>), grd, dim3(256), 0, st, G, cosv, sinv, sino, img, acc);   \
    } while (0)
    if (geom->mode == RD_PAR2) VIR_RD_BP(RD_PAR2);
    else if (geom->mode == RD_FAN2) VIR_RD_BP(RD_FAN2);
    else VIR_RD_BP(RD_CONE3);
#undef VIR_RD_BP
    return check_hip(hipGetLastError(), "vir_rd_bp launch");
}

}  // extern "C"
