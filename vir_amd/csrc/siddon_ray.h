// Siddon ray tracing, device side: one ray per lane.
//
// Follows the reference's siddons() (sys_mat/siddon.py:636-833) numerically:
// the per-ray entry/exit alphas, the snapped plane-index bounds and the gate
// are evaluated exactly as numpy evaluates them, then the sorted alpha list is
// produced by merging the three monotone plane streams on the fly instead of
// materialising and sorting it.  Plane alphas are always computed by a fresh
// division (never accumulated), so a stream can be entered at any plane and
// produce the same bits as a walk from the start of the ray.
//
// Numerics (build with -ffp-contract=off; no fast-math):
//  * ray arithmetic is IEEE f32 as numpy does it on float32 arrays;
//  * the start-plane bound i0 is f64, since numpy promotes int64 - float32
//    to float64 (siddon.py:740-741);
//  * rays with any zero direction component carry the merged list in f64,
//    since np.array([]) is float64 and concatenate promotes (siddon.py:776-787);
//  * np.isclose(x, y) = (|x-y| <= atol + rtol*|y| and isfinite(y)) or x == y,
//    in the operand dtype;
//  * astype(int32) truncates; out-of-range values become INT32_MIN.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vir {

#define VIR_INF __builtin_huge_valf()

// Per-axis ray and walker state lives in vector registers: clang vector types
// keep dynamically selected lanes as register selects instead of letting the
// optimiser turn them into indexed (scratch) memory accesses.
typedef float f3v __attribute__((ext_vector_type(3)));
typedef int i3v __attribute__((ext_vector_type(3)));

struct GridK {
    int n[3];
    float d[3];
    float o[3];
    float p0[3];      // first plane coordinate per axis
    float pN[3];      // last plane coordinate per axis
    float inv_d[3];   // RN(1/d)
    int pow2[3];      // d is a power of two: x/d == x*inv_d exactly
    float eps_f;      // epsilon as numpy uses it against float32 operands
    float inv_eps;    // RN(1/eps_f)
    double eps_d;     // epsilon against float64 operands
    // unit = 1: every d is 2^k with |k| <= 4 and every o, p0 and pN is 0
    // or of magnitude in [2^-60, 2^100], so that walks in voxel units
    // (unit_grid, unit_ray) give the same bits for rays where unit_exact
    // holds
    int unit;
};

// Correctly rounded a/b from y = RN(1/b) (Markstein): q = RN(a*y),
// r = a - q*b (exact with FMA), RN(q + r*y) == RN(a/b) provided nothing
// underflows.  Outside the guard band the IEEE division is used instead.
// Checked bit-for-bit against IEEE division on 7e8 random operand pairs.
__device__ __forceinline__ float div_rn(float a, float b, float y) {
    float q = a * y;
    float r = __builtin_fmaf(-q, b, a);
    float qq = __builtin_fmaf(r, y, q);
    float aq = fabsf(q);
    bool ok = (fabsf(a) >= 0x1p-96f && aq >= 0x1p-100f && aq <= 0x1p100f) || a == 0.0f;
    if (__builtin_expect(!ok, 0)) {
        // keep this a real (rarely taken) branch rather than a select
        asm volatile("" ::: "memory");
        qq = a / b;
    }
    return qq;
}

// div_rn without the range check, for operands the caller has shown to lie
// inside the guard band (see planes_unguarded).
__device__ __forceinline__ float div_rn_unguarded(float a, float b, float y) {
    float q = a * y;
    float r = __builtin_fmaf(-q, b, a);
    return __builtin_fmaf(r, y, q);
}

template <bool FAST>
__device__ __forceinline__ float div_rn_t(float a, float b, float y) {
    return FAST ? div_rn_unguarded(a, b, y) : div_rn(a, b, y);
}

__device__ __forceinline__ float fmin_nn(float a, float b) { return b < a ? b : a; }

// boundspace(n, d, c, float32) (core/core_funcs.py:62-81):
// float32(linspace(-n,n,n+1)) * d / 2 + c, each op rounded to float32.
// Valid for any integer k (virtual planes outside [0, n] are used as guards).
__device__ __forceinline__ float plane_at(int n, float d, float o, int k) {
    float lin = (float)(2 * k - n);
    float t = lin * d;
    t = t * 0.5f;
    return t + o;
}

// plane_at from lin = 2 k - n (a float holding that integer exactly).
__device__ __forceinline__ float plane_lin(float d, float o, float lin) {
    float t = lin * d;
    t = t * 0.5f;
    return t + o;
}

// plane_lin in voxel units (d = 1, o in voxel units): l / 2 is exact, so
// one fused multiply-add rounds once, as the add of plane_lin does.
__device__ __forceinline__ float plane_unit(float o, float lin) { return __builtin_fmaf(lin, 0.5f, o); }

__device__ __forceinline__ float plane_coord(const GridK& g, int a, int k) {
    return plane_at(g.n[a], g.d[a], g.o[a], k);
}

__device__ __forceinline__ bool isclose_f(float x, float y, float atol) {
    float tol = atol + 1e-5f * fabsf(y);
    return ((fabsf(x - y) <= tol) && __builtin_isfinite(y)) || (x == y);
}

__device__ __forceinline__ bool isclose_d(double x, double y, double atol) {
    double tol = atol + 1e-5 * fabs(y);
    return ((fabs(x - y) <= tol) && __builtin_isfinite(y)) || (x == y);
}

__device__ __forceinline__ int to_i32_d(double v) {
    return (v > -2147483649.0 && v < 2147483648.0) ? (int)v : INT32_MIN;
}

__device__ __forceinline__ int to_i32_f(float v) {
    return (v > -2147483648.0f && v < 2147483648.0f) ? (int)v : INT32_MIN;
}

// Python slice normalisation seq[start:stop] on a length-L sequence.
__device__ __forceinline__ void py_slice(int start, int stop, int L, int& s, int& e) {
    if (start < 0) { start += L; if (start < 0) start = 0; }
    if (start > L) start = L;
    if (stop < 0) { stop += L; if (stop < 0) stop = 0; }
    if (stop > L) stop = L;
    s = start;
    e = stop < start ? start : stop;
}

struct Ray {
    f3v s;
    f3v dST;
    f3v y;              // RN(1/dST) (unused where dST == 0)
    float dist;
    float amin, amax;
    i3v k0, k1;         // plane slice [k0, k1) per axis; empty when dST == 0
    bool hit;           // passed the entry/exit gate (siddon.py:759-760)
    bool f64;           // merged alpha list is float64 for this ray
};

// Per-ray setup without the plane slices: endpoints, direction, length,
// entry/exit alphas and the gate (siddon.py:705-736, 759-760).  The slices are
// left at the full plane range [0, n+1).
// FAST: the caller has established entry_unguarded for this ray, so the six
// entry/exit divisions return the same bits without div_rn's range check.
// The endpoints come as values (ray_setup_core_v) or are loaded here.
template <bool FAST = false>
__device__ __forceinline__ void ray_setup_core_v(const GridK& g, f3v sp, f3v tp, Ray& R) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        R.s[a] = sp[a];
        R.dST[a] = tp[a] - R.s[a];
        R.y[a] = 1.0f / R.dST[a];
    }
    float sq0 = R.dST[0] * R.dST[0];
    float sq1 = R.dST[1] * R.dST[1];
    float sq2 = R.dST[2] * R.dST[2];
    R.dist = __builtin_sqrtf((sq0 + sq1) + sq2);  // correctly rounded (llvm.sqrt.f32)

    float amin = 0.0f, amax = 1.0f;
    bool nmin = false, nmax = false;
    R.f64 = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        bool z = R.dST[a] == 0.0f;
        R.f64 |= z;
        float den = z ? g.eps_f : R.dST[a];
        float iden = z ? g.inv_eps : R.y[a];
        float al0 = div_rn_t<FAST>(g.p0[a] - R.s[a], den, iden);
        float alN = div_rn_t<FAST>(g.pN[a] - R.s[a], den, iden);
        float mn = 0.0f, mx = 1.0f;
        if (al0 < alN) { mn = al0; mx = alN; }
        else if (al0 > alN) { mn = alN; mx = al0; }
        nmin |= __builtin_isnan(mn);
        nmax |= __builtin_isnan(mx);
        amin = mn > amin ? mn : amin;
        amax = mx < amax ? mx : amax;
        R.k0[a] = 0;
        R.k1[a] = z ? 0 : g.n[a] + 1;
    }
    if (nmin) amin = __builtin_nanf("");
    if (nmax) amax = __builtin_nanf("");
    R.amin = amin;
    R.amax = amax;
    R.hit = (amax > amin) && !isclose_f(amax, amin, g.eps_f);
}

template <bool FAST = false>
__device__ __forceinline__ void ray_setup_core(const GridK& g, const float* __restrict__ src,
                                               const float* __restrict__ trg, Ray& R) {
    f3v sp = {src[0], src[1], src[2]};
    f3v tp = {trg[0], trg[1], trg[2]};
    ray_setup_core_v<FAST>(g, sp, tp, R);
}

// The snapped plane-index bounds i0/iN of siddon.py:740-751 as python slice
// bounds [k0, k1) of the plane arrays.
__device__ __forceinline__ void ray_setup_slices(const GridK& g, Ray& R) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float dv = R.dST[a];
        // i0: positive direction uses alpha_min, otherwise alpha_max
        float al_lo = dv > 0.0f ? R.amin : R.amax;
        float al_hi = dv > 0.0f ? R.amax : R.amin;
        float q0 = div_rn((g.pN[a] - al_lo * dv) - R.s[a], g.d[a], g.inv_d[a]);
        double v0 = (double)(g.n[a] + 1) - (double)q0;
        double r0 = rint(v0);
        int i0 = to_i32_d(isclose_d(v0, r0, g.eps_d) ? r0 : floor(v0));
        float vN = div_rn((R.s[a] + al_hi * dv) - g.p0[a], g.d[a], g.inv_d[a]);
        float rN = rintf(vN);
        int iN = to_i32_f(isclose_f(vN, rN, g.eps_f) ? rN : ceilf(vN));
        int ks, ke;
        py_slice(i0, iN, g.n[a] + 1, ks, ke);
        if (dv == 0.0f) { ks = 0; ke = 0; }
        R.k0[a] = ks;
        R.k1[a] = ke;
    }
}

// Full per-ray setup, siddon.py:705-751 and the gate at :759-760.
template <bool FAST = false>
__device__ __forceinline__ void ray_setup(const GridK& g, const float* __restrict__ src,
                                          const float* __restrict__ trg, Ray& R) {
    ray_setup_core<FAST>(g, src, trg, R);
    ray_setup_slices(g, R);
}

// Per-ray record (kRayRec floats = 80 bytes, 16-byte aligned): the outputs
// of ray_setup -- the per-ray setup and the exact plane slices -- so a
// backprojector that reads it instead of the endpoints traces the same
// elements.
//   [0..2] s  [3..5] dST  [6..8] y = RN(1/dST)  [9] dist  [10] alpha_min
//   [11] alpha_max  [12..14] k0  [15..17] k1 (int bits)  [18] hit (0/1 int bits)
//   [19] 0
constexpr int kRayRec = 20;

__device__ __forceinline__ void ray_store_record(const Ray& R, float* __restrict__ rec) {
    float4* r4 = reinterpret_cast<float4*>(rec);
    r4[0] = make_float4(R.s[0], R.s[1], R.s[2], R.dST[0]);
    r4[1] = make_float4(R.dST[1], R.dST[2], R.y[0], R.y[1]);
    r4[2] = make_float4(R.y[2], R.dist, R.amin, R.amax);
    r4[3] = make_float4(__int_as_float(R.k0[0]), __int_as_float(R.k0[1]), __int_as_float(R.k0[2]),
                        __int_as_float(R.k1[0]));
    r4[4] = make_float4(__int_as_float(R.k1[1]), __int_as_float(R.k1[2]), __int_as_float(R.hit ? 1 : 0), 0.0f);
}

__device__ __forceinline__ void ray_load_record(const float* __restrict__ rec, Ray& R) {
    const float4* r4 = reinterpret_cast<const float4*>(rec);
    const float4 a = r4[0], b = r4[1], c = r4[2], d = r4[3], e = r4[4];
    R.s = f3v{a.x, a.y, a.z};
    R.dST = f3v{a.w, b.x, b.y};
    R.y = f3v{b.z, b.w, c.x};
    R.dist = c.y;
    R.amin = c.z;
    R.amax = c.w;
    R.k0 = i3v{__float_as_int(d.x), __float_as_int(d.y), __float_as_int(d.z)};
    R.k1 = i3v{__float_as_int(d.w), __float_as_int(e.x), __float_as_int(e.y)};
    R.hit = __float_as_int(e.z) != 0;
    R.f64 = a.w == 0.0f || b.x == 0.0f || b.y == 0.0f;
}

// True when the six entry/exit divisions of ray_setup_core, (p0 - s) / den
// and (pN - s) / den per axis, pass div_rn's own range check, so
// ray_setup_core<true> computes the same bits (hit or not).
__device__ __forceinline__ bool entry_unguarded(const GridK& g, const Ray& R) {
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float iden = R.dST[a] == 0.0f ? g.inv_eps : R.y[a];
        float n0 = g.p0[a] - R.s[a], n1 = g.pN[a] - R.s[a];
        float q0 = fabsf(n0 * iden), q1 = fabsf(n1 * iden);
        ok = ok && ((fabsf(n0) >= 0x1p-96f && q0 >= 0x1p-100f && q0 <= 0x1p100f) || n0 == 0.0f);
        ok = ok && ((fabsf(n1) >= 0x1p-96f && q1 >= 0x1p-100f && q1 <= 0x1p100f) || n1 == 0.0f);
    }
    return ok;
}

// True when every plane alpha of the ray, RN((plane - s) / dST) over the
// planes p0..pN of each axis the ray crosses, lies inside div_rn's guard
// band, so the unguarded division returns the same bits.  With gap = the
// smallest non-zero |plane - s| (planes are monotone in k: outside the span
// it is the distance to the span, inside it one of the two planes bracketing
// s, found among k-1..k+2 around the f32 estimate k), every numerator is 0
// or of magnitude >= RN(gap); gap >= 2^-90, gap |1/dST| >= 2^-95 and
// max |plane - s| |1/dST| <= 2^95 keep the quotient inside the band.  Rays
// with a zero direction component (float64 lists) are excluded, as are
// grids whose planes are not distinct in f32.
__device__ __forceinline__ bool planes_unguarded(const GridK& g, const Ray& R) {
    bool ok = !R.f64;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float s = R.s[a];
        float gap;
        if (s < g.p0[a]) {
            gap = g.p0[a] - s;
        } else if (s > g.pN[a]) {
            gap = s - g.pN[a];
        } else {
            int k = (int)floorf((s - g.p0[a]) * g.inv_d[a]);
            gap = VIR_INF;
            float prev = -VIR_INF;
#pragma unroll
            for (int j = -1; j <= 2; ++j) {
                int kk = k + j;
                if (kk < 0 || kk > g.n[a]) continue;
                float pc = plane_coord(g, a, kk);
                ok = ok && pc > prev;  // distinct, increasing planes
                prev = pc;
                float dd = fabsf(pc - s);
                if (dd > 0.0f) gap = fminf(gap, dd);
            }
        }
        float hi = fmaxf(fabsf(g.p0[a] - s), fabsf(g.pN[a] - s));
        float ay = fabsf(R.y[a]);
        ok = ok && gap >= 0x1p-90f && gap * ay >= 0x1p-95f && hi * ay <= 0x1p95f && hi <= 0x1p100f;
    }
    return ok;
}

// True when the alpha window [t0, t1] lies at least four planes inside the
// ray's entry and exit on every axis it crosses.  On axis a the plane
// coordinate is affine in alpha with slope dST/d, so the condition is
// (t0 - alpha_min) |dST| / d >= 4 and (alpha_max - t1) |dST| / d >= 4 for the
// smallest |dST| / d; it is tested against 4.5, which covers the rounding of
// these few float operations.  The snapped slice bounds differ from the
// geometric entry/exit planes by at most one plane, so a walk confined to
// such a window meets the same elements with or without them.  NaN fails.
__device__ __forceinline__ bool window_interior(const GridK& g, const Ray& R, float t0, float t1) {
    float gmin = VIR_INF;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float ga = fabsf(R.dST[a]) * g.inv_d[a];
        gmin = R.dST[a] != 0.0f ? fminf(gmin, ga) : gmin;
    }
    return (t0 - R.amin) * gmin >= 4.5f && (R.amax - t1) * gmin >= 4.5f;
}

template <bool FAST = false>
__device__ __forceinline__ float plane_alpha(const GridK& g, const Ray& R, int a, int k) {
    return div_rn_t<FAST>(plane_coord(g, a, k) - R.s[a], R.dST[a], R.y[a]);
}

template <class T>
__device__ __forceinline__ T sel3(int a, T v0, T v1, T v2) {
    return a == 0 ? v0 : (a == 1 ? v1 : v2);
}

// Streaming merge of {alpha_min, alpha_max} and the three plane streams in
// ascending alpha order.  Equal values may be taken in any order: the list
// is a multiset and zero-length pairs are dropped by the reference
// (siddon.py:803-809), so the emitted entries do not depend on tie order.
// Exhausted streams hold +inf: every plane alpha in a stream lies close to
// [alpha_min, alpha_max] within [0, 1], so +inf never collides with a value.
// Plane indices are held as the float lin = 2 k - n of plane_at (exact
// integers: n <= 2^22 per axis), so advancing a stream is one add and the
// plane coordinate needs no conversion.  Per-axis state is kept in scalar
// members (x/y/z suffixes 0/1/2): a pick among the three axes then stays two
// selects on the comparison masks, where a vector element pick would be
// turned into an index and compares.
struct Walker {
    float ln0, ln1, ln2;        // lin of the next plane in stream order
    float lstop0, lstop1, lstop2;  // lin of one past the last plane in stream order
    float lstep0, lstep1, lstep2;  // +2 ascending plane index, -2 descending
    float h0, h1, h2;           // alpha of the next plane, +inf once exhausted
    int spec;      // 0: alpha_min pending, 1: alpha_max pending, 2: none
    float sp;      // the pending special value (alpha_min, alpha_max or +inf)
    float sp2;     // the special value after it (alpha_max or +inf; FAST pops)
    int remaining; // elements not yet consumed
    float prev;    // last consumed element
    int consumed;  // elements consumed so far, saturating at 2
};

// Per-axis stream state while the setup loops run (vector elements with
// constant indices once unrolled), stored into the walker's scalars at the end.
struct StreamInit {
    f3v ln, lstop, lstep, h;
};

__device__ __forceinline__ void walker_store(Walker& W, const StreamInit& S) {
    W.ln0 = S.ln[0]; W.ln1 = S.ln[1]; W.ln2 = S.ln[2];
    W.lstop0 = S.lstop[0]; W.lstop1 = S.lstop[1]; W.lstop2 = S.lstop[2];
    W.lstep0 = S.lstep[0]; W.lstep1 = S.lstep[1]; W.lstep2 = S.lstep[2];
    W.h0 = S.h[0]; W.h1 = S.h[1]; W.h2 = S.h[2];
}

// Stream a runs over plane indices kn, kn + kdir, ... up to (excluding) kstop.
__device__ __forceinline__ void set_stream(const GridK& g, StreamInit& S, int a, int kn, int kstop, int kdir) {
    S.ln[a] = (float)(2 * kn - g.n[a]);
    S.lstop[a] = (float)(2 * kstop - g.n[a]);
    S.lstep[a] = kdir > 0 ? 2.0f : -2.0f;
}

// Position all streams at their first element (walk from the ray start).
__device__ __forceinline__ void walker_begin(const GridK& g, const Ray& R, Walker& W) {
    W.remaining = 2;
    StreamInit S;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int kn, ks, kd;
        if (R.dST[a] > 0.0f) { kn = R.k0[a]; ks = R.k1[a]; kd = 1; }
        else { kn = R.k1[a] - 1; ks = R.k0[a] - 1; kd = -1; }
        if (R.dST[a] == 0.0f) { kn = 0; ks = 0; kd = 1; }
        set_stream(g, S, a, kn, ks, kd);
        W.remaining += kd > 0 ? ks - kn : kn - ks;
        S.h[a] = kn != ks ? plane_alpha(g, R, a, kn) : VIR_INF;
    }
    walker_store(W, S);
    W.spec = 0;
    W.sp = R.amin;
    W.sp2 = R.amax;
    W.consumed = 0;
    W.prev = 0.0f;
}

// Consume and return the smallest pending element.  Requires remaining > 0.
// Branch-free in the stream choice: one division per step whichever axis
// advances.  ISO: all three axes share (n, d, origin), so the plane
// coordinate needs no per-axis selection.
template <bool ISO, bool FAST, bool UNIT = false>
__device__ __forceinline__ float walker_pop(const GridK& g, const Ray& R, Walker& W) {
    float h0 = W.h0, h1 = W.h1, h2 = W.h2;
    bool c1 = h1 < h0;
    float hm = c1 ? h1 : h0;
    bool c2 = h2 < hm;
    hm = c2 ? h2 : hm;
    W.remaining -= 1;
    // the special value goes first on ties.  Once both are consumed sp is
    // +inf; on the FAST path every plane alpha is finite (|q| <= 2^96), so
    // while elements remain hm < sp.  The general path keeps the count test.
    // Branch-free: the plane step is computed either way and only committed
    // when a plane is taken.
    const bool sp_take = (FAST || W.spec < 2) && !(hm < W.sp);
    // stream choice straight from the two comparisons: c2 -> axis 2,
    // else c1 -> axis 1, else axis 0
    const bool s2 = c2 && !sp_take, s1 = !c2 && c1 && !sp_take, s0 = !c2 && !c1 && !sp_take;
    // (picks among values, not among members: a conditional over lvalues
    // would pick an address)
    const float ln0 = W.ln0, ln1 = W.ln1, ln2 = W.ln2;
    const float st0 = W.lstep0, st1 = W.lstep1, st2 = W.lstep2;
    const float ls0 = W.lstop0, ls1 = W.lstop1, ls2 = W.lstop2;
    const float ws0 = R.s[0], ws1 = R.s[1], ws2 = R.s[2];
    const float we0 = R.dST[0], we1 = R.dST[1], we2 = R.dST[2];
    const float wy0 = R.y[0], wy1 = R.y[1], wy2 = R.y[2];
    float l = (c2 ? ln2 : (c1 ? ln1 : ln0)) + (c2 ? st2 : (c1 ? st1 : st0));
    float ls = c2 ? ls2 : (c1 ? ls1 : ls0);
    float pc;
    if (UNIT) {
        pc = plane_unit(ISO ? g.o[0] : (c2 ? g.o[2] : (c1 ? g.o[1] : g.o[0])), l);
    } else if (ISO) {
        pc = plane_lin(g.d[0], g.o[0], l);
    } else {
        pc = plane_lin(c2 ? g.d[2] : (c1 ? g.d[1] : g.d[0]), c2 ? g.o[2] : (c1 ? g.o[1] : g.o[0]), l);
    }
    float sa = c2 ? ws2 : (c1 ? ws1 : ws0);
    float da = c2 ? we2 : (c1 ? we1 : we0);
    float ya = c2 ? wy2 : (c1 ? wy1 : wy0);
    float hn = div_rn_t<FAST>(pc - sa, da, ya);
    hn = l != ls ? hn : VIR_INF;
    W.ln0 = s0 ? l : ln0;
    W.ln1 = s1 ? l : ln1;
    W.ln2 = s2 ? l : ln2;
    W.h0 = s0 ? hn : h0;
    W.h1 = s1 ? hn : h1;
    W.h2 = s2 ? hn : h2;
    const float sv = W.sp;
    W.spec += sp_take ? 1 : 0;
    if constexpr (FAST) {
        W.sp = sp_take ? W.sp2 : sv;
        W.sp2 = sp_take ? VIR_INF : W.sp2;
    } else {
        W.sp = sp_take ? (W.spec == 1 ? R.amax : VIR_INF) : sv;
    }
    return sp_take ? sv : hm;
}

// Position all streams just after every element <= T (used to enter a ray
// part-way, at the guard plane of a volume brick).  Exact: each candidate
// plane's alpha is recomputed by the same division and compared to T.
template <bool FAST = false>
__device__ __forceinline__ void walker_seek(const GridK& g, const Ray& R, Walker& W, float T) {
    W.remaining = 2;
    int cnt = 0;
    float pv = -VIR_INF;
    StreamInit S;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int lo = R.k0[a], hi = R.k1[a];
        if (R.dST[a] == 0.0f || lo >= hi) {
            set_stream(g, S, a, 0, 0, 1);
            S.h[a] = VIR_INF;
            continue;
        }
        float pos = R.s[a] + T * R.dST[a];
        float fk = floorf((pos - g.p0[a]) * g.inv_d[a]);
        int k;
        float hk;
        if (R.dST[a] > 0.0f) {
            // first k in [lo, hi) with alpha(k) > T
            k = fk < -1.0e9f ? lo : (fk > 1.0e9f ? hi : (int)fk + 1);
            k = k < lo ? lo : (k > hi ? hi : k);
            // am1: alpha of plane k - 1 (the last one taken) while k > lo
            float am1 = k > lo ? plane_alpha<FAST>(g, R, a, k - 1) : -VIR_INF;
            while (k > lo && am1 > T) {
                --k;
                am1 = k > lo ? plane_alpha<FAST>(g, R, a, k - 1) : -VIR_INF;
            }
            hk = k < hi ? plane_alpha<FAST>(g, R, a, k) : VIR_INF;
            while (k < hi && hk <= T) {
                am1 = hk;
                ++k;
                hk = k < hi ? plane_alpha<FAST>(g, R, a, k) : VIR_INF;
            }
            int taken = k - lo;
            if (taken > 0) pv = fmaxf(pv, am1);
            cnt += taken;
            set_stream(g, S, a, k, hi, 1);
            W.remaining += hi - k;
        } else {
            // stream runs hi-1 .. lo; next = largest k in [lo, hi) with alpha(k) > T
            k = fk < -1.0e9f ? lo - 1 : (fk > 1.0e9f ? hi - 1 : (int)fk);
            k = k < lo - 1 ? lo - 1 : (k > hi - 1 ? hi - 1 : k);
            // ap1: alpha of plane k + 1 (the last one taken) while k < hi - 1
            float ap1 = k < hi - 1 ? plane_alpha<FAST>(g, R, a, k + 1) : -VIR_INF;
            while (k < hi - 1 && ap1 > T) {
                ++k;
                ap1 = k < hi - 1 ? plane_alpha<FAST>(g, R, a, k + 1) : -VIR_INF;
            }
            hk = k >= lo ? plane_alpha<FAST>(g, R, a, k) : VIR_INF;
            while (k >= lo && hk <= T) {
                ap1 = hk;
                --k;
                hk = k >= lo ? plane_alpha<FAST>(g, R, a, k) : VIR_INF;
            }
            int taken = (hi - 1) - k;
            if (taken > 0) pv = fmaxf(pv, ap1);
            cnt += taken;
            set_stream(g, S, a, k, lo - 1, -1);
            W.remaining += k - (lo - 1);
        }
        S.h[a] = hk;
    }
    walker_store(W, S);
    W.spec = 0;
    if (R.amin <= T) { W.spec = 1; cnt += 1; pv = fmaxf(pv, R.amin); }
    if (W.spec == 1 && R.amax <= T) { W.spec = 2; cnt += 1; pv = fmaxf(pv, R.amax); }
    W.sp = W.spec == 0 ? R.amin : (W.spec == 1 ? R.amax : VIR_INF);
    W.sp2 = W.spec == 0 ? R.amax : VIR_INF;
    W.remaining -= W.spec;
    W.consumed = cnt > 2 ? 2 : cnt;
    W.prev = pv;
}

// One merged segment (A[i], A[i+1]) -> voxel index and length, as
// siddon.py:796-809.  Returns false when the reference drops the entry
// (length not > 0).  idx holds the truncated index per axis.  P2: every
// voxel size is a power of two, so dividing by it is a multiplication.
template <bool P2, bool EXACT, bool FAST = false, bool UNIT = false>
__device__ __forceinline__ bool segment(const GridK& g, const Ray& R, float a0, float a1,
                                        int idx[3], float& len) {
    // idx: truncated index per axis.  EXACT maps values outside int32 to
    // INT32_MIN as numpy does on x86-64 (needed only where raw indices are
    // returned); otherwise they saturate, which is equally outside the grid
    // (NaN cannot occur: rays with NaN endpoints fail the gate).
    if (!FAST && R.f64) {
        double A0 = (double)a0, A1 = (double)a1;
        double dA = A1 - A0;
        double m = 0.5 * (A1 + A0);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double pos = ((double)R.s[a] + m * (double)R.dST[a]) - (double)g.p0[a];
            double v = P2 ? pos * (double)g.inv_d[a] : pos / (double)g.d[a];
            idx[a] = EXACT ? to_i32_d(v) : (int)v;
        }
        len = (float)((double)R.dist * dA);
    } else {
        float dA = a1 - a0;
        float m = 0.5f * (a1 + a0);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float pos = (R.s[a] + m * R.dST[a]) - g.p0[a];
            // (in voxel units d = 1: the position is the index coordinate)
            float v = UNIT ? pos : (P2 ? pos * g.inv_d[a] : div_rn(pos, g.d[a], g.inv_d[a]));
            idx[a] = EXACT ? to_i32_f(v) : (int)v;
        }
        len = R.dist * dA;
    }
    return len > 0.0f;
}

// End-trim tests, siddon.py:789-793 (np.isclose default atol=1e-8).
template <bool FAST = false>
__device__ __forceinline__ bool trim_close(const Ray& R, float x, float y) {
    return (!FAST && R.f64) ? isclose_d((double)x, (double)y, 1e-8) : isclose_f(x, y, 1e-8f);
}

// Walk the merged list from the current walker state and call
// visit(ix, iy, iz, len) for every entry the reference would emit, in
// order.  Stops once the left end of the next pair exceeds a_stop.  The
// first pair (A[0], A[1]) and the last pair are peeled off the loop: only
// they are subject to the end trims.
// FAST: every lane's ray satisfies planes_unguarded (float32 list, divisions
// inside the guard band); see walk_dispatch.
// ALL: visit(ix, iy, iz, len, keep) is called for every pair, keep false for
// the pairs the reference drops (branch-free visitors).
template <bool ALL, class Visit>
__device__ __forceinline__ void visit_pair(Visit& visit, const int idx[3], float len, bool keep) {
    if constexpr (ALL) visit(idx[0], idx[1], idx[2], len, keep);
    else if (keep) visit(idx[0], idx[1], idx[2], len);
}

template <bool P2, bool ISO, bool EXACT = false, bool FAST = false, bool ALL = false, bool UNIT = false,
          class Visit>
__device__ __forceinline__ void walk_entries(const GridK& g, const Ray& R, Walker& W,
                                             float a_stop, Visit&& visit) {
    int idx[3];
    float len;
    if (W.consumed == 0) {
        if (W.remaining == 0) return;
        W.prev = walker_pop<ISO, FAST, UNIT>(g, R, W);
        W.consumed = 1;
    }
    if (W.consumed == 1) {
        if (W.remaining == 0 || W.prev > a_stop) return;
        float b = walker_pop<ISO, FAST, UNIT>(g, R, W);
        bool skip = trim_close<FAST>(R, W.prev, b);                   // drop A[0]
        if (W.remaining == 0) skip |= trim_close<FAST>(R, b, W.prev);  // drop A[-1]
        bool keep = segment<P2, EXACT, FAST, UNIT>(g, R, W.prev, b, idx, len);
        visit_pair<ALL>(visit, idx, len, keep && !skip);
        W.prev = b;
        W.consumed = 2;
    }
    while (W.remaining > 1 && !(W.prev > a_stop)) {
        float b = walker_pop<ISO, FAST, UNIT>(g, R, W);
        bool keep = segment<P2, EXACT, FAST, UNIT>(g, R, W.prev, b, idx, len);
        visit_pair<ALL>(visit, idx, len, keep);
        W.prev = b;
    }
    if (W.remaining == 1 && !(W.prev > a_stop)) {
        float b = walker_pop<ISO, FAST, UNIT>(g, R, W);
        bool skip = trim_close<FAST>(R, b, W.prev);                    // drop A[-1]
        bool keep = segment<P2, EXACT, FAST, UNIT>(g, R, W.prev, b, idx, len);
        visit_pair<ALL>(visit, idx, len, keep && !skip);
        W.prev = b;
    }
}

// True when neither special value (alpha_min, alpha_max) can be popped by
// a walk over [t0, t1] entered with walker_seek(t0), and no end trim
// applies: along the axis crossing the most planes per unit alpha, at least
// 2.5 planes lie between alpha_min and t0 and 3.5 between t1 and
// alpha_max, so walker_seek consumes alpha_min and at least one plane, and
// the element that ends the walk (the first past t1) is a plane before
// alpha_max (the exact slices differ from the geometric entry and exit by
// at most one plane; the half plane covers float rounding).  NaN fails.
__device__ __forceinline__ bool window_mid(const GridK& g, const Ray& R, float t0, float t1) {
    float gmax = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) gmax = fmaxf(gmax, fabsf(R.dST[a]) * g.inv_d[a]);
    return (t0 - R.amin) * gmax >= 2.5f && (R.amax - t1) * gmax >= 3.5f;
}

// Walks in voxel units.  On a g.unit grid every d is 2^k, and dividing every
// coordinate of axis a by d_a commutes with each rounding of the walk -- the
// plane coordinates (l d/2 + o)/d = l/2 + o/d, the plane alphas (the ratio
// of two scaled coordinates, its Markstein steps scaled by d and 1/d), the
// midpoint positions and the index coordinates pos/d -- as long as every
// operand and result stays normal before and after scaling; alphas, and so
// the merged list, the trims and the lengths (dist stays in world units),
// do not change.  unit_exact states that condition for a ray: on a g.unit
// grid (|k| <= 4), each s is 0 or of magnitude in [2^-60, 2^100] and each
// dST is 0 or has |dST| alpha_min >= 2^-59 and |dST| <= 2^100; with the
// grid's coordinates likewise, every product m dST (m >= alpha_min) is at
// least 2^-59 and every sum or difference of such values 0 or a multiple of
// 2^-83, so nothing comes within 2^16 of the normal range's bottom (the
// division's remainder, a multiple of about 2^-46 of its numerator, stays
// on the subnormal grid), nor near its top.
__host__ __device__ inline GridK unit_grid(const GridK& g) {
    GridK u = g;
    for (int a = 0; a < 3; ++a) {
        u.d[a] = 1.0f;
        u.inv_d[a] = 1.0f;
        u.o[a] = g.o[a] * g.inv_d[a];
        u.p0[a] = g.p0[a] * g.inv_d[a];
        u.pN[a] = g.pN[a] * g.inv_d[a];
        u.pow2[a] = 1;
    }
    return u;
}

__device__ __forceinline__ bool unit_exact(const GridK& g, const Ray& R) {
    bool ok = g.unit != 0 && R.amin > 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float sa = fabsf(R.s[a]), ea = fabsf(R.dST[a]);
        ok = ok && (sa == 0.0f || (sa >= 0x1p-60f && sa <= 0x1p100f));
        ok = ok && (ea == 0.0f || (ea * R.amin >= 0x1p-59f && ea <= 0x1p100f));
    }
    return ok;
}

// The ray's coordinates in voxel units: s / d, dST / d and y d = RN(1/(dST/d)).
__device__ __forceinline__ void unit_ray(const GridK& g, Ray& R) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        R.s[a] = R.s[a] * g.inv_d[a];
        R.dST[a] = R.dST[a] * g.inv_d[a];
        R.y[a] = R.y[a] * g.d[a];
    }
}

// walker_pop<ISO, true> for a walk over a window_mid window: alpha_max is
// never the smallest pending element and the element count never reaches
// its end, so the special-value pick and the count are left out; streams
// that run out still yield +inf.  Same elements as walker_pop there.
template <bool ISO, bool UNIT = false>
__device__ __forceinline__ float walker_pop_mid(const GridK& g, const Ray& R, Walker& W) {
    const float h0 = W.h0, h1 = W.h1, h2 = W.h2;
    const bool c1 = h1 < h0;
    float hm = c1 ? h1 : h0;
    const bool c2 = h2 < hm;
    hm = c2 ? h2 : hm;
    const bool s2 = c2, s1 = !c2 && c1, s0 = !c2 && !c1;
    // (picks among values, not among members: a conditional over lvalues
    // would pick an address)
    const float ln0 = W.ln0, ln1 = W.ln1, ln2 = W.ln2;
    const float st0 = W.lstep0, st1 = W.lstep1, st2 = W.lstep2;
    const float ls0 = W.lstop0, ls1 = W.lstop1, ls2 = W.lstop2;
    const float l = (c2 ? ln2 : (c1 ? ln1 : ln0)) + (c2 ? st2 : (c1 ? st1 : st0));
    const float ls = c2 ? ls2 : (c1 ? ls1 : ls0);
    float pc;
    if (UNIT) {
        pc = plane_unit(ISO ? g.o[0] : (c2 ? g.o[2] : (c1 ? g.o[1] : g.o[0])), l);
    } else if (ISO) {
        pc = plane_lin(g.d[0], g.o[0], l);
    } else {
        pc = plane_lin(c2 ? g.d[2] : (c1 ? g.d[1] : g.d[0]), c2 ? g.o[2] : (c1 ? g.o[1] : g.o[0]), l);
    }
    const float ws0 = R.s[0], ws1 = R.s[1], ws2 = R.s[2];
    const float we0 = R.dST[0], we1 = R.dST[1], we2 = R.dST[2];
    const float wy0 = R.y[0], wy1 = R.y[1], wy2 = R.y[2];
    const float sa = c2 ? ws2 : (c1 ? ws1 : ws0);
    const float da = c2 ? we2 : (c1 ? we1 : we0);
    const float ya = c2 ? wy2 : (c1 ? wy1 : wy0);
    float hn = div_rn_unguarded(pc - sa, da, ya);
    hn = l != ls ? hn : VIR_INF;
    W.ln0 = s0 ? l : ln0;
    W.ln1 = s1 ? l : ln1;
    W.ln2 = s2 ? l : ln2;
    W.h0 = s0 ? hn : h0;
    W.h1 = s1 ? hn : h1;
    W.h2 = s2 ? hn : h2;
    return hm;
}

// walk_entries<P2, ISO, false, true, ALL> over a window_mid window
// [W.prev, a_stop], with the walker past alpha_min and its first pair:
// every pair is a middle pair (no end trims) and the walk stops on a_stop
// alone.  COUNT keeps W.remaining exact, for a walk_entries that carries on
// from where this one stops.
template <bool P2, bool ISO, bool ALL = false, bool COUNT = false, bool UNIT = false, class Visit>
__device__ __forceinline__ void walk_mid(const GridK& g, const Ray& R, Walker& W, float a_stop, Visit&& visit) {
    int idx[3];
    float len;
    while (!(W.prev > a_stop)) {
        const float b = walker_pop_mid<ISO, UNIT>(g, R, W);
        if (COUNT) W.remaining -= 1;
        const bool keep = segment<P2, false, true, UNIT>(g, R, W.prev, b, idx, len);
        visit_pair<ALL>(visit, idx, len, keep);
        W.prev = b;
    }
}

template <bool B>
struct Fast {
    static constexpr bool value = B;
};

// Call body(Fast<true>{}) when every active lane's ray passes
// planes_unguarded (the common case), else body(Fast<false>{}).  The choice
// is wave-uniform, and both instantiations produce the same bits.  UNG: the
// caller has established that every ray passes (vir_siddon_count_guarded),
// so only the unguarded walk is compiled -- the kernel then needs fewer
// registers.
template <bool UNG = false, class Body>
__device__ __forceinline__ void walk_dispatch(const GridK& g, const Ray& R, Body&& body) {
    if constexpr (UNG) {
        body(Fast<true>{});
    } else {
        if (__ballot(!planes_unguarded(g, R)) == 0) {
            body(Fast<true>{});
        } else {
            body(Fast<false>{});
        }
    }
}

}  // namespace vir
