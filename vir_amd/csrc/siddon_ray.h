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

struct GridK {
    int n[3];
    float d[3];
    float o[3];
    float p0[3];      // first plane coordinate per axis
    float pN[3];      // last plane coordinate per axis
    float inv_d[3];   // 1/d when d is a power of two
    int pow2[3];      // division by d may be replaced by multiplication
    float eps_f;      // epsilon as numpy uses it against float32 operands
    double eps_d;     // epsilon against float64 operands
};

// boundspace(n, d, c, float32) (core/core_funcs.py:62-81):
// float32(linspace(-n,n,n+1)) * d / 2 + c, each op rounded to float32.
// Valid for any integer k (virtual planes outside [0, n] are used as guards).
__device__ __forceinline__ float plane_coord(const GridK& g, int a, int k) {
    float lin = (float)(2 * k - g.n[a]);
    float t = lin * g.d[a];
    t = t * 0.5f;
    return t + g.o[a];
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
    float s[3];
    float dST[3];
    float dist;
    float amin, amax;
    int k0[3], k1[3];   // plane slice [k0, k1) per axis; empty when dST == 0
    bool hit;           // passed the entry/exit gate (siddon.py:759-760)
    bool f64;           // merged alpha list is float64 for this ray
};

// Per-ray setup, siddon.py:705-751 and the gate at :759-760.
__device__ __forceinline__ void ray_setup(const GridK& g, const float* __restrict__ src,
                                          const float* __restrict__ trg, Ray& R) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        R.s[a] = src[a];
        R.dST[a] = trg[a] - R.s[a];
    }
    float sq0 = R.dST[0] * R.dST[0];
    float sq1 = R.dST[1] * R.dST[1];
    float sq2 = R.dST[2] * R.dST[2];
    R.dist = __builtin_sqrtf((sq0 + sq1) + sq2);  // correctly rounded (llvm.sqrt.f32)

    float amin = 0.0f, amax = 1.0f;
    bool nmin = false, nmax = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float den = R.dST[a] != 0.0f ? R.dST[a] : g.eps_f;
        float al0 = (g.p0[a] - R.s[a]) / den;
        float alN = (g.pN[a] - R.s[a]) / den;
        float mn = 0.0f, mx = 1.0f;
        if (al0 < alN) { mn = al0; mx = alN; }
        else if (al0 > alN) { mn = alN; mx = al0; }
        nmin |= __builtin_isnan(mn);
        nmax |= __builtin_isnan(mx);
        amin = mn > amin ? mn : amin;
        amax = mx < amax ? mx : amax;
    }
    if (nmin) amin = __builtin_nanf("");
    if (nmax) amax = __builtin_nanf("");
    R.amin = amin;
    R.amax = amax;

    R.f64 = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float dv = R.dST[a];
        R.f64 |= (dv == 0.0f);
        // i0: positive direction uses alpha_min, otherwise alpha_max
        float al_lo = dv > 0.0f ? amin : amax;
        float al_hi = dv > 0.0f ? amax : amin;
        float q0 = ((g.pN[a] - al_lo * dv) - R.s[a]) / g.d[a];
        double v0 = (double)(g.n[a] + 1) - (double)q0;
        double r0 = rint(v0);
        int i0 = to_i32_d(isclose_d(v0, r0, g.eps_d) ? r0 : floor(v0));
        float vN = ((R.s[a] + al_hi * dv) - g.p0[a]) / g.d[a];
        float rN = rintf(vN);
        int iN = to_i32_f(isclose_f(vN, rN, g.eps_f) ? rN : ceilf(vN));
        int ks, ke;
        py_slice(i0, iN, g.n[a] + 1, ks, ke);
        if (dv == 0.0f) { ks = 0; ke = 0; }
        R.k0[a] = ks;
        R.k1[a] = ke;
    }
    R.hit = (amax > amin) && !isclose_f(amax, amin, g.eps_f);
}

__device__ __forceinline__ float plane_alpha(const GridK& g, const Ray& R, int a, int k) {
    return (plane_coord(g, a, k) - R.s[a]) / R.dST[a];
}

// Streaming merge of {alpha_min, alpha_max} and the three plane streams in
// ascending alpha order.  Equal values may be taken in any order: the list
// is a multiset and zero-length pairs are dropped by the reference
// (siddon.py:803-809), so the emitted entries do not depend on tie order.
struct Walker {
    int kn[3];     // next plane index in stream order
    int kstop[3];  // one past the last plane in stream order
    int kdir[3];   // +1 ascending plane index, -1 descending
    float h[3];    // alpha of plane kn (valid while kn != kstop)
    int spec;      // 0: alpha_min pending, 1: alpha_max pending, 2: none
    int remaining; // elements not yet consumed
    float prev;    // last consumed element
    int consumed;  // elements consumed so far, saturating at 2
};

__device__ __forceinline__ int stream_len(const Walker& W, int a) {
    return W.kdir[a] > 0 ? W.kstop[a] - W.kn[a] : W.kn[a] - W.kstop[a];
}

// Position all streams at their first element (walk from the ray start).
__device__ __forceinline__ void walker_begin(const GridK& g, const Ray& R, Walker& W) {
    W.remaining = 2;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (R.dST[a] > 0.0f) { W.kn[a] = R.k0[a]; W.kstop[a] = R.k1[a]; W.kdir[a] = 1; }
        else { W.kn[a] = R.k1[a] - 1; W.kstop[a] = R.k0[a] - 1; W.kdir[a] = -1; }
        if (R.dST[a] == 0.0f) { W.kn[a] = 0; W.kstop[a] = 0; W.kdir[a] = 1; }
        W.remaining += stream_len(W, a);
        W.h[a] = W.kn[a] != W.kstop[a] ? plane_alpha(g, R, a, W.kn[a]) : VIR_INF;
    }
    W.spec = 0;
    W.consumed = 0;
    W.prev = 0.0f;
}

// Consume and return the smallest pending element.  Requires remaining > 0.
__device__ __forceinline__ float walker_pop(const GridK& g, const Ray& R, Walker& W) {
    int src = -1;
    float best = VIR_INF;
    if (W.spec < 2) { best = W.spec == 0 ? R.amin : R.amax; src = 3; }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (W.kn[a] != W.kstop[a] && (src < 0 || W.h[a] < best)) { best = W.h[a]; src = a; }
    }
    if (src == 3) {
        W.spec += 1;
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            if (a == src) {
                W.kn[a] += W.kdir[a];
                W.h[a] = W.kn[a] != W.kstop[a] ? plane_alpha(g, R, a, W.kn[a]) : VIR_INF;
            }
        }
    }
    W.remaining -= 1;
    return best;
}

// Position all streams just after every element <= T (used to enter a ray
// part-way, at the guard plane of a volume brick).  Exact: each candidate
// plane's alpha is recomputed by the same division and compared to T.
__device__ __forceinline__ void walker_seek(const GridK& g, const Ray& R, Walker& W, float T) {
    walker_begin(g, R, W);
    int cnt = 0;
    float pv = -VIR_INF;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (W.kn[a] == W.kstop[a]) continue;
        float pos = R.s[a] + T * R.dST[a];
        float fk = floorf((pos - g.p0[a]) / g.d[a]);
        int lo = R.k0[a], hi = R.k1[a];
        if (W.kdir[a] > 0) {
            // first k in [lo, hi) with alpha(k) > T
            int k = fk < -1.0e9f ? lo : (fk > 1.0e9f ? hi : (int)fk + 1);
            k = k < lo ? lo : (k > hi ? hi : k);
            while (k > lo && plane_alpha(g, R, a, k - 1) > T) --k;
            while (k < hi && plane_alpha(g, R, a, k) <= T) ++k;
            int taken = k - lo;
            if (taken > 0) pv = fmaxf(pv, plane_alpha(g, R, a, k - 1));
            cnt += taken;
            W.kn[a] = k;
        } else {
            // stream runs hi-1 .. lo; next = largest k in [lo, hi) with alpha(k) > T
            int k = fk < -1.0e9f ? lo - 1 : (fk > 1.0e9f ? hi - 1 : (int)fk);
            k = k < lo - 1 ? lo - 1 : (k > hi - 1 ? hi - 1 : k);
            while (k < hi - 1 && plane_alpha(g, R, a, k + 1) > T) ++k;
            while (k >= lo && plane_alpha(g, R, a, k) <= T) --k;
            int taken = (hi - 1) - k;
            if (taken > 0) pv = fmaxf(pv, plane_alpha(g, R, a, k + 1));
            cnt += taken;
            W.kn[a] = k;
        }
        W.remaining -= 0;  // recomputed below
        W.h[a] = W.kn[a] != W.kstop[a] ? plane_alpha(g, R, a, W.kn[a]) : VIR_INF;
    }
    W.spec = 0;
    if (R.amin <= T) { W.spec = 1; cnt += 1; pv = fmaxf(pv, R.amin); }
    if (W.spec == 1 && R.amax <= T) { W.spec = 2; cnt += 1; pv = fmaxf(pv, R.amax); }
    int rem = 2 - W.spec;
#pragma unroll
    for (int a = 0; a < 3; ++a) rem += stream_len(W, a);
    W.remaining = rem;
    W.consumed = cnt > 2 ? 2 : cnt;
    W.prev = pv;
}

// One merged segment (A[i], A[i+1]) -> voxel index and length, as
// siddon.py:796-809.  Returns false when the reference drops the entry
// (length not > 0).  idx holds the truncated index per axis.
__device__ __forceinline__ bool segment(const GridK& g, const Ray& R, float a0, float a1,
                                        int idx[3], float& len) {
    if (R.f64) {
        double A0 = (double)a0, A1 = (double)a1;
        double dA = A1 - A0;
        double m = 0.5 * (A1 + A0);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double pos = ((double)R.s[a] + m * (double)R.dST[a]) - (double)g.p0[a];
            double v = g.pow2[a] ? pos * (double)g.inv_d[a] : pos / (double)g.d[a];
            idx[a] = to_i32_d(v);
        }
        len = (float)((double)R.dist * dA);
    } else {
        float dA = a1 - a0;
        float m = 0.5f * (a1 + a0);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float pos = (R.s[a] + m * R.dST[a]) - g.p0[a];
            float v = g.pow2[a] ? pos * g.inv_d[a] : pos / g.d[a];
            idx[a] = to_i32_f(v);
        }
        len = R.dist * dA;
    }
    return len > 0.0f;
}

// End-trim tests, siddon.py:789-793 (np.isclose default atol=1e-8).
__device__ __forceinline__ bool trim_close(const Ray& R, float x, float y) {
    return R.f64 ? isclose_d((double)x, (double)y, 1e-8) : isclose_f(x, y, 1e-8f);
}

// Walk the merged list from the current walker state and call
// visit(ix, iy, iz, len) for every entry the reference would emit, in
// order.  Stops once the left end of the next pair exceeds a_stop.
template <class Visit>
__device__ __forceinline__ void walk_entries(const GridK& g, const Ray& R, Walker& W,
                                             float a_stop, Visit&& visit) {
    if (W.consumed == 0) {
        if (W.remaining == 0) return;
        W.prev = walker_pop(g, R, W);
        W.consumed = 1;
    }
    while (W.remaining > 0 && !(W.prev > a_stop)) {
        float b = walker_pop(g, R, W);
        bool skip = false;
        if (W.consumed == 1) skip = trim_close(R, W.prev, b);     // drop A[0]
        if (W.remaining == 0) skip |= trim_close(R, b, W.prev);   // drop A[-1]
        int idx[3];
        float len;
        bool keep = segment(g, R, W.prev, b, idx, len);
        if (keep && !skip) visit(idx[0], idx[1], idx[2], len);
        W.prev = b;
        W.consumed = 2;
    }
}

}  // namespace vir
