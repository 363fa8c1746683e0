/*
 * ORACLE -- test infrastructure only.  Nothing in the product (vir_amd/) may
 * link, load or call this file; only tests/, the smoke check and the CPU
 * baseline leg of bench.py use it, and only as the checker.
 *
 * Plain-C restatement of the reference's Siddon system-matrix builder
 * `siddons()` in sys_mat/siddon.py:636-833 (grid planes from
 * core/core_funcs.py:62-81 and :235-265).  It follows the reference's
 * algorithm as written there -- per ray: compute ALL plane-crossing alphas,
 * sort them, trim the ends, then emit one (voxel, length) entry per
 * consecutive pair -- and reproduces numpy's dtype behaviour bit for bit:
 *
 *   - every array op on the f32 ray data is an IEEE single op, unfused
 *     (build with -ffp-contract=off);
 *   - the start-plane index i0 is evaluated in double, because numpy
 *     promotes int64 (nPixels+1) minus a float32 array to float64
 *     (siddon.py:740-741);
 *   - the merged alpha list is float64 for every ray that has a zero
 *     direction component, because np.array([]) is float64 and
 *     np.concatenate promotes (siddon.py:776-787);
 *   - np.isclose is |x-y| <= atol + rtol*|y| (& isfinite(y)) | x==y,
 *     evaluated in the dtype of its operands (numpy/_core/numeric.py);
 *   - np.linalg.norm(axis=-1) of an (...,3) float32 array is
 *     sqrtf((x*x + y*y) + z*z) (checked against numpy in tests);
 *   - astype(int32) truncates toward zero; values outside int32 become
 *     INT32_MIN as on x86-64.
 *
 * This file is deliberately a different algorithm from the GPU kernels
 * (which merge three monotone plane streams on the fly): sort-then-scan
 * versus streaming merge, so agreement between the two is evidence.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <string.h>

#define RTOL_F 1e-5f
#define RTOL_D 1e-5

/* ORACLE_PLANT selects a deliberately wrong variant, compiled only by
 * tests/test_oracle_siddon.py to show the golden vectors catch it:
 *   1 no end trimming     2 plain floor/ceil index bounds (no snapping)
 *   3 float32 list always 4 round-half-away midpoint index */
#ifndef ORACLE_PLANT
#define ORACLE_PLANT 0
#endif

static int cmp_double(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    /* numpy sorts NaN last */
    int xn = isnan(x), yn = isnan(y);
    if (xn || yn) return xn - yn;
    return (x > y) - (x < y);
}

static int32_t to_i32(double v) {
    if (!(v > -2147483649.0 && v < 2147483648.0)) return INT32_MIN;
    return (int32_t)v;
}

/* np.isclose on float32 operands with python-float atol/rtol */
static int isclose_f(float x, float y, float atol) {
    float tol = atol + RTOL_F * fabsf(y);
    return ((fabsf(x - y) <= tol) && isfinite(y)) || (x == y);
}

/* np.isclose on float64 operands */
static int isclose_d(double x, double y, double atol) {
    double tol = atol + RTOL_D * fabs(y);
    return ((fabs(x - y) <= tol) && isfinite(y)) || (x == y);
}

/* Python slice normalisation for seq[start:stop] on a sequence of length L */
static void py_slice(int64_t start, int64_t stop, int64_t L, int64_t *s, int64_t *e) {
    if (start < 0) { start += L; if (start < 0) start = 0; }
    if (start > L) start = L;
    if (stop < 0) { stop += L; if (stop < 0) stop = 0; }
    if (stop > L) stop = L;
    *s = start;
    *e = stop < start ? start : stop;
}

/* Grid plane coordinates: boundspace(n, d, c, float32) core_funcs.py:62-81
 * = float32(linspace(-n, n, n+1)) * d / 2.0 + c, each op in float32. */
void oracle_planes(int64_t n, float d, float o, float *Xb) {
    for (int64_t k = 0; k <= n; ++k) {
        float lin = (float)(2 * k - n);
        float t = lin * d;
        t = t / 2.0f;
        Xb[k] = t + o;
    }
}

/*
 * Trace nrays rays.  Outputs per ray r the entries [off[r], off[r+1]) of
 * (ix, iy, iz, len).  flags[r] bit 0 = the ray produced an index outside
 * the grid (the reference raises ValueError in ravel/flat modes); bit 1 =
 * the ray passed the entry/exit gate (the reference then stores a list,
 * possibly empty, instead of None).
 * Returns total entries, or -1 if cap would be exceeded.
 */
int64_t oracle_siddons(int64_t nrays, const float *src, const float *trg,
                       const int64_t n[3], const float d[3], const float o[3],
                       double epsilon,
                       int32_t *ox, int32_t *oy, int32_t *oz, float *olen,
                       int64_t *off, uint8_t *bad, int64_t cap) {
    float *planes[3];
    for (int a = 0; a < 3; ++a) {
        planes[a] = (float *)malloc(sizeof(float) * (size_t)(n[a] + 1));
        oracle_planes(n[a], d[a], o[a], planes[a]);
    }
    const float eps_f = (float)epsilon;
    float p0[3], pN[3];
    for (int a = 0; a < 3; ++a) { p0[a] = planes[a][0]; pN[a] = planes[a][n[a]]; }

    int64_t maxlen = 2 + (n[0] + 1) + (n[1] + 1) + (n[2] + 1);
    double *alist = (double *)malloc(sizeof(double) * (size_t)maxlen);
    float *flist = (float *)malloc(sizeof(float) * (size_t)maxlen);
    int64_t total = 0;
    off[0] = 0;

    for (int64_t r = 0; r < nrays; ++r) {
        const float *s = src + 3 * r, *t = trg + 3 * r;
        float dST[3];
        for (int a = 0; a < 3; ++a) dST[a] = t[a] - s[a];
        float sq0 = dST[0] * dST[0], sq1 = dST[1] * dST[1], sq2 = dST[2] * dST[2];
        float distance = sqrtf((sq0 + sq1) + sq2);
        bad[r] = 0;

        /* siddon.py:717-720 */
        float alpha0[3], alphaN[3];
        for (int a = 0; a < 3; ++a) {
            float num0 = p0[a] - s[a], numN = pN[a] - s[a];
            if (dST[a] != 0.0f) { alpha0[a] = num0 / dST[a]; alphaN[a] = numN / dST[a]; }
            else { alpha0[a] = num0 / eps_f; alphaN[a] = numN / eps_f; }
        }
        /* siddon.py:725-736 */
        float amin = 0.0f, amax = 1.0f;
        int nan_min = 0, nan_max = 0;
        for (int a = 0; a < 3; ++a) {
            float mn = 0.0f, mx = 1.0f;
            if (alpha0[a] < alphaN[a]) { mn = alpha0[a]; mx = alphaN[a]; }
            else if (alpha0[a] > alphaN[a]) { mn = alphaN[a]; mx = alpha0[a]; }
            if (isnan(mn)) nan_min = 1;
            if (isnan(mx)) nan_max = 1;
            if (mn > amin) amin = mn;
            if (mx < amax) amax = mx;
        }
        if (nan_min) amin = NAN;
        if (nan_max) amax = NAN;

        /* siddon.py:740-751 */
        int32_t i0[3], iN[3];
        for (int a = 0; a < 3; ++a) {
            float q0p = ((pN[a] - amin * dST[a]) - s[a]) / d[a];
            float q0n = ((pN[a] - amax * dST[a]) - s[a]) / d[a];
            double v0p = (double)(n[a] + 1) - (double)q0p;
            double v0n = (double)(n[a] + 1) - (double)q0n;
            float vNp = ((s[a] + amax * dST[a]) - p0[a]) / d[a];
            float vNn = ((s[a] + amin * dST[a]) - p0[a]) / d[a];
            double r0p = rint(v0p), r0n = rint(v0n);
            float rNp = rintf(vNp), rNn = rintf(vNn);
            int snap = ORACLE_PLANT != 2;
            int32_t i0p = to_i32(snap && isclose_d(v0p, r0p, epsilon) ? r0p : floor(v0p));
            int32_t i0n = to_i32(snap && isclose_d(v0n, r0n, epsilon) ? r0n : floor(v0n));
            int32_t iNp = to_i32(snap && isclose_f(vNp, rNp, eps_f) ? (double)rNp : (double)ceilf(vNp));
            int32_t iNn = to_i32(snap && isclose_f(vNn, rNn, eps_f) ? (double)rNn : (double)ceilf(vNn));
            i0[a] = dST[a] > 0.0f ? i0p : i0n;
            iN[a] = dST[a] > 0.0f ? iNp : iNn;
        }

        /* gate, siddon.py:759-760 */
        int hit = (amax > amin) && !isclose_f(amax, amin, eps_f);
        if (hit) {
            bad[r] |= 2;
            int f64mode = (dST[0] == 0.0f) || (dST[1] == 0.0f) || (dST[2] == 0.0f);
            if (ORACLE_PLANT == 3) f64mode = 0;
            int64_t L = 0;
            if (f64mode) { alist[L++] = amin; alist[L++] = amax; }
            else { flist[L++] = amin; flist[L++] = amax; }
            for (int a = 0; a < 3; ++a) {
                if (dST[a] == 0.0f) continue;
                int64_t ks, ke;
                py_slice(i0[a], iN[a], n[a] + 1, &ks, &ke);
                for (int64_t k = ks; k < ke; ++k) {
                    float al = (planes[a][k] - s[a]) / dST[a];
                    if (f64mode) alist[L++] = al; else flist[L++] = al;
                }
            }
            /* sort ascending; values are identical bits whether held as
             * float or double, so the order of the multiset is unique up
             * to equal values */
            if (!f64mode) for (int64_t i = 0; i < L; ++i) alist[i] = flist[i];
            qsort(alist, (size_t)L, sizeof(double), cmp_double);
            int64_t b = 0, e = L;
            /* trim, siddon.py:789-793 (default atol=1e-8) */
            if (ORACLE_PLANT == 1) {
                /* no trimming */
            } else if (f64mode) {
                if (isclose_d(alist[0], alist[1], 1e-8)) b = 1;
                if (isclose_d(alist[e - 1], alist[e - 2], 1e-8)) e -= 1;
            } else {
                if (isclose_f((float)alist[0], (float)alist[1], 1e-8f)) b = 1;
                if (isclose_f((float)alist[e - 1], (float)alist[e - 2], 1e-8f)) e -= 1;
            }
            for (int64_t i = b; i + 1 < e; ++i) {
                int32_t idx[3];
                float len;
                if (f64mode) {
                    double A0 = alist[i], A1 = alist[i + 1];
                    double dA = A1 - A0;
                    double m = 0.5 * (A1 + A0);
                    for (int a = 0; a < 3; ++a) {
                        double pos = ((double)s[a] + m * (double)dST[a]) - (double)p0[a];
                        idx[a] = ORACLE_PLANT == 4 ? to_i32(round(pos / (double)d[a]))
                                                   : to_i32(pos / (double)d[a]);
                    }
                    len = (float)((double)distance * dA);
                } else {
                    float A0 = (float)alist[i], A1 = (float)alist[i + 1];
                    float dA = A1 - A0;
                    float m = 0.5f * (A1 + A0);
                    for (int a = 0; a < 3; ++a) {
                        float pos = (s[a] + m * dST[a]) - p0[a];
                        idx[a] = ORACLE_PLANT == 4 ? to_i32(round((double)(pos / d[a])))
                                                   : to_i32((double)(pos / d[a]));
                    }
                    len = distance * dA;
                }
                if (!(len > 0.0f)) continue;
                if (total >= cap) {
                    for (int a = 0; a < 3; ++a) free(planes[a]);
                    free(alist); free(flist);
                    return -1;
                }
                for (int a = 0; a < 3; ++a)
                    if (idx[a] < 0 || idx[a] >= n[a]) bad[r] |= 1;
                ox[total] = idx[0]; oy[total] = idx[1]; oz[total] = idx[2];
                olen[total] = len;
                ++total;
            }
        }
        off[r + 1] = total;
    }
    for (int a = 0; a < 3; ++a) free(planes[a]);
    free(alist); free(flist);
    return total;
}

/* Upper bound on entries per ray (planes + 2 bounds - 1 segments). */
int64_t oracle_siddons_maxlen(const int64_t n[3]) {
    return (n[0] + 1) + (n[1] + 1) + (n[2] + 1) + 1;
}

/*
 * CPU baseline of the reference's CPU path for one batch of rays: the
 * siddons() trace (this restatement), the CSR forward projection with a
 * float64 accumulator (projection/proj_numba.py:11-32, parallel over rays
 * like fast_forward_project_p) and the CSR backprojection (float32,
 * projection/projection_c.c:176-219), over `nthreads` OpenMP threads.  The
 * backprojection goes into per-thread volumes summed at the end.  Rays are
 * traced in chunks so each call of oracle_siddons amortises its set-up.
 * Returns the number of (ray, voxel) entries, or -1 on allocation failure.
 */
int64_t oracle_fp_bp(int64_t nrays, const float *src, const float *trg,
                     const int64_t n[3], const float d[3], const float o[3], double epsilon,
                     const float *vol, float *sino, float *back, int nthreads) {
    const int64_t nvox = n[0] * n[1] * n[2];
    const int64_t maxlen = oracle_siddons_maxlen(n);
    const int64_t chunk = 256;
    if (nthreads < 1) nthreads = 1;
    float *part = (float *)calloc((size_t)nthreads * (size_t)nvox, sizeof(float));
    if (!part) return -1;
    int64_t total = 0;
    int failed = 0;
#pragma omp parallel num_threads(nthreads) reduction(+ : total)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        float *mine = part + (size_t)tid * (size_t)nvox;
        const int64_t cap = chunk * maxlen;
        int32_t *x = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
        int32_t *y = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
        int32_t *z = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
        float *ln = (float *)malloc(sizeof(float) * (size_t)cap);
        int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (size_t)(chunk + 1));
        uint8_t *bad = (uint8_t *)malloc((size_t)chunk);
        if (!x || !y || !z || !ln || !off || !bad) {
#pragma omp atomic write
            failed = 1;
        } else {
#pragma omp for schedule(dynamic, 1)
            for (int64_t r0 = 0; r0 < nrays; r0 += chunk) {
                const int64_t cnt = r0 + chunk <= nrays ? chunk : nrays - r0;
                int64_t m = oracle_siddons(cnt, src + 3 * r0, trg + 3 * r0, n, d, o, epsilon, x, y, z, ln,
                                           off, bad, cap);
                if (m < 0) continue;
                total += m;
                for (int64_t k = 0; k < cnt; ++k) {
                    double s = 0.0;
                    for (int64_t e = off[k]; e < off[k + 1]; ++e) {
                        if (x[e] < 0 || y[e] < 0 || z[e] < 0 || x[e] >= n[0] || y[e] >= n[1] || z[e] >= n[2])
                            continue;
                        const int64_t vi = ((int64_t)x[e] * n[1] + y[e]) * n[2] + z[e];
                        s += (double)vol[vi] * (double)ln[e];
                    }
                    const float sv = (float)s;
                    sino[r0 + k] = sv;
                    for (int64_t e = off[k]; e < off[k + 1]; ++e) {
                        if (x[e] < 0 || y[e] < 0 || z[e] < 0 || x[e] >= n[0] || y[e] >= n[1] || z[e] >= n[2])
                            continue;
                        mine[((int64_t)x[e] * n[1] + y[e]) * n[2] + z[e]] += sv * ln[e];
                    }
                }
            }
        }
        free(x); free(y); free(z); free(ln); free(off); free(bad);
    }
    if (!failed) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t v = 0; v < nvox; ++v) {
            float s = 0.0f;
            for (int t = 0; t < nthreads; ++t) s += part[(size_t)t * (size_t)nvox + v];
            back[v] = s;
        }
    }
    free(part);
    return failed ? -1 : total;
}

/*
 * Float64 products with a stored CSR matrix (int32 column index, float32
 * value, int64 row offsets) on `nthreads` OpenMP threads: y = A x, parallel
 * over rows, and x = A^T y into per-thread float64 partial vectors summed
 * in thread order (a fixed static schedule, so the result is reproducible).
 * The textbook sparse products of projection/proj.py:56-150 in float64, for
 * SIRT references at workload size (oracle/projection.CsrOperator).
 * oracle_csr_bp64 returns -1 when the partial vectors cannot be allocated.
 */
void oracle_csr_fp64(int64_t nrows, const int32_t *ind, const float *len, const int64_t *off,
                     const double *x, double *y, int nthreads) {
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t r = 0; r < nrows; ++r) {
        double s = 0.0;
        for (int64_t e = off[r]; e < off[r + 1]; ++e) s += x[ind[e]] * (double)len[e];
        y[r] = s;
    }
}

int oracle_csr_bp64(int64_t nrows, const int32_t *ind, const float *len, const int64_t *off,
                    int64_t ncols, const double *y, double *x, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    double *part = (double *)calloc((size_t)nthreads * (size_t)ncols, sizeof(double));
    if (!part) return -1;
#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
        const int tid = omp_get_thread_num();
#else
        const int tid = 0;
#endif
        double *mine = part + (size_t)tid * (size_t)ncols;
#pragma omp for schedule(static)
        for (int64_t r = 0; r < nrows; ++r) {
            const double v = y[r];
            for (int64_t e = off[r]; e < off[r + 1]; ++e) mine[ind[e]] += v * (double)len[e];
        }
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t c = 0; c < ncols; ++c) {
        double s = 0.0;
        for (int t = 0; t < nthreads; ++t) s += part[(size_t)t * (size_t)ncols + c];
        x[c] = s;
    }
    free(part);
    return 0;
}
