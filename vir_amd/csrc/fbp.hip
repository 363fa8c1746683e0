// Filtered backprojection for gfx950: projection filtering as a direct
// convolution with the taps of a windowed ramp filter, and a pixel-driven
// weighted backprojection with linear interpolation (parallel beam, fan beam
// with a flat or an equiangular detector).
//
// Filtering.  filtration.filter_sino (filtration.py:267-329) zero-pads each
// row to z >= 2 nu, multiplies its rFFT by 2 H(f) (H the windowed ramp of
// ramp_filter / window_filter, filtration.py:4-264) and crops the inverse to
// nu samples.  H is real and even, so that is the linear convolution of the
// row with the symmetric kernel g = irfft(2 H) -- no wrap-around reaches the
// first nu outputs -- evaluated here as out[i] = sum_j w[j] row[j] g[|i - j|]
// from LDS copies of the row and the taps, in double.  The host designs g
// (and applies the equiangular (g/sin g)^2 correction to it).
//
// Backprojection.  Kak & Slaney's fan-beam formulas in the projector
// conventions of sys_mat/dd.py (view angle a, c = cos a, s = sin a; pixel
// centre (x, y); p = -s x + c y across the ray, o = c x + s y toward the
// source at DSO):
//   parallel : coordinate p,                                 weight 1
//   flat fan : virtual-detector coordinate DSO p / (DSO - o), weight (DSO / (DSO - o))^2
//   arc fan  : fan angle atan2(p, DSO - o),                  weight 1 / ((DSO - o)^2 + p^2)
// sampled by linear interpolation between bin centres (zero outside the
// detector; positions, weights and sums in double), times the caller's scale.
#include <algorithm>
#include <cmath>

#include "vir_common.h"

namespace vir {

constexpr int kFbpMaxBins = 4096;  // row + taps staged in LDS (32 KiB)

__global__ __launch_bounds__(256) void k_fbp_filter(const float* __restrict__ sino, int64_t nrows, int nu,
                                                    const float* __restrict__ weight,
                                                    const float* __restrict__ taps, float* __restrict__ out) {
    extern __shared__ float fbp_lds[];
    float* row = fbp_lds;
    float* g = fbp_lds + nu;
    const int64_t r = blockIdx.x;
    if (r >= nrows) return;
    const float* in = sino + r * nu;
    for (int j = threadIdx.x; j < nu; j += blockDim.x) {
        row[j] = weight ? in[j] * weight[j] : in[j];
        g[j] = taps[j];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nu; i += blockDim.x) {
        double acc = 0.0;
        for (int j = 0; j < nu; ++j) {
            const int k = i >= j ? i - j : j - i;
            acc += (double)row[j] * (double)g[k];
        }
        out[r * nu + i] = (float)acc;
    }
}

struct FbpGeo {
    int nx, ny, nu, na, mode;
    float d, DSO;
    double c0, inv_du;  // filter-domain coordinate of bin 0's centre, 1 / bin spacing
    float scale;
};

template <int MODE>
__global__ __launch_bounds__(256) void k_fbp_bp(FbpGeo G, const float* __restrict__ cs,
                                                const float* __restrict__ sn, const float* __restrict__ q,
                                                float* __restrict__ img, int accumulate) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)G.nx * G.ny) return;
    const int ix = (int)(t / G.ny), iy = (int)(t % G.ny);
    // geometry and interpolation in double: the filtered projections are
    // rough (ramp-filtered noise), so the sampling position must not round
    const double x = (double)G.d * (((double)ix + 0.5) - 0.5 * (double)G.nx);
    const double y = (double)G.d * (((double)iy + 0.5) - 0.5 * (double)G.ny);
    double acc = 0.0;
    for (int ia = 0; ia < G.na; ++ia) {
        const double c = cs[ia], s = sn[ia];
        const double p = -s * x + c * y;
        double coord, w;
        if (MODE == 0) {
            coord = p;
            w = 1.0;
        } else {
            const double den = (double)G.DSO - (c * x + s * y);
            if (MODE == 1) {
                const double m = (double)G.DSO / den;
                coord = m * p;
                w = m * m;
            } else {
                coord = atan2(p, den);
                w = 1.0 / (den * den + p * p);
            }
        }
        const double tf = (coord - G.c0) * G.inv_du;
        const double fl = floor(tf);
        const double fr = tf - fl;
        if (!(fl >= -1.0 && fl <= (double)G.nu - 1.0)) continue;
        const int j = (int)fl;
        const float* row = q + (int64_t)ia * G.nu;
        const double a = j >= 0 ? (double)row[j] : 0.0;
        const double b = j + 1 < G.nu ? (double)row[j + 1] : 0.0;
        acc += w * ((1.0 - fr) * a + fr * b);
    }
    float r = (float)acc * G.scale;
    if (accumulate) r += img[t];
    img[t] = r;
}

}  // namespace vir

using namespace vir;

extern "C" {

int vir_fbp_filter(const float* sino, int64_t nrows, int32_t nu, const float* weight, const float* taps, float* out,
                   void* stream) {
    if (nrows < 0 || nu < 1) return fail_arg("nrows must be >= 0 and nu >= 1");
    if (nu > kFbpMaxBins) return fail_arg("nu exceeds 4096 detector bins");
    if (nrows == 0) return VIR_OK;
    if (!sino || !taps || !out) return fail_arg("NULL pointer argument");
    if (sino == out) return fail_arg("in-place filtering is not supported");
    const size_t lds = (size_t)2 * nu * sizeof(float);
    hipLaunchKernelGGL(k_fbp_filter, dim3((unsigned)nrows), dim3(256), lds, as_stream(stream), sino, nrows, nu,
                       weight, taps, out);
    return check_hip(hipGetLastError(), "vir_fbp_filter launch");
}

int vir_fbp_bp(const vir_dd_geom* geom, const float* cosv, const float* sinv, const float* q, float* img,
               float scale, uint32_t flags, void* stream) {
    if (!geom) return fail_arg("geometry is NULL");
    if (geom->mode < 0 || geom->mode > 2) return fail_arg("mode must be 0 (par), 1 (fan flat) or 2 (fan arc)");
    if (geom->nx < 1 || geom->ny < 1 || geom->nu < 1 || geom->na < 0) return fail_arg("bad sizes");
    if (!(geom->d_pix > 0.0f) || !(geom->du > 0.0f)) return fail_arg("spacings must be > 0");
    if (geom->mode > 0 && !(geom->DSO > 0.0f && geom->DSD > 0.0f)) return fail_arg("DSO and DSD must be > 0");
    if (!img || (geom->na > 0 && (!cosv || !sinv || !q))) return fail_arg("NULL pointer argument");
    FbpGeo G;
    G.nx = geom->nx; G.ny = geom->ny; G.nu = geom->nu; G.na = geom->na; G.mode = geom->mode;
    G.d = geom->d_pix; G.DSO = geom->DSO; G.scale = scale;
    // bin-0 centre and spacing in the filter domain (detector, virtual
    // detector through the isocentre, or fan angle)
    const double u0 = (double)geom->du * (0.5 - 0.5 * geom->nu + (double)geom->su);
    double f = 1.0;
    if (geom->mode == 1) f = (double)geom->DSO / geom->DSD;
    if (geom->mode == 2) f = 1.0 / geom->DSD;
    G.c0 = u0 * f;
    G.inv_du = 1.0 / ((double)geom->du * f);
    const int64_t n = (int64_t)G.nx * G.ny;
    dim3 grd((unsigned)((n + 255) / 256));
    hipStream_t st = as_stream(stream);
    const int acc = (flags & VIR_ACCUMULATE) ? 1 : 0;
    if (G.mode == 0) hipLaunchKernelGGL(k_fbp_bp<0>, grd, dim3(256), 0, st, G, cosv, sinv, q, img, acc);
    else if (G.mode == 1) hipLaunchKernelGGL(k_fbp_bp<1>, grd, dim3(256), 0, st, G, cosv, sinv, q, img, acc);
    else hipLaunchKernelGGL(k_fbp_bp<2>, grd, dim3(256), 0, st, G, cosv, sinv, q, img, acc);
    return check_hip(hipGetLastError(), "vir_fbp_bp launch");
}

}  // extern "C"
