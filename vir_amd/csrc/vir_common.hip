// Library-wide state: per-thread error message, version, grid setup.
#include <cmath>
#include <cstring>
#include <string>

#include "vir_common.h"

namespace vir {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail_arg(const std::string& msg) {
    g_err = msg;
    return VIR_EINVAL;
}

int check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return VIR_OK;
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return VIR_EHIP;
}

// Host copy of boundspace (core/core_funcs.py:62-81); this translation unit
// is built with -ffp-contract=off so the f32 ops round as numpy's do.
static float host_plane(int n, float d, float o, int k) {
    float lin = (float)(2 * k - n);
    float t = lin * d;
    t = t * 0.5f;
    return t + o;
}

int make_grid(const vir_grid* gp, GridK& g) {
    if (!gp) return fail_arg("grid is NULL");
    for (int a = 0; a < 3; ++a) {
        if (gp->n[a] < 1 || gp->n[a] > (1 << 22)) return fail_arg("grid n out of range");
        if (!(gp->d[a] > 0.0f) || !std::isfinite(gp->d[a])) return fail_arg("grid d must be > 0");
        if (!std::isfinite(gp->o[a])) return fail_arg("grid origin not finite");
        g.n[a] = gp->n[a];
        g.d[a] = gp->d[a];
        g.o[a] = gp->o[a];
        g.p0[a] = host_plane(gp->n[a], gp->d[a], gp->o[a], 0);
        g.pN[a] = host_plane(gp->n[a], gp->d[a], gp->o[a], gp->n[a]);
        int ex;
        float m = std::frexp(gp->d[a], &ex);
        g.pow2[a] = (m == 0.5f) ? 1 : 0;
        g.inv_d[a] = 1.0f / gp->d[a];
    }
    if (!(gp->eps >= 0.0) || !std::isfinite(gp->eps)) return fail_arg("epsilon must be finite, >= 0");
    g.eps_f = (float)gp->eps;
    g.inv_eps = 1.0f / g.eps_f;
    g.eps_d = gp->eps;
    g.unit = 1;
    for (int a = 0; a < 3; ++a) {
        int ex;
        const float m = std::frexp(gp->d[a], &ex);  // d = 2^(ex-1) when m == 0.5
        if (!(m == 0.5f && ex - 1 >= -4 && ex - 1 <= 4)) g.unit = 0;
        for (float c : {g.o[a], g.p0[a], g.pN[a]}) {
            const float ac = std::fabs(c);
            if (!(ac == 0.0f || (ac >= 0x1p-60f && ac <= 0x1p100f))) g.unit = 0;
        }
    }
    return VIR_OK;
}

}  // namespace vir

extern "C" {

int vir_version(void) { return 1; }

const char* vir_last_error(void) { return vir::g_err.c_str(); }

}
