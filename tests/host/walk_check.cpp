// Host check of the Siddon walker (vir_amd/csrc/siddon_ray.h compiled for
// the CPU): the walks the unguarded kernels take must visit exactly the
// pairs of the general walk in world units.
//   * forward: the three-phase walk of k_siddon_fp (general walk to three
//     planes past the entry, walk_mid through the middle, general walk to
//     the exit; in voxel units on g.unit power-of-two grids for rays where
//     unit_exact holds) against walk_entries from walker_begin, every pair
//     and its keep flag;
//   * backward: for every 16x8x8 sub-brick a ray crosses, the guard-box
//     window of k_siddon_bp_wave; walk_mid (where window_mid holds) or
//     walk_entries after walker_seek, in voxel units likewise, against
//     walk_entries after walker_seek in world units.
// Argument 2 "loose" drops the margins of the middle phase and of
// window_mid, "unscaled_y" leaves the reciprocal direction in world units in
// the voxel-unit walk (planted mistakes): the check must then report
// mismatches.
// Input: int32 n[3], nrays; float32 d[3], o[3]; float64 eps; float32
// src[nrays][3], trg[nrays][3].  Output: one line of counts; exit status 1
// on any mismatch.
#include <cstdio>
#include <cstring>
#include <set>
#include <tuple>
#include <vector>
#include "siddon_ray.h"
using namespace vir;

struct Pair {
    int x, y, z;
    float len;
    bool keep;
};

static bool same(const std::vector<Pair>& a, const std::vector<Pair>& b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i) {
        if (a[i].keep != b[i].keep) return false;
        if (a[i].keep && (a[i].x != b[i].x || a[i].y != b[i].y || a[i].z != b[i].z ||
                          memcmp(&a[i].len, &b[i].len, 4) != 0))
            return false;
    }
    return true;
}

template <bool P2, bool ISO>
static int run(const GridK& g, const std::vector<float>& src, const std::vector<float>& trg, long long nr,
               bool loose, bool unscaled_y) {
    long long rays = 0, fp_pairs = 0, fp_mid = 0, fp_bad = 0, windows = 0, bp_mid = 0, bp_bad = 0, unit_rays = 0;
    const GridK gu = unit_grid(g);
    for (long long r = 0; r < nr; ++r) {
        Ray R;
        ray_setup(g, &src[3 * r], &trg[3 * r], R);
        if (!R.hit || !planes_unguarded(g, R)) continue;
        ++rays;
        // the kernels' voxel-unit walk: power-of-two grids, unit_exact rays
        const bool unit = P2 && g.unit && unit_exact(g, R);
        unit_rays += unit;
        Ray Ru = R;
        if (unit) unit_ray(g, Ru);
        if (unit && unscaled_y) Ru.y = R.y;
        // forward
        std::vector<Pair> A, B;
        auto into = [](std::vector<Pair>& v) {
            return [&v](int x, int y, int z, float l, bool k) { v.push_back({x, y, z, l, k}); };
        };
        Walker W1;
        walker_begin(g, R, W1);
        walk_entries<P2, ISO, false, true, true>(g, R, W1, VIR_INF, into(A));
        auto fp_walk = [&](const GridK& gw, const Ray& Rw, auto unit_tag) {
            constexpr bool U = decltype(unit_tag)::value;
            Walker W;
            walker_begin(gw, Rw, W);
            float gmax = 0.0f;
            for (int a = 0; a < 3; ++a) gmax = fmaxf(gmax, fabsf(Rw.dST[a]) * gw.inv_d[a]);
            const float a1 = loose ? Rw.amin : Rw.amin + 3.0f / gmax, a2 = loose ? Rw.amax : Rw.amax - 4.0f / gmax;
            walk_entries<P2, ISO, false, true, true, U>(gw, Rw, W, a1, into(B));
            const size_t before = B.size();
            if (W.consumed == 2 && (loose || window_mid(gw, Rw, W.prev, a2)))
                walk_mid<P2, ISO, true, true, U>(gw, Rw, W, a2, into(B));
            fp_mid += B.size() - before;
            walk_entries<P2, ISO, false, true, true, U>(gw, Rw, W, VIR_INF, into(B));
        };
        if (unit) fp_walk(gu, Ru, Fast<true>{});
        else fp_walk(g, R, Fast<false>{});
        fp_pairs += A.size();
        if (!same(A, B)) ++fp_bad;
        // backward windows of the sub-bricks holding the ray's entries
        std::set<std::tuple<int, int, int>> bricks;
        for (const Pair& p : A)
            if (p.keep && p.x >= 0 && p.y >= 0 && p.z >= 0 && p.x < g.n[0] && p.y < g.n[1] && p.z < g.n[2])
                bricks.insert(std::make_tuple(p.x / 16, p.y / 8, p.z / 8));
        for (const auto& b : bricks) {
            const int b0[3] = {std::get<0>(b) * 16, std::get<1>(b) * 8, std::get<2>(b) * 8};
            const int ext[3] = {16, 8, 8};
            float alo = -VIR_INF, ahi = VIR_INF;
            bool miss = false;
            for (int a = 0; a < 3; ++a) {
                const int b1 = b0[a] + ext[a] < g.n[a] ? b0[a] + ext[a] : g.n[a];
                const float lo = plane_coord(g, a, b0[a]) - 0.05f * g.d[a];
                const float hi = plane_coord(g, a, b1) + 0.05f * g.d[a];
                if (R.dST[a] != 0.0f) {
                    const float t1 = (lo - R.s[a]) * R.y[a], t2 = (hi - R.s[a]) * R.y[a];
                    alo = fmaxf(alo, fminf(t1, t2));
                    ahi = fminf(ahi, fmaxf(t1, t2));
                } else if (R.s[a] < lo || R.s[a] > hi) {
                    miss = true;
                }
            }
            if (miss || alo > ahi) continue;
            ++windows;
            const bool mid = loose || window_mid(g, R, alo, ahi);
            bp_mid += mid;
            std::vector<Pair> C, D;
            auto kept = [](std::vector<Pair>& v) {
                return [&v](int x, int y, int z, float l) { v.push_back({x, y, z, l, true}); };
            };
            Walker W2;
            walker_seek<true>(g, R, W2, alo);
            walk_entries<P2, ISO, false, true>(g, R, W2, ahi, kept(C));
            auto bp_walk = [&](const GridK& gw, const Ray& Rw, auto unit_tag) {
                constexpr bool U = decltype(unit_tag)::value;
                Walker W3;
                walker_seek<true>(gw, Rw, W3, alo);
                if (mid) walk_mid<P2, ISO, false, false, U>(gw, Rw, W3, ahi, kept(D));
                else walk_entries<P2, ISO, false, true, false, U>(gw, Rw, W3, ahi, kept(D));
            };
            if (unit) bp_walk(gu, Ru, Fast<true>{});
            else bp_walk(g, R, Fast<false>{});
            if (!same(C, D)) ++bp_bad;
        }
    }
    printf("rays %lld unit_rays %lld fp_pairs %lld fp_mid_pairs %lld fp_mismatch %lld bp_windows %lld "
           "bp_mid_windows %lld bp_mismatch %lld\n",
           rays, unit_rays, fp_pairs, fp_mid, fp_bad, windows, bp_mid, bp_bad);
    return (fp_bad || bp_bad) ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const bool loose = argc > 2 && strcmp(argv[2], "loose") == 0;
    const bool unscaled_y = argc > 2 && strcmp(argv[2], "unscaled_y") == 0;
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[4];
    float fd[6];
    double eps;
    if (fread(hdr, 4, 4, f) != 4 || fread(fd, 4, 6, f) != 6 || fread(&eps, 8, 1, f) != 1) return 2;
    const long long nr = hdr[3];
    std::vector<float> src(3 * nr), trg(3 * nr);
    if (fread(src.data(), 4, src.size(), f) != src.size() || fread(trg.data(), 4, trg.size(), f) != trg.size())
        return 2;
    fclose(f);
    // the grid as make_grid (vir_common.hip) builds it
    GridK g;
    g.unit = 1;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = hdr[a];
        g.d[a] = fd[a];
        g.o[a] = fd[3 + a];
        g.p0[a] = plane_at(g.n[a], g.d[a], g.o[a], 0);
        g.pN[a] = plane_at(g.n[a], g.d[a], g.o[a], g.n[a]);
        int ex;
        const float m = frexpf(g.d[a], &ex);
        g.pow2[a] = m == 0.5f;
        g.inv_d[a] = 1.0f / g.d[a];
        if (!(m == 0.5f && ex - 1 >= -4 && ex - 1 <= 4)) g.unit = 0;
        for (float c : {g.o[a], g.p0[a], g.pN[a]}) {
            const float ac = fabsf(c);
            if (!(ac == 0.0f || (ac >= 0x1p-60f && ac <= 0x1p100f))) g.unit = 0;
        }
    }
    g.eps_f = (float)eps;
    g.inv_eps = 1.0f / g.eps_f;
    g.eps_d = eps;
    const bool p2 = g.pow2[0] && g.pow2[1] && g.pow2[2];
    const bool iso = g.n[0] == g.n[1] && g.n[1] == g.n[2] && g.d[0] == g.d[1] && g.d[1] == g.d[2] &&
                     g.o[0] == g.o[1] && g.o[1] == g.o[2];
    if (p2 && iso) return run<true, true>(g, src, trg, nr, loose, unscaled_y);
    if (p2) return run<true, false>(g, src, trg, nr, loose, unscaled_y);
    return run<false, false>(g, src, trg, nr, loose, unscaled_y);
}
