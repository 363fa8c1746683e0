// Host check of the Siddon walker (vir_amd/csrc/siddon_ray.h compiled for
// the CPU): the lean walks must visit exactly the pairs of the general walk.
//   * forward: the three-phase walk of k_siddon_fp (general walk to three
//     planes past the entry, walk_mid -- with voxel-unit indices on
//     power-of-two grids -- through the middle, general walk to the exit)
//     against walk_entries from walker_begin, every pair and its keep flag;
//   * backward: for every 16x8x8 sub-brick a ray crosses, the guard-box
//     window of k_siddon_bp_wave; where window_mid holds, walk_mid after
//     walker_seek against walk_entries after walker_seek.
// Argument 2 "loose" drops the margins of the middle phase and of
// window_mid (a planted mistake): the check must then report mismatches.
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
               bool loose) {
    long long rays = 0, fp_pairs = 0, fp_mid = 0, fp_bad = 0, windows = 0, bp_mid = 0, bp_bad = 0;
    for (long long r = 0; r < nr; ++r) {
        Ray R;
        ray_setup(g, &src[3 * r], &trg[3 * r], R);
        if (!R.hit || !planes_unguarded(g, R)) continue;
        ++rays;
        // forward
        std::vector<Pair> A, B;
        auto into = [](std::vector<Pair>& v) {
            return [&v](int x, int y, int z, float l, bool k) { v.push_back({x, y, z, l, k}); };
        };
        Walker W1;
        walker_begin(g, R, W1);
        walk_entries<P2, ISO, false, true, true>(g, R, W1, VIR_INF, into(A));
        Walker W;
        walker_begin(g, R, W);
        float gmax = 0.0f;
        for (int a = 0; a < 3; ++a) gmax = fmaxf(gmax, fabsf(R.dST[a]) * g.inv_d[a]);
        const float a1 = loose ? R.amin : R.amin + 3.0f / gmax, a2 = loose ? R.amax : R.amax - 4.0f / gmax;
        walk_entries<P2, ISO, false, true, true>(g, R, W, a1, into(B));
        const size_t before = B.size();
        if (W.consumed == 2 && (loose || window_mid(g, R, W.prev, a2)) && (!P2 || scaled_exact(g, R)))
            walk_mid<P2, ISO, true, true, P2>(g, R, W, a2, into(B));
        fp_mid += B.size() - before;
        walk_entries<P2, ISO, false, true, true>(g, R, W, VIR_INF, into(B));
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
            if (!loose && !window_mid(g, R, alo, ahi)) continue;
            ++bp_mid;
            std::vector<Pair> C, D;
            auto kept = [](std::vector<Pair>& v) {
                return [&v](int x, int y, int z, float l) { v.push_back({x, y, z, l, true}); };
            };
            Walker W2;
            walker_seek<true>(g, R, W2, alo);
            walk_entries<P2, ISO, false, true>(g, R, W2, ahi, kept(C));
            Walker W3;
            walker_seek<true>(g, R, W3, alo);
            walk_mid<P2, ISO>(g, R, W3, ahi, kept(D));
            if (!same(C, D)) ++bp_bad;
        }
    }
    printf("rays %lld fp_pairs %lld fp_mid_pairs %lld fp_mismatch %lld bp_windows %lld bp_mid_windows %lld "
           "bp_mismatch %lld\n",
           rays, fp_pairs, fp_mid, fp_bad, windows, bp_mid, bp_bad);
    return (fp_bad || bp_bad) ? 1 : 0;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const bool loose = argc > 2 && strcmp(argv[2], "loose") == 0;
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
    g.scaled = 1;
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
        const float p0 = fabsf(g.p0[a]);
        if (!(m == 0.5f && ex - 1 >= -20 && ex - 1 <= 20)) g.scaled = 0;
        if (!(p0 == 0.0f || (p0 >= 0x1p-80f && p0 <= 0x1p100f))) g.scaled = 0;
        g.p0s[a] = g.p0[a] * g.inv_d[a];
    }
    g.eps_f = (float)eps;
    g.inv_eps = 1.0f / g.eps_f;
    g.eps_d = eps;
    const bool p2 = g.pow2[0] && g.pow2[1] && g.pow2[2];
    const bool iso = g.n[0] == g.n[1] && g.n[1] == g.n[2] && g.d[0] == g.d[1] && g.d[1] == g.d[2] &&
                     g.o[0] == g.o[1] && g.o[1] == g.o[2];
    if (p2 && iso) return run<true, true>(g, src, trg, nr, loose);
    if (p2) return run<true, false>(g, src, trg, nr, loose);
    return run<false, false>(g, src, trg, nr, loose);
}
