// Errors, version, PCG32 and the seeded workload generators (host code).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <utility>

#include "internal.h"

namespace rumi {

static thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { g_err[0] = 0; }

const char* last_error() { return g_err; }

}  // namespace rumi

using namespace rumi;

extern "C" const char* rumi_last_error(void) { return rumi::last_error(); }

extern "C" const char* rumi_version(void) { return "rumi-mi355x 0.1 (gfx950)"; }

extern "C" int rumi_pcg32_sequence(uint64_t initstate, uint64_t initseq, uint32_t n,
                                   uint32_t* host_out) {
    if (!host_out && n) return fail(RUMI_ERR_INVALID, "rumi_pcg32_sequence: null output");
    Pcg32 g(initstate, initseq);
    for (uint32_t i = 0; i < n; ++i) host_out[i] = g.next();
    return RUMI_OK;
}

// ------------------------------------------------------------ icosphere
// Class-I geodesic subdivision of the icosahedron at frequency f.  Vertex
// numbering: the 12 corners, then the f-1 interior points of each of the 30
// edges (edge order of first appearance, points from the lower to the higher
// corner index), then the interior points of each face.  Every point is
// placed on the unit sphere (in double) and pushed to radius
// 1 + amplitude * (2u - 1), u = PCG32(seed, vertex index).uniform().
namespace {

struct V3d {
    double x, y, z;
};

V3d lerp3(const V3d& a, const V3d& b, const V3d& c, double wa, double wb, double wc) {
    return {a.x * wa + b.x * wb + c.x * wc, a.y * wa + b.y * wb + c.y * wc,
            a.z * wa + b.z * wb + c.z * wc};
}

}  // namespace

extern "C" int rumi_make_icosphere(uint32_t f, uint64_t seed, float amplitude, float* xyz,
                                   uint32_t* idx) {
    if (f < 1 || f > 4096) return fail(RUMI_ERR_INVALID, "icosphere frequency %u not in [1,4096]", f);
    if (!xyz || !idx) return fail(RUMI_ERR_INVALID, "icosphere: null output");
    const double p = (1.0 + std::sqrt(5.0)) / 2.0;
    const V3d corner[12] = {{-1, p, 0}, {1, p, 0}, {-1, -p, 0}, {1, -p, 0},
                            {0, -1, p}, {0, 1, p}, {0, -1, -p}, {0, 1, -p},
                            {p, 0, -1}, {p, 0, 1}, {-p, 0, -1}, {-p, 0, 1}};
    // Counter-clockwise seen from outside.
    const int face[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                             {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                             {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                             {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
    const uint64_t nv = 10ull * f * f + 2;
    std::vector<V3d> pos(nv);
    for (int i = 0; i < 12; ++i) pos[i] = corner[i];
    // Edges in order of first appearance.
    std::map<std::pair<int, int>, int> edge_id;
    std::vector<std::pair<int, int>> edges;
    for (int t = 0; t < 20; ++t)
        for (int k = 0; k < 3; ++k) {
            int a = face[t][k], b = face[t][(k + 1) % 3];
            std::pair<int, int> key(std::min(a, b), std::max(a, b));
            if (!edge_id.count(key)) {
                edge_id[key] = (int)edges.size();
                edges.push_back(key);
            }
        }
    const uint64_t edge_base = 12;
    for (size_t e = 0; e < edges.size(); ++e)
        for (uint32_t k = 1; k < f; ++k) {
            double w = (double)k / f;
            pos[edge_base + e * (f - 1) + (k - 1)] =
                lerp3(corner[edges[e].first], corner[edges[e].second], corner[0], 1.0 - w, w, 0.0);
        }
    const uint64_t face_base = edge_base + 30ull * (f - 1);
    const uint64_t per_face = f >= 3 ? (uint64_t)(f - 1) * (f - 2) / 2 : 0;
    // Index of grid point (i, j) of face t: i steps toward corner B, j toward C.
    auto vid = [&](int t, uint32_t i, uint32_t j) -> uint64_t {
        const int A = face[t][0], B = face[t][1], C = face[t][2];
        if (i == 0 && j == 0) return A;
        if (i == f && j == 0) return B;
        if (i == 0 && j == f) return C;
        auto on_edge = [&](int a, int b, uint32_t k) -> uint64_t {  // k steps from a toward b
            int e = edge_id[std::pair<int, int>(std::min(a, b), std::max(a, b))];
            uint32_t s = a < b ? k : f - k;
            return edge_base + (uint64_t)e * (f - 1) + (s - 1);
        };
        if (j == 0) return on_edge(A, B, i);
        if (i == 0) return on_edge(A, C, j);
        if (i + j == f) return on_edge(B, C, j);
        // interior: rows j = 1..f-2, i = 1..f-1-j
        uint64_t off = (uint64_t)(j - 1) * (f - 1) - (uint64_t)(j - 1) * j / 2;
        return face_base + (uint64_t)t * per_face + off + (i - 1);
    };
    for (int t = 0; t < 20; ++t)
        for (uint32_t j = 1; j + 1 < f; ++j)
            for (uint32_t i = 1; i + j < f; ++i) {
                double wb = (double)i / f, wc = (double)j / f;
                pos[vid(t, i, j)] = lerp3(corner[face[t][0]], corner[face[t][1]],
                                          corner[face[t][2]], 1.0 - wb - wc, wb, wc);
            }
    for (uint64_t v = 0; v < nv; ++v) {
        const V3d& q = pos[v];
        double len = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
        Pcg32 g(seed, v);
        double r = 1.0 + (double)amplitude * (2.0 * (double)g.uniform() - 1.0);
        xyz[3 * v + 0] = (float)(q.x / len * r);
        xyz[3 * v + 1] = (float)(q.y / len * r);
        xyz[3 * v + 2] = (float)(q.z / len * r);
    }
    uint64_t tcount = 0;
    for (int t = 0; t < 20; ++t)
        for (uint32_t j = 0; j < f; ++j)
            for (uint32_t i = 0; i + j < f; ++i) {
                uint64_t a = vid(t, i, j), b = vid(t, i + 1, j), c = vid(t, i, j + 1);
                idx[3 * tcount + 0] = (uint32_t)a;
                idx[3 * tcount + 1] = (uint32_t)b;
                idx[3 * tcount + 2] = (uint32_t)c;
                ++tcount;
                if (i + j + 1 < f) {
                    uint64_t d = vid(t, i + 1, j + 1);
                    idx[3 * tcount + 0] = (uint32_t)b;
                    idx[3 * tcount + 1] = (uint32_t)d;
                    idx[3 * tcount + 2] = (uint32_t)c;
                    ++tcount;
                }
            }
    if (tcount != 20ull * f * f) return fail(RUMI_ERR_INVALID, "icosphere: triangle count mismatch");
    return RUMI_OK;
}

extern "C" int rumi_make_rays_uniform(uint64_t n, uint64_t seed, float half_extent,
                                      rumi_ray* rays) {
    if (!rays && n) return fail(RUMI_ERR_INVALID, "rumi_make_rays_uniform: null output");
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t i = 0; i < n; ++i) {
        Pcg32 g(seed, i);
        double ux = g.uniform(), uy = g.uniform(), uz = g.uniform();
        double u = g.uniform(), v = g.uniform();
        double z = 1.0 - 2.0 * u;
        double s = std::sqrt(std::max(0.0, 1.0 - z * z));
        double phi = two_pi * v;
        rumi_ray& r = rays[i];
        r.ox = (float)((2.0 * ux - 1.0) * half_extent);
        r.oy = (float)((2.0 * uy - 1.0) * half_extent);
        r.oz = (float)((2.0 * uz - 1.0) * half_extent);
        r.tmin = 0.0f;
        r.dx = (float)(s * std::cos(phi));
        r.dy = (float)(s * std::sin(phi));
        r.dz = (float)z;
        r.tmax = std::numeric_limits<float>::infinity();
    }
    return RUMI_OK;
}
