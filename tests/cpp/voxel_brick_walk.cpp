// The voxel walk's block moves (voxel_walk in vkhrt_amd/csrc/trace_impl.hpp) against the plain
// cell-by-cell 3D DDA, on the CPU with the same float arithmetic: random grids with sparse and
// clustered cells, empty 16x16x16 and 4x4x4 blocks left in one move, and rays that are
// axis-parallel, diagonal (equal face t on several axes), start on cell boundaries or end early.
// Prints the mismatch count. Argument "flip" plants a wrong tie rule in the move (the higher
// axis first) to show that the rays find it.
#include <cstdio>
#include <cstdint>
#include <cmath>
#include <vector>
#include <random>
#include <algorithm>
#include <cstring>
struct Res { float t; uint32_t idx; int ax; bool hit; };
struct Grid { int R[3]; float O[3]; float c; std::vector<uint8_t> occ, brick, super_; int bx, by, bz, sx, sy, sz; };
static bool filled(const Grid& g, int x, int y, int z) { return g.occ[x + y * g.R[0] + (size_t)z * g.R[0] * g.R[1]]; }
static bool g_flip = false;
template <bool JUMP>
Res walk(const Grid& g, const float oo[3], const float dd[3], float tmin, float tmax) {
    Res r{0, 0, -1, false};
    const int* R = g.R; const float* O = g.O; const float c = g.c;
    float inv[3]; float t0 = tmin, t1 = tmax; int ea = -1;
    for (int k = 0; k < 3; ++k) {
        const float hi = O[k] + float(R[k]) * c;
        if (dd[k] == 0.0f) { if (oo[k] < O[k] || oo[k] > hi) return r; inv[k] = 0.0f; continue; }
        inv[k] = 1.0f / dd[k];
        const float ta = (O[k] - oo[k]) * inv[k], tb = (hi - oo[k]) * inv[k];
        const float lo = ta < tb ? ta : tb, up = ta < tb ? tb : ta;
        if (lo > t0) { t0 = lo; ea = k; }
        if (up < t1) t1 = up;
    }
    if (!(t0 <= t1)) return r;
    int i[3], step[3]; float tn[3];
    for (int k = 0; k < 3; ++k) {
        const float f = floorf((oo[k] + dd[k] * t0 - O[k]) / c);
        i[k] = f < 0.0f ? 0 : f > float(R[k] - 1) ? R[k] - 1 : int(f);
        step[k] = dd[k] > 0.0f ? 1 : dd[k] < 0.0f ? -1 : 0;
        tn[k] = step[k] == 0 ? INFINITY : (O[k] + float(i[k] + (step[k] > 0 ? 1 : 0)) * c - oo[k]) * inv[k];
    }
    float te = t0;
    for (int it = 0, lim = R[0] + R[1] + R[2] + 3; it < lim; ++it) {
        int S = 1;  // size of the empty block around the cell: 16, 4, or 1 (the cell itself)
        if (JUMP) {
            if (!g.super_[(i[0] >> 4) + (i[1] >> 4) * g.sx + (i[2] >> 4) * g.sx * g.sy]) S = 16;
            else if (!g.brick[(i[0] >> 2) + (i[1] >> 2) * g.bx + (i[2] >> 2) * g.bx * g.by]) S = 4;
        }
        if (S == 1 && filled(g, i[0], i[1], i[2])) {
            int ax = ea;
            if (ax < 0) { ax = 0; if (fabsf(dd[1]) > fabsf(dd[ax])) ax = 1; if (fabsf(dd[2]) > fabsf(dd[ax])) ax = 2; }
            r.t = te; r.idx = i[0] + i[1] * R[0] + i[2] * R[0] * R[1]; r.ax = ax; r.hit = true; return r;
        }
        if (S > 1) {
            float fe = INFINITY; int a = -1, ja = 0; bool nan = false;
            for (int k = 0; k < 3; ++k) {
                if (step[k] == 0) continue;
                const int j = step[k] > 0 ? std::min(i[k] | (S - 1), R[k] - 1) : (i[k] & ~(S - 1));
                const float f = (O[k] + float(j + (step[k] > 0 ? 1 : 0)) * c - oo[k]) * inv[k];
                if (f != f) nan = true;
                if (a < 0 || f < fe) { fe = f; a = k; ja = j; }
            }
            if (!nan && a >= 0) {
                if (!(fe <= t1)) return r;
                for (int k = 0; k < 3; ++k) {
                    if (k == a || step[k] == 0) continue;
                    for (int s = 0; s < S - 1 && (tn[k] < fe || (tn[k] == fe && (g_flip ? k > a : k < a))); ++s) {
                        i[k] += step[k];
                        tn[k] = (O[k] + float(i[k] + (step[k] > 0 ? 1 : 0)) * c - oo[k]) * inv[k];
                    }
                }
                te = fe > t0 ? fe : t0;
                ea = a;
                i[a] = ja + step[a];
                if (i[a] < 0 || i[a] >= R[a]) return r;
                tn[a] = (O[a] + float(i[a] + (step[a] > 0 ? 1 : 0)) * c - oo[a]) * inv[a];
                continue;
            }
            if (filled(g, i[0], i[1], i[2])) return r;  // unreachable: an empty block holds no filled cell
        }
        int a = 0;
        if (tn[1] < tn[a]) a = 1;
        if (tn[2] < tn[a]) a = 2;
        if (!(tn[a] <= t1)) return r;
        te = tn[a] > t0 ? tn[a] : t0;
        ea = a;
        i[a] += step[a];
        if (i[a] < 0 || i[a] >= R[a]) return r;
        tn[a] = (O[a] + float(i[a] + (step[a] > 0 ? 1 : 0)) * c - oo[a]) * inv[a];
    }
    return r;
}
int main(int argc, char** argv) {
    g_flip = argc > 1 && std::strcmp(argv[1], "flip") == 0;
    std::mt19937 rng(1);
    std::uniform_real_distribution<float> U(0, 1);
    long long bad = 0, total = 0, hits = 0;
    for (int gi = 0; gi < 60; ++gi) {
        Grid g;
        for (int k = 0; k < 3; ++k) g.R[k] = 1 + int(U(rng) * 110);
        g.c = gi % 3 == 0 ? 0.5f : gi % 3 == 1 ? 0.1f : 0.37f;
        for (int k = 0; k < 3; ++k) g.O[k] = (U(rng) - 0.5f) * 20;
        size_t n = size_t(g.R[0]) * g.R[1] * g.R[2];
        g.occ.assign(n, 0);
        const float dens = gi % 4 == 0 ? 0.0005f : gi % 4 == 1 ? 0.005f : gi % 4 == 2 ? 0.05f : 0.0f;
        for (size_t q = 0; q < n; ++q) g.occ[q] = U(rng) < dens;
        // a few clusters
        for (int cl = 0; cl < 3 && gi % 4 != 3; ++cl) {
            int cx = int(U(rng) * g.R[0]), cy = int(U(rng) * g.R[1]), cz = int(U(rng) * g.R[2]);
            for (int z = std::max(0, cz - 2); z < std::min(g.R[2], cz + 3); ++z)
                for (int y = std::max(0, cy - 2); y < std::min(g.R[1], cy + 3); ++y)
                    for (int x = std::max(0, cx - 2); x < std::min(g.R[0], cx + 3); ++x) g.occ[x + y * g.R[0] + (size_t)z * g.R[0] * g.R[1]] = 1;
        }
        g.bx = (g.R[0] + 3) >> 2; g.by = (g.R[1] + 3) >> 2; g.bz = (g.R[2] + 3) >> 2;
        g.brick.assign(size_t(g.bx) * g.by * g.bz, 0);
        for (int z = 0; z < g.R[2]; ++z) for (int y = 0; y < g.R[1]; ++y) for (int x = 0; x < g.R[0]; ++x)
            if (filled(g, x, y, z)) g.brick[(x >> 2) + (y >> 2) * g.bx + (z >> 2) * g.bx * g.by] = 1;
        g.sx = (g.R[0] + 15) >> 4; g.sy = (g.R[1] + 15) >> 4; g.sz = (g.R[2] + 15) >> 4;
        g.super_.assign(size_t(g.sx) * g.sy * g.sz, 0);
        for (int z = 0; z < g.R[2]; ++z) for (int y = 0; y < g.R[1]; ++y) for (int x = 0; x < g.R[0]; ++x)
            if (filled(g, x, y, z)) g.super_[(x >> 4) + (y >> 4) * g.sx + (z >> 4) * g.sx * g.sy] = 1;
        for (int ri = 0; ri < 40000; ++ri) {
            float o[3], d[3];
            const int kind = ri % 6;
            for (int k = 0; k < 3; ++k) {
                const float ext = g.R[k] * g.c;
                o[k] = kind == 0 ? g.O[k] + U(rng) * ext : g.O[k] + (U(rng) * 3 - 1) * ext;
                if (kind == 5) o[k] = g.O[k] + float(int(U(rng) * g.R[k])) * g.c;  // on cell boundaries
                d[k] = U(rng) * 2 - 1;
            }
            if (kind == 1) d[ri % 3] = 0.0f;
            if (kind == 2) { d[ri % 3] = 0.0f; d[(ri + 1) % 3] = 0.0f; }
            if (kind == 3) { d[0] = d[1] = d[2] = (ri & 1) ? 1.0f : -1.0f; if (ri & 2) d[1] = -d[1]; }  // diagonal: ties
            if (kind == 4) { d[0] = 1.0f; d[1] = 1.0f; d[2] = 0.5f; o[0] = g.O[0]; o[1] = g.O[1]; }
            if (d[0] == 0 && d[1] == 0 && d[2] == 0) d[0] = 1;
            const float tmax = (ri % 7 == 0) ? U(rng) * 10 : 1e4f;
            Res a = walk<false>(g, o, d, 0.001f, tmax), b = walk<true>(g, o, d, 0.001f, tmax);
            ++total; hits += a.hit;
            if (a.hit != b.hit || (a.hit && (a.idx != b.idx || a.ax != b.ax || std::memcmp(&a.t, &b.t, 4) != 0))) {
                if (bad < 5) printf("mismatch grid %d ray %d: %d/%d idx %u/%u t %.9g/%.9g ax %d/%d\n", gi, ri, a.hit, b.hit, a.idx, b.idx, a.t, b.t, a.ax, b.ax);
                ++bad;
            }
        }
    }
    printf("%lld rays, %lld hits, %lld mismatches\n", total, hits, bad);
    return 0;
}
