// Oracle of the batched path tracer (test infrastructure; no reference counterpart: the
// reference shades primary hits with Shade() alone, shading.glsl:1-10, and has no path
// tracer; this restates the extension's path rule from its definition).
//
// Path rule: light from L = (0, 1, 0) with radiance 1 (the reversed lightDirection of
// shading.glsl:4). At a hit with normal n and ray direction d: facing normal nf = n if
// n . d <= 0, else -n; albedo a = tint[material] * Shade(n, 1); light contribution
// T * a * nf.y when nf.y > 0 and the shadow ray (hit point, L, tmin 0.001) is unoccluded;
// below the bounce limit the next ray leaves the hit point in a cosine-distributed
// direction about nf and the throughput becomes T * a; a ray that misses adds T * env.
// Random numbers: SplitMix64's output function of seed ^ (p << 24 | depth << 16 | k).
//
// Pins (tests/test_path_oracle.py): the generator's integer output against an independent
// SplitMix64 and its uniformity; the basis against orthonormality and handedness for random,
// axis and signed-zero normals; the direction distribution against the cosine lobe's
// moments (E[cos] = 2/3, E[cos^2] = 1/2, zero tangent means) and the hemisphere; the path
// accumulation against closed forms built from the oracle's own traced hits (direct light
// only, and one bounce off a lone convex fibre whose bounce rays must escape).
#pragma once
#include <cmath>
#include <cstdint>

#include "oracle_math.hpp"

namespace orc {

inline uint64_t path_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Uniform float in [0, 1): the top 24 bits of the mixed key, times 2^-24 (exact).
inline float path_rand(uint64_t seed, uint32_t p, uint32_t depth, uint32_t k) {
    const uint64_t key = (uint64_t(p) << 24) | (uint64_t(depth) << 16) | uint64_t(k);
    const uint32_t bits = uint32_t(path_mix64((seed ^ key) + 0x9E3779B97F4A7C15ull) >> 40);
    return float(bits) * std::ldexp(1.0f, -24);
}

// Tangents t1, t2 of the unit normal n (Duff, Burgess, Christensen, Hery, Kensler, Liani,
// Villemin, "Building an Orthonormal Basis, Revisited", JCGT 6(1) 2017, listing 3).
inline void path_basis(V3<float> n, V3<float>& t1, V3<float>& t2) {
    const float sign = std::copysign(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    t1 = mk(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    t2 = mk(b, sign + n.y * n.y * a, -n.y);
}

// Cosine-distributed direction about n (Malley: uniform disk point lifted to the
// hemisphere); the disk point by rejection from at most 16 pairs, the centre if all fail.
inline V3<float> path_cosine_dir(V3<float> n, uint64_t seed, uint32_t p, uint32_t depth) {
    float dx = 0.0f, dy = 0.0f, r2 = 0.0f;
    for (uint32_t k = 0; k < 16; ++k) {
        const float u = 2.0f * path_rand(seed, p, depth, 2 * k) - 1.0f;
        const float v = 2.0f * path_rand(seed, p, depth, 2 * k + 1) - 1.0f;
        const float q = u * u + v * v;
        if (q < 1.0f) {
            dx = u;
            dy = v;
            r2 = q;
            break;
        }
    }
    const float h = std::sqrt(1.0f - r2);
    V3<float> t1, t2;
    path_basis(n, t1, t2);
    const V3<float> w = mk((dx * t1.x + dy * t2.x) + h * n.x, (dx * t1.y + dy * t2.y) + h * n.y,
                           (dx * t1.z + dy * t2.z) + h * n.z);
    return normalize(w);
}

}  // namespace orc
