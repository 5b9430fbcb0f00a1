// xpbd_oracle.cpp -- CPU oracle for the ABI41 XPBD substep path.
//
// TEST INFRASTRUCTURE ONLY.  This is a scalar C++ restatement of the reference
// solver (native/src/ssbl_xpbd_cuda_abi41.cu, cited as A41:<line>) used as the
// checker for the HIP library and as the CPU baseline in bench.py.  It is
// never linked into, loaded by, or called from the product path.
//
// It exports the same ABI41 entry points (include/ssbl_xpbd_hip.h) and the
// ssbl_ext_* extension entry points, so a single host harness drives it and the
// HIP library side by side.
//
// Semantics: the reference kernels that race on float atomics are restated in
// the deterministic form this project defines for them (the HIP library
// implements the same definitions):
//   * direct stretch, LRA tethers and uncoloured bends: Jacobi sweep -- every
//     constraint reads the positions from before the pass; each vertex sums its
//     corrections onto its old value in ascending constraint order;
//   * self-collision accumulation: per-vertex sums in 2^-32 m fixed point
//     (order independent), converted back to float once per pass;
//   * spatial hashes: bucket contents in (primitive index, cell loop) order,
//     the first 32 entries of a bucket visible to queries, as the reference's
//     32-slot buckets keep the first 32 arrivals;
//   * PCG: block rows summed in CSR (edge) order, inner products in FP64.
// rsqrtf(x) is restated as 1/sqrtf(x); the reference's fast-math intrinsics
// are not reproduced.  Built with -ffp-contract=off: every fused multiply-add
// below is an explicit std::fma, as in the reference's fma_rn helpers.
//
// Precision: Solver<float> is the FP32 reference-semantics oracle;
// Solver<double> runs the same algorithm in FP64 to bound FP32 rounding error
// (selected by ssbl_oracle_set_precision(64) before create).

#include "ssbl_xpbd_hip.h"
#include "ssbl_xpbd_hip_ext.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_error;
int g_precision_bits = 32;
bool g_dynamic_hash = true;   // test hook: false makes every dynamic query a brute-force scan

bool fail(const char* message) {
    g_error = message ? message : "oracle error";
    return false;
}

// ---------------------------------------------------------------- constants
// A41:23-180 (only the ones on the restated path).
constexpr float kEpsF = 1.0e-7f;
constexpr int kSelfHashMinCount = 128;           // A41:28
constexpr int kBucketSlots = 32;                 // A41:29,32,36
constexpr int kMinBuckets = 1024;                // A41:30
constexpr int kSelfTriangleHashMinCount = 64;    // A41:31
constexpr int kSelfEdgeHashMinCount = 64;        // A41:35
constexpr int kSelfHashMaxPrimitiveCells = 64;   // A41:34
constexpr int kNeighborSlots = 32;               // A41:38
constexpr int kMaxForceFields = 64;              // A41:166
constexpr int kSelfModeFast = 1;                 // A41:140

enum CountSlot {
    kCountSoft = 0, kCountExactImpulse = 1, kCountEdgeEdge = 2, kCountHardFallbacks = 3,
    kCountSelfCandidates = 9, kCountSelfOverflow = 10, kCountStaticSdf = 11, kCountLraTacks = 12,
    kCountBendingWings = 13, kCountTackGuards = 14, kCountBendingGuards = 15,
    // fast self-collision mode, A41:56-72
    kCountIslandCandidates = 16, kCountIslandClusters = 17, kCountIslandVertexRefs = 18, kCountIslandApplied = 19,
    kCountIslandGuarded = 20, kCountIslandMaxDeltaMicro = 21, kCountFilterSeen = 22, kCountFilterAcceptedVv = 23,
    kCountFilterAcceptedVt = 24, kCountFilterAcceptedEe = 25, kCountFilterRejectedRest = 26, kCountFilterRejectedDuplicate = 27,
    kCountFilterRejectedOwnership = 28, kCountFilterCacheHits = 29, kCountFilterCacheMisses = 30, kCountClusterCount = 31,
    kCountClusterOwnedContacts = 32,
    kCountFriction = 33, kCountSlots = 38
};
// fast self-collision filters, A41:142-149
constexpr int kFilterLocalSlots = 64;
constexpr int kClusterMaxLocal = 16;
constexpr int kClusterMaxContactsPerOwner = 1;
constexpr int kIslandMinContacts = 3;
constexpr int kIslandMaxContacts = 48;
constexpr float kIslandCorrectionScale = 0.50f;
constexpr float kIslandMaxDeltaScale = 0.30f;
constexpr float kIslandMinAvgDepthScale = 0.04f;
enum PairType { kPairVv = 0, kPairVt = 1, kPairEe = 2 };

template <class R> struct V3 { R x, y, z; };
template <class R> inline V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> inline V3<R> mul(V3<R> a, R s) { return mk(a.x * s, a.y * s, a.z * s); }
// A41:622-628 dot with fused multiply-adds in z(y(x)) order.
template <class R> inline R dot(V3<R> a, V3<R> b) {
    return std::fma(a.z, b.z, std::fma(a.y, b.y, std::fma(a.x, b.x, R(0))));
}
// A41:630-636
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return mk(std::fma(a.y, b.z, -(a.z * b.y)), std::fma(a.z, b.x, -(a.x * b.z)), std::fma(a.x, b.y, -(a.y * b.x)));
}
template <class R> inline R length(V3<R> a) { return std::sqrt(std::max(dot(a, a), R(0))); }
template <class R> inline bool finite(V3<R> a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
template <class R> inline R clamp01(R v) { return std::min(std::max(v, R(0)), R(1)); }
template <class R> inline R rsqrt(R v) { return R(1) / std::sqrt(v); }

// A41:654-664 cell coordinate and bucket hash (bit-exact integer arithmetic).
template <class R> inline int cell_coord(R value, R cell) {
    return static_cast<int>(std::floor(value / std::max(cell, R(kEpsF))));
}
inline unsigned hash_cell(int x, int y, int z, int buckets) {
    const unsigned h = static_cast<unsigned>(x) * 73856093u ^ static_cast<unsigned>(y) * 19349663u ^
                       static_cast<unsigned>(z) * 83492791u;
    return buckets > 0 ? (h & static_cast<unsigned>(buckets - 1)) : 0u;
}
inline int next_pow2(int v) {  // A41:5966-5973
    int r = 1;
    v = std::max(v, 1);
    while (r < v && r < (1 << 30)) r <<= 1;
    return r;
}

// A41:1589-1635 closest point on a triangle (Ericson's region test).
template <class R> V3<R> closest_point_on_triangle(V3<R> p, V3<R> a, V3<R> b, V3<R> c) {
    const R eps = R(kEpsF);
    V3<R> ab = b - a, ac = c - a, ap = p - a;
    R d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= R(0) && d2 <= R(0)) return a;
    V3<R> bp = p - b;
    R d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= R(0) && d4 <= d3) return b;
    R vc = d1 * d4 - d3 * d2;
    if (vc <= R(0) && d1 >= R(0) && d3 <= R(0)) return a + mul(ab, d1 / std::max(d1 - d3, eps));
    V3<R> cp = p - c;
    R d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= R(0) && d5 <= d6) return c;
    R vb = d5 * d2 - d1 * d6;
    if (vb <= R(0) && d2 >= R(0) && d6 <= R(0)) return a + mul(ac, d2 / std::max(d2 - d6, eps));
    R va = d3 * d6 - d5 * d4;
    if (va <= R(0) && (d4 - d3) >= R(0) && (d5 - d6) >= R(0)) {
        R w = (d4 - d3) / std::max((d4 - d3) + (d5 - d6), eps);
        return b + mul(c - b, w);
    }
    R denom = R(1) / std::max(va + vb + vc, eps);
    R v = vb * denom, w = vc * denom;
    return a + (mul(ab, v) + mul(ac, w));
}

// A41:4594-4617
template <class R> void triangle_barycentric(V3<R> p, V3<R> a, V3<R> b, V3<R> c, R* wa, R* wb, R* wc) {
    V3<R> v0 = b - a, v1 = c - a, v2 = p - a;
    R d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    R denom = d00 * d11 - d01 * d01;
    if (std::fabs(denom) <= R(kEpsF)) {
        *wa = *wb = *wc = R(1) / R(3);
        return;
    }
    R v = (d11 * d20 - d01 * d21) / denom;
    R w = (d00 * d21 - d01 * d20) / denom;
    v = std::min(std::max(v, R(0)), R(1));
    w = std::min(std::max(w, R(0)), R(1) - v);
    R u = std::max(R(1) - v - w, R(0));
    R sum = std::max(u + v + w, R(kEpsF));
    *wa = u / sum;
    *wb = v / sum;
    *wc = w / sum;
}

// A41:5458-5505 closest parameters between two segments.
template <class R> void closest_segment_parameters(V3<R> p1, V3<R> q1, V3<R> p2, V3<R> q2, R* s, R* t) {
    const R eps = R(kEpsF);
    V3<R> d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    R a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    if (a <= eps && e <= eps) { *s = R(0); *t = R(0); return; }
    if (a <= eps) { *s = R(0); *t = clamp01(f / std::max(e, eps)); return; }
    R c = dot(d1, r);
    if (e <= eps) { *t = R(0); *s = clamp01(-c / std::max(a, eps)); return; }
    R b = dot(d1, d2);
    R denom = a * e - b * b;
    *s = std::fabs(denom) > eps ? clamp01((b * f - c * e) / denom) : R(0);
    R tnom = b * (*s) + f;
    if (tnom < R(0)) { *t = R(0); *s = clamp01(-c / std::max(a, eps)); }
    else if (tnom > e) { *t = R(1); *s = clamp01((b - c) / std::max(a, eps)); }
    else { *t = tnom / e; }
}

template <class R> V3<R> limit_delta(V3<R> d, R max_len) {  // A41:951-960
    if (!finite(d) || !std::isfinite(max_len) || max_len <= R(0)) return mk<R>(0, 0, 0);
    R len = length(d);
    if (!std::isfinite(len) || len <= R(kEpsF)) return mk<R>(0, 0, 0);
    return len > max_len ? mul(d, max_len / len) : d;
}

template <class R> V3<R> normalize_or(V3<R> v, V3<R> fb) {  // A41:1205-1211
    R len = length(v);
    if (!std::isfinite(len) || len <= R(kEpsF)) return fb;
    return mul(v, R(1) / len);
}

// ------------------------------------------------------------- force fields
// A41:181-214 permutation table used by Blender's improved Perlin noise.
const unsigned char kNoiseHash[512] = {
#define NOISE_ROW_A 0xA2,0xA0,0x19,0x3B,0xF8,0xEB,0xAA,0xEE,0xF3,0x1C,0x67,0x28,0x1D,0xED,0x0,0xDE,0x95,0x2E,0xDC,0x3F,0x3A,0x82,0x35,0x4D,0x6C,0xBA,0x36,0xD0,0xF6,0xC,0x79,0x32,0xD1,0x59,0xF4,0x8,0x8B,0x63,0x89,0x2F,0xB8,0xB4,0x97,0x83,0xF2,0x8F,0x18,0xC7,0x51,0x14,0x65,0x87,0x48,0x20,0x42,0xA8,0x80,0xB5,0x40,0x13,0xB2,0x22,0x7E,0x57
#define NOISE_ROW_B 0xBC,0x7F,0x6B,0x9D,0x86,0x4C,0xC8,0xDB,0x7C,0xD5,0x25,0x4E,0x5A,0x55,0x74,0x50,0xCD,0xB3,0x7A,0xBB,0xC3,0xCB,0xB6,0xE2,0xE4,0xEC,0xFD,0x98,0xB,0x96,0xD3,0x9E,0x5C,0xA1,0x64,0xF1,0x81,0x61,0xE1,0xC4,0x24,0x72,0x49,0x8C,0x90,0x4B,0x84,0x34,0x38,0xAB,0x78,0xCA,0x1F,0x1,0xD7,0x93,0x11,0xC1,0x58,0xA9,0x31,0xF9,0x44,0x6D
#define NOISE_ROW_C 0xBF,0x33,0x9C,0x5F,0x9,0x94,0xA3,0x85,0x6,0xC6,0x9A,0x1E,0x7B,0x46,0x15,0x30,0x27,0x2B,0x1B,0x71,0x3C,0x5B,0xD6,0x6F,0x62,0xAC,0x4F,0xC2,0xC0,0xE,0xB1,0x23,0xA7,0xDF,0x47,0xB0,0x77,0x69,0x5,0xE9,0xE6,0xE7,0x76,0x73,0xF,0xFE,0x6E,0x9B,0x56,0xEF,0x12,0xA5,0x37,0xFC,0xAE,0xD9,0x3,0x8E,0xDD,0x10,0xB9,0xCE,0xC9,0x8D
#define NOISE_ROW_D 0xDA,0x2A,0xBD,0x68,0x17,0x9F,0xBE,0xD4,0xA,0xCC,0xD2,0xE8,0x43,0x3D,0x70,0xB7,0x2,0x7D,0x99,0xD8,0xD,0x60,0x8A,0x4,0x2C,0x3E,0x92,0xE5,0xAF,0x53,0x7,0xE0,0x29,0xA6,0xC5,0xE3,0xF5,0xF7,0x4A,0x41,0x26,0x6A,0x16,0x5E,0x52,0x2D,0x21,0xAD,0xF0,0x91,0xFF,0xEA,0x54,0xFA,0x66,0x1A,0x45,0x39,0xCF,0x75,0xA4,0x88,0xFB,0x5D
    NOISE_ROW_A, NOISE_ROW_B, NOISE_ROW_C, NOISE_ROW_D,
    NOISE_ROW_A, NOISE_ROW_B, NOISE_ROW_C, NOISE_ROW_D
#undef NOISE_ROW_A
#undef NOISE_ROW_B
#undef NOISE_ROW_C
#undef NOISE_ROW_D
};

template <class R> R fade(R t) { return t * t * t * (t * (t * R(6) - R(15)) + R(10)); }
template <class R> R nlerp(R t, R a, R b) { return a + t * (b - a); }
template <class R> R grad(int h, R x, R y, R z) {
    h &= 15;
    R u = h < 8 ? x : y;
    R v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}
template <class R> R perlin(R x, R y, R z) {  // A41:1282-1317
    R fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int X = static_cast<int>(fx) & 255, Y = static_cast<int>(fy) & 255, Z = static_cast<int>(fz) & 255;
    x -= fx; y -= fy; z -= fz;
    R u = fade(x), v = fade(y), w = fade(z);
    int A = kNoiseHash[X] + Y, AA = kNoiseHash[A] + Z, AB = kNoiseHash[A + 1] + Z;
    int B = kNoiseHash[X + 1] + Y, BA = kNoiseHash[B] + Z, BB = kNoiseHash[B + 1] + Z;
    return nlerp(w,
                 nlerp(v, nlerp(u, grad(kNoiseHash[AA], x, y, z), grad(kNoiseHash[BA], x - R(1), y, z)),
                       nlerp(u, grad(kNoiseHash[AB], x, y - R(1), z), grad(kNoiseHash[BB], x - R(1), y - R(1), z))),
                 nlerp(v, nlerp(u, grad(kNoiseHash[AA + 1], x, y, z - R(1)), grad(kNoiseHash[BA + 1], x - R(1), y, z - R(1))),
                       nlerp(u, grad(kNoiseHash[AB + 1], x, y - R(1), z - R(1)),
                             grad(kNoiseHash[BB + 1], x - R(1), y - R(1), z - R(1)))));
}
template <class R> R turbulence(R size, R x, R y, R z) {  // A41:1323-1337
    if (size != R(0) && std::isfinite(size)) {
        R inv = R(1) / size;
        x *= inv; y *= inv; z *= inv;
    }
    R sum = 0, amp = 1, fs = 1;
    for (int i = 0; i <= 2; ++i, amp *= R(0.5), fs *= R(2)) sum += (R(0.5) + R(0.5) * perlin(fs * x, fs * y, fs * z)) * amp;
    return sum * (R(4) / R(7));
}
// sin and pow of the force fields are evaluated in double and rounded once,
// so host and device maths libraries agree on the float result.
template <class R> R sin_r(R x) { return static_cast<R>(std::sin(static_cast<double>(x))); }
template <class R> R pow_r(R x, R y) { return static_cast<R>(std::pow(static_cast<double>(x), static_cast<double>(y))); }
template <class R> R field_noise(V3<R> p, int seed, R salt) {  // A41:1339-1347
    R value = sin_r(p.x * R(12.9898) + p.y * R(78.233) + p.z * R(37.719) + (R(seed) + salt) * R(19.191)) * R(43758.5453);
    return (value - std::floor(value)) * R(2) - R(1);
}
template <class R> R falloff(const SsblXpbdForceField& f, R distance, R radial) {  // A41:1355-1386
    if (!std::isfinite(distance)) return R(0);
    R mn = std::max(R(f.distance_min), R(0)), mx = std::max(R(f.distance_max), R(0));
    R rmn = std::max(R(f.radial_min), R(0)), rmx = std::max(R(f.radial_max), R(0));
    if (f.use_max_distance && mx > R(0) && distance >= mx) return R(0);
    if (f.use_radial_max && rmx > R(0) && radial >= rmx) return R(0);
    R att = 1;
    if (f.use_max_distance && mx > mn && distance > mn) att *= clamp01((mx - distance) / std::max(mx - mn, R(kEpsF)));
    if (f.use_min_distance && f.falloff_power > 0.0f && distance > mn) {
        R ref = std::max(mn, R(1.0e-4));
        att *= pow_r(ref / std::max(distance, ref), R(f.falloff_power));
    }
    if (f.use_radial_max && rmx > rmn && radial > rmn) att *= clamp01((rmx - radial) / std::max(rmx - rmn, R(kEpsF)));
    if (f.use_radial_min && f.radial_falloff > 0.0f && radial > rmn) {
        R ref = std::max(rmn, R(1.0e-4));
        att *= pow_r(ref / std::max(radial, ref), R(f.radial_falloff));
    }
    return std::isfinite(att) ? att : R(0);
}
template <class R> V3<R> limit_accel(V3<R> a) {  // A41:1388-1394
    R len = length(a);
    if (!std::isfinite(len) || len <= R(5000)) return std::isfinite(len) ? a : mk<R>(0, 0, 0);
    return mul(a, R(5000) / std::max(len, R(kEpsF)));
}
template <class R> V3<R> with_flow(const SsblXpbdForceField& f, V3<R> force, V3<R> vel, R fo) {  // A41:1396-1409
    if (f.type != 6 && f.type != 9) {
        R flow = std::isfinite(f.flow) ? R(f.flow) : R(0);
        if (std::fabs(flow) > R(1.0e-6) && fo != R(0)) force = force + mul(vel, -flow * fo);
    }
    return limit_accel(force);
}
template <class R> V3<R> arr3(const float v[3]) { return mk<R>(v[0], v[1], v[2]); }

template <class R> V3<R> eval_field(const SsblXpbdForceField& f, V3<R> p, V3<R> vel, V3<R> n, int has_n) {  // A41:1411-1531
    const V3<R> zero = mk<R>(0, 0, 0);
    if (!std::isfinite(f.strength) || f.strength == 0.0f || f.apply_to_location == 0) return zero;
    V3<R> delta = p - arr3<R>(f.origin);
    V3<R> axis = normalize_or(arr3<R>(f.axis), mk<R>(0, 0, 1));
    V3<R> radial_delta = f.use_2d_force ? delta - mul(axis, dot(delta, axis)) : delta;
    R distance = length(delta), radial_distance = length(radial_delta);
    R fo = falloff<R>(f, distance, radial_distance);
    R strength = R(f.strength) * fo;
    if (strength == R(0) || !std::isfinite(strength)) return zero;
    switch (f.type) {
    case 1: {  // wind
        if (!has_n) return zero;
        V3<R> dir = normalize_or(arr3<R>(f.direction), zero);
        R proj = dot(dir, n);
        if (!std::isfinite(proj) || proj <= R(0)) return zero;
        R mag = strength * proj;
        if (f.noise > 0.0f) mag *= std::max(R(0), R(1) + R(f.noise) * field_noise(p, f.seed, R(0)));
        return with_flow(f, mul(dir, mag), vel, fo);
    }
    case 2: case 5: {  // force / charge
        V3<R> sd = f.use_2d_force ? radial_delta : delta;
        R sdist = f.use_2d_force ? radial_distance : distance;
        if (sdist <= R(1.0e-6)) return zero;
        R scale = strength / std::max(sdist, R(kEpsF));
        if (f.type == 5) scale /= std::max(sdist * sdist, R(0.01));
        return with_flow(f, mul(sd, scale), vel, fo);
    }
    case 3: {  // vortex
        V3<R> radial = delta - mul(axis, dot(delta, axis));
        if (dot(radial, radial) <= R(1.0e-10)) return zero;
        V3<R> tangent = normalize_or(cross(axis, radial), zero);
        return with_flow(f, mul(tangent, strength), vel, fo);
    }
    case 6: {  // harmonic
        R sdist = f.use_2d_force ? radial_distance : distance;
        V3<R> sd = f.use_2d_force ? radial_delta : delta;
        if (sdist <= R(1.0e-6)) return zero;
        V3<R> dir = mul(sd, R(1) / std::max(sdist, R(kEpsF)));
        R rest = std::max(R(f.rest_length), R(0));
        R spring = -strength * (sdist - rest);
        R damping = -dot(vel, dir) * std::max(R(f.harmonic_damping), R(0)) * std::fabs(strength);
        return with_flow(f, mul(dir, spring + damping), vel, fo);
    }
    case 7: {  // Lennard-Jones
        if (distance <= R(1.0e-6)) return zero;
        V3<R> dir = mul(delta, R(1) / std::max(distance, R(kEpsF)));
        R radius = std::max(R(f.rest_length), std::max(R(f.size), R(0.1)));
        R ratio = std::min(radius / std::max(distance, R(1.0e-3)), R(6));
        R r2 = ratio * ratio, r6 = r2 * r2 * r2;
        return with_flow(f, mul(dir, strength * (r6 * r6 - r6)), vel, fo);
    }
    case 8: {  // magnet
        V3<R> ax = normalize_or(arr3<R>(f.direction), zero);
        return with_flow(f, mul(cross(vel, ax), strength), vel, fo);
    }
    case 9: {  // drag
        R speed = length(vel);
        if (speed <= R(1.0e-6)) return zero;
        R lin = std::max(R(f.linear_drag), R(0)), quad = std::max(R(f.quadratic_drag), R(0));
        if (lin <= R(0) && quad <= R(0)) lin = std::fabs(strength);
        return with_flow(f, mul(vel, -(lin + quad * speed)), vel, fo);
    }
    case 4: {  // turbulence
        V3<R> q = f.use_global_coords ? p : delta + axis;
        V3<R> nv = mk(R(-1) + R(2) * turbulence(R(f.size), q.x, q.y, q.z), R(-1) + R(2) * turbulence(R(f.size), q.y, q.z, q.x),
                      R(-1) + R(2) * turbulence(R(f.size), q.z, q.x, q.y));
        return with_flow(f, mul(nv, strength), vel, fo);
    }
    case 10: {  // texture
        R freq = std::max(std::max(R(f.noise), R(f.texture_nabla)), R(0.25));
        if (f.size > 1.0e-6f) freq = std::max(freq, R(1) / R(f.size));
        V3<R> q = mul(delta, freq);
        V3<R> nv = mk(field_noise(q, f.seed, R(0)), field_noise(q, f.seed, R(7)), field_noise(q, f.seed, R(13)));
        return with_flow(f, mul(nv, strength), vel, fo);
    }
    default:
        return zero;
    }
}

// ------------------------------------------------------- dynamic colliders
// Triangle soups and particles that other objects (other cloths through the
// global dynamic scene, or animated colliders) present to a cloth, with their
// spatial hashes: uploads A41:8319-8420 and 8575-8697, bounds and hashes
// A41:3694-3932, 4325-4352 and 6220-6449, global scene A41:6597-6782.
// Bucket contents are kept in ascending primitive order (the reference's
// atomic fills leave that order to scheduling); a particle bucket shows its
// first 16 particles (A41:4345-4350).
constexpr int kParticleHashMinCount = 128;       // A41:25
constexpr int kParticleBucketSlots = 16;         // A41:26
constexpr int kMinParticleBuckets = 1024;        // A41:27
constexpr int kTriangleHashMinCount = 64;        // A41:39
constexpr int kMinTriangleBuckets = 1024;        // A41:40
constexpr int kDynamicMaxHashCells = 64;         // A41:79
constexpr float kNeighborImpulseScale = 0.5f;    // A41:129
// signed (indexed) dynamic triangles, A41:80-83, 130-137
constexpr int kSignedInsidePasses = 6;
constexpr float kSignedQueryMargin = 0.04f;
constexpr float kSignedMaxQueryMargin = 0.08f;
constexpr float kSignedInsideTargetMargin = -0.001f;
constexpr float kSignedRelaxation = 1.0f / 3.0f;
constexpr float kSignedOutwardRetention = 1.0f / 8.0f;
constexpr float kSignedContactMarginScale = 0.75f;
constexpr float kSignedInsideRelaxation = 1.0f / 3.0f;
constexpr float kSignedInsideMaxDeltaScale = 16.0f;
constexpr float kSignedInsideMaxDeltaFloor = 3.0e-3f;
constexpr float kDynamicMaxDeltaScale = 3.0f;
constexpr float kDynamicMaxDeltaFloor = 5.0e-4f;
constexpr int kCrossOff = 0, kCrossLower = 2;    // A41:167-169
enum DynCountSlot {
    kCountDynParticleCandidates = 4, kCountDynParticleContacts = 5, kCountDynParticleOverflow = 6, kCountTrianglePairs = 7,
    kCountDynTriCandidates = 34, kCountDynTriBucketOverflow = 35, kCountDynTriLarge = 36, kCountDynTriAabbRejects = 37
};

int next_pow2_int(int v) {
    int r = 1;
    while (r < v && r < (1 << 30)) r <<= 1;
    return r;
}

int cell_coord_f(float value, float cell) { return cell_coord<float>(value, cell); }

// A41:6590-6601
float dynamic_triangle_cell_size(float max_extent, float margin, float thickness, bool fine) {
    const float coarse = std::max(std::max(max_extent + margin * 2.0f, thickness), 1.0e-3f);
    if (!fine) return coarse;
    const float fine_cell = std::max(std::max(thickness * 1.5f, margin * 4.0f), 0.02f);
    return std::max(std::min(coarse, fine_cell), 1.0e-3f);
}

struct DynamicSet {
    // triangles (9 floats each)
    int nt = 0;
    std::vector<float> tv;
    std::vector<int> tsrc, tlayer;           // empty: no source filter
    float tri_cell = 0.0f, span_margin = 0.0f;
    bool tri_hash = false;
    int tri_buckets = 0;
    std::vector<float> lo, hi, ctr, rad;     // 3T, 3T, 3T, T
    std::vector<int> cmin, cmax;             // 3T (hashed sets only)
    std::vector<int> tri_off, tri_idx;       // buckets + 1, entries ascending per bucket
    std::vector<int> large;                  // ascending
    int max_occupancy = 0;
    // particles
    int np = 0;
    std::vector<float> pp, pprev, pr;        // 3P, 3P, P
    std::vector<int> psrc, player;
    std::vector<float> radii_seen;
    float part_cell = 0.0f;
    bool part_hash = false;
    int part_buckets = 0;
    std::vector<int> pcell;                  // 3P
    std::vector<int> pcount, pidx;           // buckets, buckets * 16
    // signed (indexed) triangles: A41:8433-8590
    bool signed_set = false;
    bool tri_motion = false, part_motion = false;   // previous centres / positions valid for one step
    std::vector<float> prev_ctr;             // 3T

    // A41:8338-8420 (soup) and 6692-6722 (scene): cell size from the largest extent
    void set_triangles(const float* t, int n, const int* src, const int* layer, float margin, float thickness, bool fine) {
        nt = std::max(n, 0);
        tv.assign(t, t + 9 * static_cast<size_t>(nt));
        tsrc.clear();
        tlayer.clear();
        if (src) tsrc.assign(src, src + nt);
        if (layer) tlayer.assign(layer, layer + nt);
        float max_extent = 0.0f;
        for (int k = 0; k < nt; ++k) {
            const float* T = &tv[9 * static_cast<size_t>(k)];
            const float ext = std::max(std::max(T[0], std::max(T[3], T[6])) - std::min(T[0], std::min(T[3], T[6])),
                                       std::max(std::max(T[1], std::max(T[4], T[7])) - std::min(T[1], std::min(T[4], T[7])),
                                                std::max(T[2], std::max(T[5], T[8])) - std::min(T[2], std::min(T[5], T[8]))));
            if (std::isfinite(ext) && ext >= 0.0f) max_extent = std::max(max_extent, ext);
        }
        tri_cell = dynamic_triangle_cell_size(max_extent, std::max(margin, 0.0f), std::max(thickness, 0.0f), fine);
        span_margin = std::max(margin, 0.0f);
        signed_set = false;
        tri_motion = false;
        compute_bounds();
    }
    // A41:3694-3718: bounds, centre and bounding radius of every triangle
    void compute_bounds() {
        lo.assign(3 * static_cast<size_t>(nt), 0.0f);
        hi = lo;
        ctr = lo;
        rad.assign(nt, 0.0f);
        for (int k = 0; k < nt; ++k) {
            const float* T = &tv[9 * static_cast<size_t>(k)];
            const V3<float> a = mk(T[0], T[1], T[2]), b = mk(T[3], T[4], T[5]), c = mk(T[6], T[7], T[8]);
            const V3<float> cen = mul((a + b) + c, 1.0f / 3.0f);
            lo[3 * k] = std::min(a.x, std::min(b.x, c.x)); lo[3 * k + 1] = std::min(a.y, std::min(b.y, c.y));
            lo[3 * k + 2] = std::min(a.z, std::min(b.z, c.z));
            hi[3 * k] = std::max(a.x, std::max(b.x, c.x)); hi[3 * k + 1] = std::max(a.y, std::max(b.y, c.y));
            hi[3 * k + 2] = std::max(a.z, std::max(b.z, c.z));
            ctr[3 * k] = cen.x; ctr[3 * k + 1] = cen.y; ctr[3 * k + 2] = cen.z;
            rad[k] = std::sqrt(std::max(std::max(dot(a - cen, a - cen), dot(b - cen, b - cen)), dot(c - cen, c - cen)));
        }
    }
    // A41:8433-8590: corners gathered from indexed vertices (triangles with an
    // index out of range collapse to the origin), the fine cell from the
    // largest extent, and the centres of the previous upload kept as the
    // previous centres when the topology and count are unchanged.
    void set_indexed(const float* verts, int nverts, const int* idx, int ntri, bool topology_changed, float margin, float thickness) {
        nverts = std::max(nverts, 0);
        ntri = std::max(ntri, 0);
        if (nverts == 0 || ntri == 0) {
            clear_triangles();
            return;
        }
        float max_extent = 0.0f;
        std::vector<float> soup(9 * static_cast<size_t>(ntri), 0.0f);
        for (int t = 0; t < ntri; ++t) {
            const int ia = idx[3 * t], ib = idx[3 * t + 1], ic = idx[3 * t + 2];
            if (ia < 0 || ib < 0 || ic < 0 || ia >= nverts || ib >= nverts || ic >= nverts) continue;
            const int id[3] = {ia, ib, ic};
            for (int j = 0; j < 3; ++j)
                for (int a = 0; a < 3; ++a) soup[9 * static_cast<size_t>(t) + 3 * j + a] = verts[3 * id[j] + a];
            const float* T = &soup[9 * static_cast<size_t>(t)];
            const float ext = std::max(std::max(T[0], std::max(T[3], T[6])) - std::min(T[0], std::min(T[3], T[6])),
                                       std::max(std::max(T[1], std::max(T[4], T[7])) - std::min(T[1], std::min(T[4], T[7])),
                                                std::max(T[2], std::max(T[5], T[8])) - std::min(T[2], std::min(T[5], T[8]))));
            if (std::isfinite(ext) && ext >= 0.0f) max_extent = std::max(max_extent, ext);
        }
        const bool had_previous = !topology_changed && nt == ntri && ctr.size() == 3 * static_cast<size_t>(ntri);
        const std::vector<float> old_ctr = ctr;
        nt = ntri;
        tv = soup;
        tsrc.clear();
        tlayer.clear();
        tri_cell = dynamic_triangle_cell_size(max_extent, std::max(margin, 0.0f), std::max(thickness, 0.0f), true);
        span_margin = std::max(margin, 0.0f);
        compute_bounds();
        prev_ctr = had_previous ? old_ctr : ctr;
        tri_motion = had_previous;
        signed_set = true;
    }
    // A41:8575-8697 (per cloth) and 6620-6668 (scene)
    void set_particles(const float* p, const float* prev, const float* r, const int* src, const int* layer, int n, float margin,
                       float thickness) {
        n = std::max(n, 0);
        std::vector<float> clean(n);
        float max_radius = 0.0f;
        for (int i = 0; i < n; ++i) {
            float v = std::isfinite(r[i]) ? r[i] : 0.0f;
            clean[i] = std::max(v, 0.0f);
            max_radius = std::max(max_radius, clean[i]);
        }
        if (part_cell <= 0.0f || clean != radii_seen || np != n) {
            part_cell = std::max(std::max(max_radius + std::max(margin, 0.0f), std::max(thickness, 0.0f)), 1.0e-3f);
            radii_seen = clean;
        }
        const bool had_previous = np == n && !pp.empty();
        std::vector<float> old = pp;
        part_motion = had_previous;   // A41:8625-8690
        np = n;
        pp.assign(p, p + 3 * static_cast<size_t>(n));
        if (prev) pprev.assign(prev, prev + 3 * static_cast<size_t>(n));
        else pprev = had_previous ? old : pp;
        pr.assign(r, r + n);
        psrc.clear();
        player.clear();
        if (src) psrc.assign(src, src + n);
        if (layer) player.assign(layer, layer + n);
    }
    void clear_triangles() {
        nt = 0; tv.clear(); tsrc.clear(); tlayer.clear(); tri_hash = false; large.clear(); max_occupancy = 0;
        ctr.clear(); signed_set = false; tri_motion = false;
    }
    void clear_particles() { np = 0; pp.clear(); pprev.clear(); pr.clear(); psrc.clear(); player.clear(); part_hash = false; part_motion = false; }

    // A41:3694-3760 bounds, 3762-3932 hash (compact), 4325-4352 particle hash
    void build(unsigned long long* counts) {
        lo.assign(3 * static_cast<size_t>(nt), 0.0f);
        hi = lo;
        ctr = lo;
        rad.assign(nt, 0.0f);
        tri_hash = g_dynamic_hash && nt >= kTriangleHashMinCount;
        cmin.clear();
        cmax.clear();
        large.clear();
        max_occupancy = 0;
        const float cs = std::max(tri_cell, 1.0e-3f);
        if (tri_hash) {
            cmin.assign(3 * static_cast<size_t>(nt), 0);
            cmax = cmin;
            tri_buckets = next_pow2_int(std::max(kMinTriangleBuckets, nt * 4));
        }
        std::vector<std::vector<int>> lists(tri_hash ? tri_buckets : 0);
        for (int k = 0; k < nt; ++k) {
            const float* T = &tv[9 * static_cast<size_t>(k)];
            const V3<float> a = mk(T[0], T[1], T[2]), b = mk(T[3], T[4], T[5]), c = mk(T[6], T[7], T[8]);
            const V3<float> mn = mk(std::min(a.x, std::min(b.x, c.x)), std::min(a.y, std::min(b.y, c.y)), std::min(a.z, std::min(b.z, c.z)));
            const V3<float> mx = mk(std::max(a.x, std::max(b.x, c.x)), std::max(a.y, std::max(b.y, c.y)), std::max(a.z, std::max(b.z, c.z)));
            const V3<float> cen = mul((a + b) + c, 1.0f / 3.0f);
            const float r = std::sqrt(std::max(std::max(dot(a - cen, a - cen), dot(b - cen, b - cen)), dot(c - cen, c - cen)));
            lo[3 * k] = mn.x; lo[3 * k + 1] = mn.y; lo[3 * k + 2] = mn.z;
            hi[3 * k] = mx.x; hi[3 * k + 1] = mx.y; hi[3 * k + 2] = mx.z;
            ctr[3 * k] = cen.x; ctr[3 * k + 1] = cen.y; ctr[3 * k + 2] = cen.z;
            rad[k] = r;
            if (!tri_hash) continue;
            const float m = span_margin;
            int x0 = cell_coord_f(mn.x - m, cs), y0 = cell_coord_f(mn.y - m, cs), z0 = cell_coord_f(mn.z - m, cs);
            int x1 = cell_coord_f(mx.x + m, cs), y1 = cell_coord_f(mx.y + m, cs), z1 = cell_coord_f(mx.z + m, cs);
            const long long sx = 1LL * x1 - x0 + 1, sy = 1LL * y1 - y0 + 1, sz = 1LL * z1 - z0 + 1;
            long long cells = (sx > 0 && sy > 0 && sz > 0) ? sx * sy * sz : 0;
            if (cells <= 0) {
                x0 = x1 = cell_coord_f(cen.x, cs);
                y0 = y1 = cell_coord_f(cen.y, cs);
                z0 = z1 = cell_coord_f(cen.z, cs);
                cells = 1;
            }
            cmin[3 * k] = x0; cmin[3 * k + 1] = y0; cmin[3 * k + 2] = z0;
            cmax[3 * k] = x1; cmax[3 * k + 1] = y1; cmax[3 * k + 2] = z1;
            if (cells > kDynamicMaxHashCells) {
                counts[kCountDynTriLarge] += 1;
                large.push_back(k);
                continue;
            }
            for (int z = z0; z <= z1; ++z)
                for (int y = y0; y <= y1; ++y)
                    for (int x = x0; x <= x1; ++x) lists[hash_cell(x, y, z, tri_buckets)].push_back(k);
        }
        tri_off.assign(tri_hash ? tri_buckets + 1 : 0, 0);
        tri_idx.clear();
        for (size_t bk = 0; bk < lists.size(); ++bk) {
            tri_off[bk] = static_cast<int>(tri_idx.size());
            tri_idx.insert(tri_idx.end(), lists[bk].begin(), lists[bk].end());
            max_occupancy = std::max(max_occupancy, static_cast<int>(lists[bk].size()));
        }
        if (tri_hash) tri_off[tri_buckets] = static_cast<int>(tri_idx.size());
        // particles
        part_hash = g_dynamic_hash && np >= kParticleHashMinCount;
        if (!part_hash) return;
        part_buckets = next_pow2_int(std::max(kMinParticleBuckets, np * 4));
        pcell.assign(3 * static_cast<size_t>(np), 0);
        pcount.assign(part_buckets, 0);
        pidx.assign(static_cast<size_t>(part_buckets) * kParticleBucketSlots, -1);
        for (int i = 0; i < np; ++i) {
            const int cx = cell_coord_f(pp[3 * i], part_cell), cy = cell_coord_f(pp[3 * i + 1], part_cell),
                      cz = cell_coord_f(pp[3 * i + 2], part_cell);
            pcell[3 * i] = cx; pcell[3 * i + 1] = cy; pcell[3 * i + 2] = cz;
            const unsigned bk = hash_cell(cx, cy, cz, part_buckets);
            const int slot = pcount[bk]++;
            if (slot < kParticleBucketSlots) pidx[bk * kParticleBucketSlots + slot] = i;
            else counts[kCountDynParticleOverflow] += 1;
        }
    }
};

// A global dynamic scene plus the counters its builds write (A41:309-340).
struct Scene {
    DynamicSet set;
    float margin = 0.0f, thickness = 0.0f;
    unsigned long long counts[kCountSlots] = {};
    long long hash_overflow = 0;
};

// ------------------------------------------------------------------ solver
struct Pin { int index; float target[3]; float weight; };

template <class R> struct Solver {
    using Vec = V3<R>;
    SsblXpbdConfig cfg{};
    SsblXpbdRuntimeColliders colliders{};
    int stretch_mode = SSBL_EXT_STRETCH_JACOBI;
    int pcg_cadence = 1;   // SSBL_ABI41_PCG_SUBSTEP_CADENCE when set above 0 (A41:284-292)
    std::vector<Vec> pos, prev, vel, rest;
    std::vector<float> inv_mass;       // as uploaded
    std::vector<unsigned char> pinned; // inv_mass <= 0 (A41:8861-8865)
    struct Edge { int i, j; float rest; };
    std::vector<Edge> springs, bends, lras;
    std::vector<int> edge_colors, bend_colors;   // host colour offsets (empty when none)
    bool bend_coloring_valid = false;
    std::vector<int> tris;              // 3T
    long long wing_count = 0;           // dihedral "bending wings" (diagnostics only, A41:7193-7302)
    std::vector<int> tri_offsets, tri_incident;  // vertex -> incident triangles (A41:7051-7109)
    std::vector<float> area_share, rest_scale;
    float pressure_accel_scale = 1792.0f, pressure_impulse_rest_scale = 0.85f;
    float pressure_velocity_rest_scale = 2.25f, pressure_velocity_length_scale = 1.0f;
    std::vector<Pin> pins;
    std::vector<SsblXpbdForceField> fields;
    int unsupported_fields = 0;
    // static SDF
    std::vector<Vec> static_tris;
    std::vector<R> sdf;
    int sdf_ready = 0, sdf_dim[3] = {0, 0, 0}, sdf_signed = 0;
    Vec sdf_origin{};
    R sdf_voxel = 0;
    long long sdf_rebuilds = 0, sdf_unsigned = 0;
    float sdf_build_ms = 0.0f;
    // self collision
    std::vector<int> self_cells;              // 3V from the last vertex-hash build
    std::vector<unsigned> nb_counts, nb_index;  // V, V*32
    std::vector<Vec> snap_pos, snap_prev;     // positions the last candidate table was built from
    // hard-stretch PCG (A41:3032-3464, 7304-7394, 7626-7759)
    struct Sym { R m11, m12, m13, m22, m23, m33; };
    bool edge_coloring_valid = false;
    std::vector<int> inc_off, inc_edge;       // vertex -> incident valid edges, ascending
    std::vector<Sym> pcg_diag, pcg_pinv, pcg_off;
    std::vector<Vec> pcg_rhs, pcg_sol, pcg_res, pcg_z, pcg_dir, pcg_adir;
    double pcg_rz = 0, pcg_dad = 0, pcg_rznext = 0, pcg_initial = 0;
    int pcg_status = 0;
    long long pcg_guard = 0, pcg_runs = 0;
    float pcg_max_delta = 0.0f;
    std::vector<int64_t> acc_q;               // 3V fixed point
    std::vector<int> acc_w;
    float max_smoothed = 0.0f;
    unsigned long long counts[kCountSlots] = {};
    SsblXpbdDiagnostics diag{};
    // dynamic colliders: the cloth's own, or an attached global scene
    DynamicSet dyn;
    bool dyn_tri_allocated = false;
    Scene* scene = nullptr;
    int target_src = 0, target_layer = 0, cross_mode = kCrossOff;
    bool use_scene() const { return scene != nullptr && cross_mode != kCrossOff; }

    void count(int slot, unsigned long long n = 1) { counts[slot] += n; }
    bool w_zero(int v) const { return pinned[v] != 0; }
    R weight(int v) const { return pinned[v] ? R(0) : R(inv_mass[v]); }
    int V() const { return cfg.vertex_count; }

    // ------------------------------------------------------------- create
    bool create(const SsblXpbdConfig* config, const SsblXpbdMesh* mesh);
    bool build_sdf(const float* triangles, int count);
    bool set_sdf_grid(const int* dims, const float* ov, const float* values, int count);
    void pressure_setup(const std::vector<float>& rest_f);

    // -------------------------------------------------------------- step
    bool step(int substeps, int iterations, int fetch);
    void integrate(R dt);
    void pin_project(R exponent, R dt);
    void stretch_jacobi(R dt);
    void stretch_colored(R dt);
    void lra_jacobi(R dt);
    void bend_pass(R dt);
    bool bend_one(const Edge& b, R dt, const std::vector<Vec>& P, Vec* di, Vec* dj);
    void analytic(R dt);
    void static_sdf_contacts(R dt);
    void velocity_update(R dt);
    void self_collision();
    template <class H> void island_aggregate(const H& TH, R cs);
    void self_accumulate(int v, V3<R> d);
    void dynamic_contacts(R dt);
    void signed_contacts(R dt, bool inside);
    void signed_response(int i, V3<R> orig, V3<R> raw, V3<R> cvel, R proxy_depth, R dt, bool inside);
    bool contact_response(Vec n, R proxy, Vec* vel_io);
    // helpers
    Vec normal_at(int v, const std::vector<Vec>& P, int* has) const;
    void populate_sdf_diag() {  // A41:6104-6121
        diag.static_sdf_rebuild_count = sdf_rebuilds;
        diag.static_sdf_voxel_count = sdf_ready ? 1LL * sdf_dim[0] * sdf_dim[1] * sdf_dim[2] : 0;
        diag.static_sdf_grid_x = sdf_dim[0];
        diag.static_sdf_grid_y = sdf_dim[1];
        diag.static_sdf_grid_z = sdf_dim[2];
        diag.static_sdf_build_ms = sdf_build_ms;
        diag.static_sdf_unsigned_fallback_count = sdf_unsigned;
        diag.static_triangle_count = cfg.static_triangle_count;
    }
    void pcg_run(R dt);
    void polish(R dt);
    void final_caps(R dt);
    void cap_ranges(int kind, int passes, bool reverse_too, R dt);
    void edge_jacobi(int kind, int lo, int hi, R dt);
    void gather_average(int kind, R dt);
    R self_radius() const {
        R req = std::max(R(cfg.self_collision_distance), R(0));
        if (req > R(0)) return req;
        return std::max(std::max(R(cfg.cloth_thickness), R(cfg.collision_margin)), R(0));
    }
    bool sdf_sample(Vec p, R* out) const;
    R sdf_sample_clamped(Vec p) const;
    Vec sdf_normal(Vec p) const;
};

template <class R> bool Solver<R>::create(const SsblXpbdConfig* config, const SsblXpbdMesh* mesh) {
    cfg = *config;  // A41:8797-8848 sanitising
    cfg.vertex_count = std::max(cfg.vertex_count, 0);
    cfg.edge_count = std::max(cfg.edge_count, 0);
    cfg.bend_count = std::max(cfg.bend_count, 0);
    cfg.lra_count = std::max(cfg.lra_count, 0);
    cfg.triangle_count = std::max(cfg.triangle_count, 0);
    cfg.damping = std::isfinite(cfg.damping) ? cfg.damping : 1.0f;
    cfg.damping = std::max(0.0f, std::min(cfg.damping, 1.0f));
    if (!std::isfinite(cfg.bend_compliance) || cfg.bend_compliance < 0.0f) cfg.bend_compliance = 0.0f;
    if (!std::isfinite(cfg.lra_compliance) || cfg.lra_compliance < 0.0f) cfg.lra_compliance = 0.0f;
    if (!std::isfinite(cfg.stretch_optimization_strength)) cfg.stretch_optimization_strength = 0.0f;
    cfg.stretch_optimization_strength = std::min(std::max(cfg.stretch_optimization_strength, 0.0f), 1.0f);
    cfg.stretch_optimization_enabled =
        (cfg.edge_count > 0 && cfg.stretch_optimization_enabled && cfg.stretch_optimization_strength > 0.0f) ? 1 : 0;
    colliders.use_ground = cfg.use_ground;
    colliders.ground_height = cfg.ground_height;
    colliders.use_wall = cfg.use_wall;
    std::memcpy(colliders.wall_origin, cfg.wall_origin, sizeof(float) * 3);
    std::memcpy(colliders.wall_normal, cfg.wall_normal, sizeof(float) * 3);
    colliders.use_sphere = cfg.use_sphere;
    std::memcpy(colliders.sphere_center, cfg.sphere_center, sizeof(float) * 3);
    colliders.sphere_radius = cfg.sphere_radius;

    const int n = cfg.vertex_count;
    std::vector<float> rest_f(mesh->positions, mesh->positions + 3 * static_cast<size_t>(n));
    pos.resize(n); prev.resize(n); vel.assign(n, mk<R>(0, 0, 0)); rest.resize(n);
    inv_mass.assign(mesh->inv_mass, mesh->inv_mass + n);
    pinned.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        rest[i] = mk<R>(rest_f[3 * i], rest_f[3 * i + 1], rest_f[3 * i + 2]);
        pos[i] = prev[i] = rest[i];
        if (inv_mass[i] <= 0.0f) pinned[i] = 1;
    }
    springs.clear();
    for (int e = 0; e < cfg.edge_count; ++e) {  // A41:8866-8873
        int i = mesh->edges ? mesh->edges[2 * e] : 0, j = mesh->edges ? mesh->edges[2 * e + 1] : 0;
        float r = mesh->edge_rest_lengths ? mesh->edge_rest_lengths[e] : 0.0f;
        springs.push_back({std::max(i, 0), std::max(j, 0), r});
    }
    if (cfg.bend_count > 0 && (!mesh->bends || !mesh->bend_rest_lengths))
        return fail("ABI41 bend constraints require bend pairs and rest lengths");
    if (cfg.lra_count > 0 && (!mesh->lra_edges || !mesh->lra_rest_lengths))
        return fail("ABI41 LRA constraints require edge pairs and rest lengths");
    bends.clear();
    for (int b = 0; b < cfg.bend_count; ++b) bends.push_back({mesh->bends[2 * b], mesh->bends[2 * b + 1], mesh->bend_rest_lengths[b]});
    lras.clear();
    for (int b = 0; b < cfg.lra_count; ++b) lras.push_back({mesh->lra_edges[2 * b], mesh->lra_edges[2 * b + 1], mesh->lra_rest_lengths[b]});
    edge_colors.clear();
    if (cfg.edge_color_count > 0) {
        if (!mesh->edge_color_offsets) return fail("ABI41 edge color offsets are required");
        edge_colors.assign(mesh->edge_color_offsets, mesh->edge_color_offsets + cfg.edge_color_count + 1);
    }
    bend_colors.clear();
    if (cfg.bend_color_count > 0) {
        if (!mesh->bend_color_offsets) return fail("ABI41 bend color offsets are required");
        bend_colors.assign(mesh->bend_color_offsets, mesh->bend_color_offsets + cfg.bend_color_count + 1);
    }
    // A coloured bend pass is only a well-defined Gauss-Seidel sweep when no
    // colour touches a vertex twice; otherwise the bends run as a Jacobi sweep.
    bend_coloring_valid = !bend_colors.empty();
    if (bend_coloring_valid) {
        std::vector<int> stamp(n, -1);
        for (int c = 0; c < cfg.bend_color_count && bend_coloring_valid; ++c) {
            int lo = std::max(bend_colors[c], 0), hi = std::min(bend_colors[c + 1], cfg.bend_count);
            for (int b = lo; b < hi; ++b) {
                int i = bends[b].i, j = bends[b].j;
                if (i < 0 || j < 0 || i >= n || j >= n) continue;
                if (stamp[i] == c || stamp[j] == c) { bend_coloring_valid = false; break; }
                stamp[i] = c; stamp[j] = c;
            }
        }
    }
    // incident valid edges per vertex, ascending (the PCG CSR row order, A41:7304-7394)
    inc_off.assign(n + 1, 0);
    for (const Edge& e : springs) if (e.i < n && e.j < n && e.i != e.j) { inc_off[e.i + 1]++; inc_off[e.j + 1]++; }
    for (int i = 0; i < n; ++i) inc_off[i + 1] += inc_off[i];
    inc_edge.assign(inc_off[n], 0);
    {
        std::vector<int> cur(inc_off.begin(), inc_off.end() - 1);
        for (int e = 0; e < cfg.edge_count; ++e) {
            const Edge& E = springs[e];
            if (E.i < n && E.j < n && E.i != E.j) { inc_edge[cur[E.i]++] = e; inc_edge[cur[E.j]++] = e; }
        }
    }
    edge_coloring_valid = !edge_colors.empty();
    if (edge_coloring_valid) {
        std::vector<int> stamp(n, -1);
        for (int c = 0; c < cfg.edge_color_count && edge_coloring_valid; ++c) {
            const int lo = std::max(edge_colors[c], 0), hi = std::min(edge_colors[c + 1], cfg.edge_count);
            for (int e = lo; e < hi; ++e) {
                const int i = springs[e].i, j = springs[e].j;
                if (i >= n || j >= n) continue;
                if (stamp[i] == c || stamp[j] == c) { edge_coloring_valid = false; break; }
                stamp[i] = c; stamp[j] = c;
            }
        }
    }
    if (cfg.stretch_optimization_enabled) {
        if (inc_off[n] == 0) return fail("PCG stretch optimization found no valid CSR edges");
        pcg_diag.assign(n, Sym{}); pcg_pinv.assign(n, Sym{}); pcg_off.assign(inc_off[n], Sym{});
        for (auto* v : {&pcg_rhs, &pcg_sol, &pcg_res, &pcg_z, &pcg_dir, &pcg_adir}) v->assign(n, mk<R>(0, 0, 0));
    }
    tris.assign(3 * static_cast<size_t>(cfg.triangle_count), 0);
    for (int t = 0; t < 3 * cfg.triangle_count; ++t) tris[t] = mesh->triangles ? std::max(mesh->triangles[t], 0) : 0;
    // A41:7193-7302: an edge seen again with a different opposite vertex is a
    // bending wing; the reference builds them for diagnostics only.
    wing_count = 0;
    if (cfg.triangle_count > 1) {
        std::unordered_map<uint64_t, int> owner;
        for (int t = 0; t < cfg.triangle_count; ++t) {
            const int ids[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
            if (ids[0] >= n || ids[1] >= n || ids[2] >= n) continue;
            for (int k = 0; k < 3; ++k) {
                const int x = ids[k], y = ids[(k + 1) % 3], opp = ids[(k + 2) % 3];
                const uint64_t key = (uint64_t(unsigned(std::min(x, y))) << 32) | uint64_t(unsigned(std::max(x, y)));
                auto it = owner.find(key);
                if (it == owner.end()) { owner.emplace(key, opp); continue; }
                if (it->second != opp) ++wing_count;
            }
        }
    }
    // incident-triangle CSR, A41:7051-7109
    tri_offsets.assign(n + 1, 0);
    for (int k = 0; k < 3 * cfg.triangle_count; ++k) if (tris[k] < n) tri_offsets[tris[k] + 1]++;
    for (int i = 0; i < n; ++i) tri_offsets[i + 1] += tri_offsets[i];
    tri_incident.assign(tri_offsets[n], -1);
    {
        std::vector<int> cur(tri_offsets.begin(), tri_offsets.end() - 1);
        for (int t = 0; t < cfg.triangle_count; ++t)
            for (int k = 0; k < 3; ++k) { int v = tris[3 * t + k]; if (v < n) tri_incident[cur[v]++] = t; }
    }
    auto envf = [](const char* name, float def) {
        const char* raw = std::getenv(name);
        if (!raw || !raw[0]) return def;
        char* end = nullptr;
        float v = std::strtof(raw, &end);
        return (end == raw || !std::isfinite(v)) ? def : v;
    };
    // A41:266-292: an integer after strtol, used when above 0, default cadence 1
    if (const char* raw = std::getenv("SSBL_ABI41_PCG_SUBSTEP_CADENCE")) {
        char* end = nullptr;
        const long v = std::strtol(raw, &end, 10);
        pcg_cadence = (end != raw && v > 0 && v <= std::numeric_limits<int>::max()) ? static_cast<int>(v) : 1;
    }
    pressure_accel_scale = std::max(1.0f, std::min(envf("SSBL_ABI41_PRESSURE_UI_ACCEL_SCALE", 1792.0f), 8192.0f));
    pressure_impulse_rest_scale = std::max(0.05f, std::min(envf("SSBL_ABI41_PRESSURE_IMPULSE_REST_SCALE", 0.85f), 8.0f));
    pressure_velocity_rest_scale = std::max(0.10f, std::min(envf("SSBL_ABI41_PRESSURE_VELOCITY_REST_SCALE", 2.25f), 16.0f));
    pressure_setup(rest_f);
    self_cells.assign(3 * static_cast<size_t>(n), 0);
    nb_counts.assign(n, 0);
    nb_index.assign(static_cast<size_t>(n) * kNeighborSlots, 0);
    acc_q.assign(3 * static_cast<size_t>(n), 0);
    acc_w.assign(n, 0);
    diag.finite_flag = 1;
    const int stc = std::max(cfg.static_triangle_count, 0);
    if (stc > 0 && !mesh->static_triangles) return fail("static triangle data is required when triangle count is nonzero");
    return build_sdf(mesh->static_triangles, stc);
}

// A41:7413-7584 per-vertex pressure area shares, rest scales and the
// velocity length scale (host set-up; the total area sums in FP64).
template <class R> void Solver<R>::pressure_setup(const std::vector<float>& rest_f) {
    const int n = cfg.vertex_count;
    using VF = V3<float>;
    auto P = [&](int i) { return mk<float>(rest_f[3 * i], rest_f[3 * i + 1], rest_f[3 * i + 2]); };
    area_share.assign(n, 0.0f);
    double total_area = 0.0;
    int valid = 0;
    for (int t = 0; t < cfg.triangle_count; ++t) {
        int i0 = tris[3 * t], i1 = tris[3 * t + 1], i2 = tris[3 * t + 2];
        if (i0 >= n || i1 >= n || i2 >= n) continue;
        VF a = P(i0), b = P(i1), c = P(i2);
        VF av = cross(b - a, c - a);
        float sq = dot(av, av);
        if (!std::isfinite(sq) || sq <= 0.0f) continue;
        float area = 0.5f * std::sqrt(sq);
        if (!std::isfinite(area) || area <= 0.0f) continue;
        float share = area / 3.0f;
        area_share[i0] += share; area_share[i1] += share; area_share[i2] += share;
        total_area += static_cast<double>(area);
        ++valid;
    }
    if (valid <= 0 || total_area <= 0.0 || !std::isfinite(total_area)) {
        std::fill(area_share.begin(), area_share.end(), 0.0f);
    } else {
        float mean = static_cast<float>(total_area / static_cast<double>(std::max(n, 1)));
        float mx = std::max(mean * 16.0f, 1.0e-12f);
        for (float& s : area_share) s = (!std::isfinite(s) || s <= 0.0f) ? 0.0f : std::min(s, mx);
    }
    rest_scale.assign(n, 0.0f);
    std::vector<double> sums(n, 0.0);
    std::vector<int> cnt(n, 0);
    for (const Edge& e : springs) {
        if (e.i >= n || e.j >= n || !std::isfinite(e.rest) || e.rest <= 0.0f) continue;
        sums[e.i] += e.rest; sums[e.j] += e.rest; cnt[e.i]++; cnt[e.j]++;
    }
    double total = 0.0;
    int vcount = 0;
    for (int i = 0; i < n; ++i) {
        float s = 0.0f;
        if (cnt[i] > 0) s = static_cast<float>(sums[i] / static_cast<double>(cnt[i]));
        else if (std::isfinite(area_share[i]) && area_share[i] > 0.0f) s = std::sqrt(area_share[i]);
        if (std::isfinite(s) && s > 0.0f) { rest_scale[i] = s; total += s; ++vcount; }
    }
    float mean_scale = vcount > 0 ? static_cast<float>(total / static_cast<double>(vcount)) : 1.0f;
    float floor_s = std::max(mean_scale * 0.02f, 1.0e-7f), cap_s = std::max(mean_scale * 16.0f, floor_s);
    for (float& s : rest_scale) {
        if (!std::isfinite(s) || s <= 0.0f) s = mean_scale;
        s = std::max(floor_s, std::min(s, cap_s));
    }
    const char* raw = std::getenv("SSBL_ABI41_PRESSURE_VELOCITY_OBJECT_SCALE");
    float obj = 0.08f;
    if (raw && raw[0]) { char* end = nullptr; float v = std::strtof(raw, &end); if (end != raw && std::isfinite(v)) obj = v; }
    obj = std::max(0.005f, std::min(obj, 1.0f));
    float ms = 0.0f; int mc = 0;
    for (float s : rest_scale) if (std::isfinite(s) && s > 0.0f) { ms += s; ++mc; }
    ms = mc > 0 ? ms / static_cast<float>(mc) : 1.0f;
    pressure_velocity_length_scale = std::max(ms, 1.0e-6f);
    if (n > 0) {
        VF lo{}, hi{};
        bool any = false;
        for (int i = 0; i < n; ++i) {
            VF p = P(i);
            if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.z)) continue;
            if (!any) { lo = hi = p; any = true; }
            else {
                lo.x = std::min(lo.x, p.x); lo.y = std::min(lo.y, p.y); lo.z = std::min(lo.z, p.z);
                hi.x = std::max(hi.x, p.x); hi.y = std::max(hi.y, p.y); hi.z = std::max(hi.z, p.z);
            }
        }
        if (any) {
            VF span = hi - lo;
            float diagl = std::sqrt(std::max(dot(span, span), 0.0f));
            if (std::isfinite(diagl) && diagl > 0.0f) pressure_velocity_length_scale = std::max(diagl * obj, ms * 4.0f);
        }
    }
}

// ------------------------------------------------------------ static SDF
struct QKey { long long x, y, z; bool operator==(const QKey& o) const { return x == o.x && y == o.y && z == o.z; } };
struct QKeyHash {
    size_t operator()(const QKey& k) const {
        uint64_t h = static_cast<uint64_t>(k.x) * 73856093ull ^ static_cast<uint64_t>(k.y) * 19349663ull ^ static_cast<uint64_t>(k.z) * 83492791ull;
        h ^= h >> 33; h *= 0xff51afd7ed558ccdull; h ^= h >> 33;
        return static_cast<size_t>(h);
    }
};
// A41:6026-6063: a triangle soup is closed when, after welding vertices on a
// 1e-6 grid, every undirected edge is used by exactly two triangles.
bool soup_is_closed(const float* t, int count) {
    if (!t || count <= 0) return false;
    std::unordered_map<QKey, int, QKeyHash> ids;
    std::unordered_map<uint64_t, int> edges;
    auto vid = [&](const float* v) {
        QKey k{std::llround(double(v[0]) * 1e6), std::llround(double(v[1]) * 1e6), std::llround(double(v[2]) * 1e6)};
        auto f = ids.find(k);
        if (f != ids.end()) return f->second;
        int id = static_cast<int>(ids.size());
        ids.emplace(k, id);
        return id;
    };
    auto ek = [](int a, int b) { return (uint64_t(unsigned(std::min(a, b))) << 32) | uint64_t(unsigned(std::max(a, b))); };
    for (int i = 0; i < count; ++i) {
        const float* a = t + 9 * i;
        for (int k = 0; k < 9; ++k) if (!std::isfinite(a[k])) return false;
        int ia = vid(a), ib = vid(a + 3), ic = vid(a + 6);
        if (ia == ib || ib == ic || ic == ia) continue;
        edges[ek(ia, ib)]++; edges[ek(ib, ic)]++; edges[ek(ic, ia)]++;
    }
    if (edges.empty()) return false;
    for (auto& kv : edges) if (kv.second != 2) return false;
    return true;
}

// A41:1725-1747 ray along +x against a triangle (Moller-Trumbore).
template <class R> bool ray_hits(V3<R> p, V3<R> a, V3<R> b, V3<R> c) {
    const V3<R> dir = mk<R>(1, 0, 0);
    V3<R> e1 = b - a, e2 = c - a;
    V3<R> h = cross(dir, e2);
    R det = dot(e1, h);
    if (std::fabs(det) <= R(1.0e-8)) return false;
    R inv = R(1) / det;
    V3<R> s = p - a;
    R u = inv * dot(s, h);
    if (u < R(-1.0e-5) || u > R(1) + R(1.0e-5)) return false;
    V3<R> q = cross(s, e1);
    R v = inv * dot(dir, q);
    if (v < R(-1.0e-5) || u + v > R(1) + R(1.0e-5)) return false;
    return inv * dot(e2, q) > R(1.0e-6);
}

// A41:6123-6218 grid set-up, A41:1787-1829 brute-force voxel distances.
// A static SDF given as a grid (a counterpart of build_sdf's output for
// callers that hold one; the static triangles stay for the normal fallback).
template <class R> bool Solver<R>::set_sdf_grid(const int* dims, const float* ov, const float* values, int count) {
    if (!dims || !ov || !values) return fail("static SDF grid, origin and values are required");
    const long long voxels = 1LL * dims[0] * dims[1] * dims[2];
    if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2 || voxels > std::numeric_limits<int>::max() || count != voxels)
        return fail("static SDF grid needs at least 2 voxels per axis and dx*dy*dz values");
    if (!std::isfinite(ov[0]) || !std::isfinite(ov[1]) || !std::isfinite(ov[2]) || !std::isfinite(ov[3]) || ov[3] <= 0.0f)
        return fail("static SDF origin and voxel size must be finite, the voxel size positive");
    sdf.assign(values, values + voxels);
    for (int k = 0; k < 3; ++k) sdf_dim[k] = dims[k];
    sdf_origin = mk<R>(ov[0], ov[1], ov[2]);
    sdf_voxel = R(ov[3]);
    sdf_signed = 1;
    sdf_ready = 1;
    sdf_rebuilds++;
    populate_sdf_diag();
    return true;
}

template <class R> bool Solver<R>::build_sdf(const float* triangles, int count) {
    sdf_ready = 0;
    sdf_dim[0] = sdf_dim[1] = sdf_dim[2] = 0;
    sdf_signed = 0;
    static_tris.clear();
    cfg.static_triangle_count = count;
    if (count <= 0) { populate_sdf_diag(); return true; }
    auto t0 = std::chrono::steady_clock::now();
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < 3 * count; ++i) {
        const float* v = triangles + 3 * i;
        if (!std::isfinite(v[0]) || !std::isfinite(v[1]) || !std::isfinite(v[2]))
            return fail("static triangle data contains non-finite coordinates");
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], v[k]); hi[k] = std::max(hi[k], v[k]); }
    }
    float voxel = (std::isfinite(cfg.static_sdf_voxel_size) && cfg.static_sdf_voxel_size > 0.0f)
                      ? cfg.static_sdf_voxel_size : std::max(std::max(cfg.collision_margin, 0.0f) * 0.5f, 0.002f);
    int band = std::max(cfg.static_sdf_band_voxels > 0 ? cfg.static_sdf_band_voxels : 4, 1);
    int max_res = std::max(cfg.static_sdf_max_resolution > 0 ? cfg.static_sdf_max_resolution : 160, 16);
    float margin = std::max(cfg.collision_margin, 0.0f);
    float expand = margin + static_cast<float>(band) * voxel + voxel * 2.0f;
    float ext[3];
    for (int k = 0; k < 3; ++k) { lo[k] -= expand; hi[k] += expand; ext[k] = std::max(hi[k] - lo[k], voxel); }
    float largest = std::max(ext[0], std::max(ext[1], ext[2]));
    if (largest / voxel + 1.0f > static_cast<float>(max_res)) voxel = largest / static_cast<float>(std::max(max_res - 1, 1));
    voxel = std::max(voxel, 1.0e-5f);
    int dim[3];
    for (int k = 0; k < 3; ++k) dim[k] = std::min(std::max(static_cast<int>(std::ceil(ext[k] / voxel)) + 1, 2), max_res);
    long long voxels = 1LL * dim[0] * dim[1] * dim[2];
    if (voxels <= 0 || voxels > std::numeric_limits<int>::max()) return fail("static SDF grid is too large");
    for (int k = 0; k < 3; ++k) sdf_dim[k] = dim[k];
    sdf_origin = mk<R>(lo[0], lo[1], lo[2]);
    sdf_voxel = R(voxel);
    sdf_signed = soup_is_closed(triangles, count) ? 1 : 0;
    if (!sdf_signed) sdf_unsigned++;
    static_tris.resize(3 * static_cast<size_t>(count));
    for (int i = 0; i < 3 * count; ++i) static_tris[i] = mk<R>(triangles[3 * i], triangles[3 * i + 1], triangles[3 * i + 2]);
    sdf.assign(static_cast<size_t>(voxels), R(0));
    for (long long vx = 0; vx < voxels; ++vx) {
        int x = static_cast<int>(vx % dim[0]), y = static_cast<int>((vx / dim[0]) % dim[1]), z = static_cast<int>(vx / (1LL * dim[0] * dim[1]));
        V3<R> p = sdf_origin + mk(R(x) * sdf_voxel, R(y) * sdf_voxel, R(z) * sdf_voxel);
        V3<R> rp = p + mk(R(0), sdf_voxel * R(1.37e-4), sdf_voxel * R(2.11e-4));
        R best = R(3.402823466e+38);
        int parity = 0;
        for (int t = 0; t < count; ++t) {
            V3<R> a = static_tris[3 * t], b = static_tris[3 * t + 1], c = static_tris[3 * t + 2];
            V3<R> d = p - closest_point_on_triangle(p, a, b, c);
            best = std::min(best, dot(d, d));
            if (sdf_signed && ray_hits(rp, a, b, c)) parity ^= 1;
        }
        R dist = std::sqrt(std::max(best, R(0)));
        if (sdf_signed && parity) dist = -dist;
        sdf[static_cast<size_t>(vx)] = dist;
    }
    sdf_ready = 1;
    sdf_rebuilds++;
    sdf_build_ms = static_cast<float>(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    populate_sdf_diag();
    return true;
}

template <class R> bool Solver<R>::sdf_sample(V3<R> p, R* out) const {  // A41:1645-1691
    if (!sdf_ready || sdf_dim[0] < 2 || sdf_dim[1] < 2 || sdf_dim[2] < 2 || sdf_voxel <= R(0)) return false;
    R inv = R(1) / sdf_voxel;
    R gx = (p.x - sdf_origin.x) * inv, gy = (p.y - sdf_origin.y) * inv, gz = (p.z - sdf_origin.z) * inv;
    if (gx < R(0) || gy < R(0) || gz < R(0) || gx > R(sdf_dim[0] - 1) || gy > R(sdf_dim[1] - 1) || gz > R(sdf_dim[2] - 1)) return false;
    auto cl = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
    int x0 = cl(static_cast<int>(std::floor(gx)), 0, sdf_dim[0] - 1), y0 = cl(static_cast<int>(std::floor(gy)), 0, sdf_dim[1] - 1),
        z0 = cl(static_cast<int>(std::floor(gz)), 0, sdf_dim[2] - 1);
    int x1 = cl(x0 + 1, 0, sdf_dim[0] - 1), y1 = cl(y0 + 1, 0, sdf_dim[1] - 1), z1 = cl(z0 + 1, 0, sdf_dim[2] - 1);
    R tx = gx - R(x0), ty = gy - R(y0), tz = gz - R(z0);
    auto at = [&](int x, int y, int z) { return sdf[(static_cast<size_t>(z) * sdf_dim[1] + y) * sdf_dim[0] + x]; };
    auto lerp = [](R a, R b, R t) { return a + (b - a) * t; };
    R c00 = lerp(at(x0, y0, z0), at(x1, y0, z0), tx), c10 = lerp(at(x0, y1, z0), at(x1, y1, z0), tx);
    R c01 = lerp(at(x0, y0, z1), at(x1, y0, z1), tx), c11 = lerp(at(x0, y1, z1), at(x1, y1, z1), tx);
    R c0 = lerp(c00, c10, ty), c1 = lerp(c01, c11, ty);
    *out = lerp(c0, c1, tz);
    return std::isfinite(*out);
}

template <class R> R Solver<R>::sdf_sample_clamped(V3<R> p) const {  // A41:1693-1723
    R inv = R(1) / std::max(sdf_voxel, R(kEpsF));
    R gx = std::min(std::max((p.x - sdf_origin.x) * inv, R(0)), R(sdf_dim[0] - 1));
    R gy = std::min(std::max((p.y - sdf_origin.y) * inv, R(0)), R(sdf_dim[1] - 1));
    R gz = std::min(std::max((p.z - sdf_origin.z) * inv, R(0)), R(sdf_dim[2] - 1));
    auto cl = [](int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); };
    int x0 = cl(static_cast<int>(std::floor(gx)), 0, sdf_dim[0] - 1), y0 = cl(static_cast<int>(std::floor(gy)), 0, sdf_dim[1] - 1),
        z0 = cl(static_cast<int>(std::floor(gz)), 0, sdf_dim[2] - 1);
    int x1 = cl(x0 + 1, 0, sdf_dim[0] - 1), y1 = cl(y0 + 1, 0, sdf_dim[1] - 1), z1 = cl(z0 + 1, 0, sdf_dim[2] - 1);
    R tx = gx - R(x0), ty = gy - R(y0), tz = gz - R(z0);
    auto at = [&](int x, int y, int z) { return sdf[(static_cast<size_t>(z) * sdf_dim[1] + y) * sdf_dim[0] + x]; };
    auto lerp = [](R a, R b, R t) { return a + (b - a) * t; };
    return lerp(lerp(lerp(at(x0, y0, z0), at(x1, y0, z0), tx), lerp(at(x0, y1, z0), at(x1, y1, z0), tx), ty),
                lerp(lerp(at(x0, y0, z1), at(x1, y0, z1), tx), lerp(at(x0, y1, z1), at(x1, y1, z1), tx), ty), tz);
}

template <class R> V3<R> Solver<R>::sdf_normal(V3<R> p) const {  // A41:1771-1785, fallback A41:1749-1769
    R h = std::max(sdf_voxel, R(1.0e-5));
    R dx = sdf_sample_clamped(p + mk(h, R(0), R(0))) - sdf_sample_clamped(p + mk(-h, R(0), R(0)));
    R dy = sdf_sample_clamped(p + mk(R(0), h, R(0))) - sdf_sample_clamped(p + mk(R(0), -h, R(0)));
    R dz = sdf_sample_clamped(p + mk(R(0), R(0), h)) - sdf_sample_clamped(p + mk(R(0), R(0), -h));
    V3<R> n = mk(dx, dy, dz);
    R lsq = dot(n, n);
    if (std::isfinite(lsq) && lsq > R(1.0e-12)) return mul(n, rsqrt(lsq));
    V3<R> best_n = mk<R>(0, 0, 1);
    R best = R(3.402823466e+38);
    const int count = static_cast<int>(static_tris.size() / 3);
    for (int t = 0; t < count; ++t) {
        V3<R> a = static_tris[3 * t], b = static_tris[3 * t + 1], c = static_tris[3 * t + 2];
        V3<R> d = p - closest_point_on_triangle(p, a, b, c);
        R dsq = dot(d, d);
        if (dsq < best) {
            best = dsq;
            best_n = dsq > R(kEpsF) * R(kEpsF) ? normalize_or(d, best_n) : normalize_or(cross(b - a, c - a), best_n);
        }
    }
    return normalize_or(best_n, mk<R>(0, 0, 1));
}

// ------------------------------------------------------------- substeps
// A41:910-949
template <class R> bool Solver<R>::contact_response(V3<R> n, R proxy, V3<R>* v) {
    if (!finite(n) || !finite(*v)) return false;
    R vn = dot(*v, n);
    if (vn < R(0)) *v = *v - mul(n, vn);
    bool corrected = false;
    V3<R> tangent = *v - mul(n, dot(*v, n));
    R tl = length(tangent);
    R damping = clamp01(R(cfg.contact_tangent_damping));
    if (damping > R(0) && tl > R(1.0e-6)) { *v = *v - mul(tangent, damping); corrected = true; }
    tangent = *v - mul(n, dot(*v, n));
    tl = length(tangent);
    R friction = std::max(R(cfg.contact_friction), R(0));
    if (friction > R(0) && proxy > R(0) && tl > R(1.0e-6)) {
        R drop = std::min(tl, friction * proxy);
        if (drop > R(1.0e-7)) { *v = *v - mul(tangent, drop / std::max(tl, R(kEpsF))); corrected = true; }
    }
    if (corrected) count(kCountFriction);
    return corrected;
}

// A41:1213-1261, with rsqrtf restated as 1/sqrtf; P is the snapshot read.
template <class R> V3<R> Solver<R>::normal_at(int v, const std::vector<V3<R>>& P, int* has) const {
    *has = 0;
    int s = tri_offsets[v], e = tri_offsets[v + 1];
    if (e <= s) return mk<R>(0, 0, 0);
    V3<R> acc = mk<R>(0, 0, 0);
    const int n = cfg.vertex_count;
    for (int c = s; c < e; ++c) {
        int t = tri_incident[c];
        if (t < 0 || t >= cfg.triangle_count) continue;
        int i0 = tris[3 * t], i1 = tris[3 * t + 1], i2 = tris[3 * t + 2];
        if (i0 >= n || i1 >= n || i2 >= n) continue;
        V3<R> nn = cross(P[i1] - P[i0], P[i2] - P[i0]);
        R lsq = dot(nn, nn);
        if (!std::isfinite(lsq) || lsq <= R(1.0e-10)) continue;
        acc = acc + mul(nn, rsqrt(lsq));
    }
    R asq = dot(acc, acc);
    if (!std::isfinite(asq) || asq <= R(1.0e-10)) return mk<R>(0, 0, 0);
    *has = 1;
    return mul(acc, rsqrt(asq));
}

// A41:1869-1925; neighbours' positions for the surface normal are read from the
// positions before the pass.
template <class R> void Solver<R>::integrate(R dt) {
    const int n = V();
    const bool need_normal = (cfg.use_volume_pressure && cfg.pressure_strength > 0.0f) || !fields.empty();
    std::vector<V3<R>> P = pos;
    for (int i = 0; i < n; ++i) {
        prev[i] = P[i];
        if (inv_mass[i] <= 0.0f || pinned[i]) { vel[i] = mk<R>(0, 0, 0); pos[i] = rest[i]; continue; }
        int has = 0;
        V3<R> nrm = need_normal ? normal_at(i, P, &has) : mk<R>(0, 0, 0);
        V3<R> acc = mk<R>(cfg.gravity[0], cfg.gravity[1], cfg.gravity[2]);
        const R im = R(inv_mass[i]);
        const bool pressure = cfg.use_volume_pressure && cfg.pressure_strength > 0.0f && has;
        if (pressure) {
            R pf = std::max(R(cfg.pressure_strength), R(0));
            R share = R(area_share[i]);
            if (std::isfinite(share) && share > R(0)) acc = acc + mul(nrm, pf * R(pressure_accel_scale) * share * im);
        }
        if (!fields.empty()) {
            V3<R> fa = mk<R>(0, 0, 0);
            int cnt = std::min(static_cast<int>(fields.size()), kMaxForceFields);
            for (int k = 0; k < cnt; ++k) fa = fa + eval_field<R>(fields[k], P[i], vel[i], nrm, has);
            acc = acc + limit_accel(fa);
        }
        V3<R> v = vel[i] + mul(acc, dt);
        v = mul(v, R(cfg.damping));
        if (pressure) {
            R local_rest = R(rest_scale[i]);
            if (std::isfinite(local_rest) && local_rest > R(kEpsF) && std::isfinite(dt) && dt > R(kEpsF)) {
                R pa = std::max(R(cfg.pressure_strength), R(0)) * std::max(R(pressure_accel_scale), R(0)) *
                       std::max(R(area_share[i]), R(0)) * std::max(im, R(0));
                R max_pd = std::max(local_rest * std::max(R(pressure_impulse_rest_scale), R(0)), R(1.0e-7));
                R pd = pa * dt * dt;
                if (std::isfinite(pd) && pd > max_pd) v = v - mul(nrm, (pd - max_pd) / dt);
                R vr = std::max(local_rest, std::max(R(pressure_velocity_length_scale), R(0)));
                R max_out = std::max(vr * std::max(R(pressure_velocity_rest_scale), R(0)) / dt, R(1.0e-6));
                R out = dot(v, nrm);
                if (std::isfinite(out) && out > max_out) v = v - mul(nrm, out - max_out);
            }
        }
        vel[i] = v;
        pos[i] = P[i] + mul(v, dt);
    }
}

// A41:3466-3509; pins apply in ascending pin order.
template <class R> void Solver<R>::pin_project(R exponent, R dt) {
    for (const Pin& p : pins) {
        int i = p.index;
        if (i < 0 || i >= V()) continue;
        float w = p.weight;
        if (!std::isfinite(w) || w < 0.05f) continue;
        V3<R> target = mk<R>(p.target[0], p.target[1], p.target[2]);
        if (w >= 0.75f || inv_mass[i] <= 0.0f) { pos[i] = target; prev[i] = target; vel[i] = mk<R>(0, 0, 0); continue; }
        // soft relaxation; computed in float exactly as the device would see it
        float total = std::min(std::max(w / 0.75f, 0.0f), 1.0f);
        float relax;
        if (total >= 1.0f) relax = 1.0f;
        else {
            float remaining = std::max(1.0f - total, 0.0f);
            float r = 1.0f - std::pow(remaining, std::max(static_cast<float>(exponent), 1.0e-4f));
            relax = std::isfinite(r) ? std::min(std::max(r, 0.0f), 1.0f) : total;
        }
        if (relax <= 0.0f) continue;
        R rl = R(relax);
        V3<R> corrected = pos[i] + mul(target - pos[i], rl);
        V3<R> damped = mul(vel[i], std::max(R(1) - rl, R(0)));
        pos[i] = corrected;
        prev[i] = corrected - mul(damped, std::max(dt, R(kEpsF)));
        vel[i] = damped;
    }
}

// A41:1927-1980 as a Jacobi sweep.
template <class R> void Solver<R>::stretch_jacobi(R dt) {
    const int n = V();
    const std::vector<V3<R>> P = pos;
    const R alpha = std::max(R(cfg.stretch_compliance), R(0)) / std::max(dt * dt, R(kEpsF));
    for (const Edge& s : springs) {
        int i = s.i, j = s.j;
        if (i >= n || j >= n) continue;
        R wi = weight(i), wj = weight(j);
        R w = wi + wj;
        if (w <= R(0)) continue;
        V3<R> delta = P[j] - P[i];
        R len = std::sqrt(std::max(dot(delta, delta), R(kEpsF)));
        R c = len - R(s.rest);
        R dl = -c / (w + alpha);
        V3<R> corr = mul(delta, R(0.18) * dl / std::max(len, R(kEpsF)));
        R cl = std::sqrt(std::max(dot(corr, corr), R(0)));
        R mc = std::max(R(0.0025), std::min(std::max(R(s.rest), R(0)) * R(0.10), R(0.025)));
        if (cl > mc) { corr = mul(corr, mc / std::max(cl, R(kEpsF))); count(kCountHardFallbacks); }
        const bool over = c > R(0);
        if (wi > R(0)) {
            V3<R> di = mul(corr, -wi);
            pos[i] = pos[i] + di;
            if (over) prev[i] = prev[i] + mul(di, R(0.08));
        }
        if (wj > R(0)) {
            V3<R> dj = mul(corr, wj);
            pos[j] = pos[j] + dj;
            if (over) prev[j] = prev[j] + mul(dj, R(0.08));
        }
    }
}

// Coloured Gauss-Seidel form of the same projection: colours in order, each
// constraint reading the current positions (the mode requires a valid
// colouring, so a colour's constraints touch disjoint vertices).
template <class R> void Solver<R>::stretch_colored(R dt) {
    const int n = V();
    const R alpha = std::max(R(cfg.stretch_compliance), R(0)) / std::max(dt * dt, R(kEpsF));
    for (int c = 0; c < cfg.edge_color_count; ++c) {
        int lo = std::max(edge_colors[c], 0);
        int hi = std::min(edge_colors[c + 1], cfg.edge_count);
        for (int e = lo; e < hi; ++e) {
            const Edge& s = springs[e];
            int i = s.i, j = s.j;
            if (i >= n || j >= n) continue;
            R wi = weight(i), wj = weight(j), w = wi + wj;
            if (w <= R(0)) continue;
            V3<R> delta = pos[j] - pos[i];
            R len = std::sqrt(std::max(dot(delta, delta), R(kEpsF)));
            R cc = len - R(s.rest);
            R dl = -cc / (w + alpha);
            V3<R> corr = mul(delta, R(0.18) * dl / std::max(len, R(kEpsF)));
            R cl = std::sqrt(std::max(dot(corr, corr), R(0)));
            R mc = std::max(R(0.0025), std::min(std::max(R(s.rest), R(0)) * R(0.10), R(0.025)));
            if (cl > mc) { corr = mul(corr, mc / std::max(cl, R(kEpsF))); count(kCountHardFallbacks); }
            const bool over = cc > R(0);
            if (wi > R(0)) { V3<R> di = mul(corr, -wi); pos[i] = pos[i] + di; if (over) prev[i] = prev[i] + mul(di, R(0.08)); }
            if (wj > R(0)) { V3<R> dj = mul(corr, wj); pos[j] = pos[j] + dj; if (over) prev[j] = prev[j] + mul(dj, R(0.08)); }
        }
    }
}

// A41:2710-2803 as a Jacobi sweep.
template <class R> void Solver<R>::lra_jacobi(R dt) {
    const int n = V();
    const std::vector<V3<R>> P = pos;
    for (int k = 0; k < cfg.lra_count; ++k) {
        const Edge& L = lras[k];
        int i = L.i, j = L.j;
        if (i < 0 || j < 0 || i >= n || j >= n || i == j) { count(kCountTackGuards); continue; }
        R wi = weight(i), wj = weight(j), w = wi + wj;
        if (!std::isfinite(w) || w <= R(0)) { count(kCountTackGuards); continue; }
        V3<R> a = P[i], d = P[j];
        if (!finite(a) || !finite(d)) { count(kCountTackGuards); continue; }
        V3<R> dx = d - a;
        R lsq = std::fma(dx.x, dx.x, std::fma(dx.y, dx.y, dx.z * dx.z));
        R restl = R(L.rest);
        if (!std::isfinite(lsq) || lsq <= R(1.0e-10) || !std::isfinite(restl) || restl <= R(kEpsF)) { count(kCountTackGuards); continue; }
        R inv_len = rsqrt(lsq);
        R len = lsq * inv_len;
        R stretch = len - restl;
        if (!std::isfinite(stretch) || stretch <= R(0)) continue;
        dx = mul(dx, inv_len);
        R bc = std::max(R(cfg.lra_compliance), R(1.0e-12));
        R stiff = (R(1) / bc) * R(500000);
        R ealpha = (bc / R(500000)) / std::max(dt * dt, R(kEpsF));
        R sw = stiff / std::max(stiff + R(500000), R(1));
        R dl = -stretch / std::max(w + ealpha, R(kEpsF));
        V3<R> corr = mul(dx, R(0.35) * sw * dl);
        if (!finite(corr)) { count(kCountTackGuards); continue; }
        R cl = length(corr);
        R mc = std::max(R(5.0e-4), std::min(std::max(restl, R(cfg.cloth_thickness)) * R(0.20), R(0.035)));
        if (!std::isfinite(cl) || cl <= R(0)) continue;
        if (cl > mc) { corr = mul(corr, mc / std::max(cl, R(kEpsF))); count(kCountTackGuards); }
        const bool th = cfg.cloth_thickness > 0.0f;
        R cap = th ? R(0.0025) : R(0.012);
        R rl = std::max(restl, R(0)) * (th ? R(0.02) : R(0.08));
        R tl = th ? R(cfg.cloth_thickness) * R(0.12) : cap;
        R al = std::max(R(5.0e-4), std::min(std::min(rl, tl), cap));
        if (wi > R(0)) {
            V3<R> di = limit_delta(mul(corr, -wi / std::max(w, R(kEpsF))), al);
            pos[i] = pos[i] + di; prev[i] = prev[i] + mul(di, R(0.32));
        }
        if (wj > R(0)) {
            V3<R> dj = limit_delta(mul(corr, wj / std::max(w, R(kEpsF))), al);
            pos[j] = pos[j] + dj; prev[j] = prev[j] + mul(dj, R(0.32));
        }
        count(kCountLraTacks);
    }
}

// A41:2886-2970: one bend constraint; returns false when it does nothing.
template <class R> bool Solver<R>::bend_one(const Edge& b, R dt, const std::vector<V3<R>>& P, V3<R>* di, V3<R>* dj) {
    const int n = V();
    int i = b.i, j = b.j;
    *di = *dj = mk<R>(0, 0, 0);
    if (i < 0 || j < 0 || i >= n || j >= n || i == j) { count(kCountBendingGuards); return false; }
    R wi = weight(i), wj = weight(j), w = wi + wj;
    if (!std::isfinite(w) || w <= R(0)) return false;
    R restl = R(b.rest);
    if (!std::isfinite(restl) || restl <= R(kEpsF)) { count(kCountBendingGuards); return false; }
    V3<R> pi = P[i], pj = P[j];
    if (!finite(pi) || !finite(pj)) { count(kCountBendingGuards); return false; }
    V3<R> delta = pj - pi;
    R lsq = std::fma(delta.x, delta.x, std::fma(delta.y, delta.y, delta.z * delta.z));
    if (!std::isfinite(lsq) || lsq <= R(1.0e-10)) { count(kCountBendingGuards); return false; }
    R inv_len = rsqrt(lsq);
    R len = lsq * inv_len;
    R c = len - restl;
    if (!std::isfinite(c) || std::fabs(c) <= R(1.0e-6)) return false;
    R alpha = std::max(R(cfg.bend_compliance), R(0)) / std::max(dt * dt, R(kEpsF));
    R dl = -c / std::max(w + alpha, R(kEpsF));
    V3<R> corr = mul(delta, R(0.10) * dl * inv_len);
    if (!finite(corr)) { count(kCountBendingGuards); return false; }
    R cl = length(corr);
    R mc = std::max(R(2.5e-4), std::min(std::max(restl, R(cfg.cloth_thickness)) * R(0.08), R(0.020)));
    if (!std::isfinite(cl) || cl <= R(0)) return false;
    if (cl > mc) { corr = mul(corr, mc / std::max(cl, R(kEpsF))); count(kCountBendingGuards); }
    if (wi > R(0)) *di = mul(corr, -wi);
    if (wj > R(0)) *dj = mul(corr, wj);
    count(kCountBendingWings);
    return true;
}

// A41:9205-9217: coloured pass (A41:2886-2970) when the colouring is valid,
// otherwise the uncoloured kernel (A41:2805-2884) as a Jacobi sweep.
template <class R> void Solver<R>::bend_pass(R dt) {
    if (bend_coloring_valid) {
        for (int c = 0; c < cfg.bend_color_count; ++c) {
            int lo = std::max(bend_colors[c], 0), hi = std::min(bend_colors[c + 1], cfg.bend_count);
            for (int k = lo; k < hi; ++k) {
                V3<R> di, dj;
                if (!bend_one(bends[k], dt, pos, &di, &dj)) continue;
                const int i = bends[k].i, j = bends[k].j;
                if (weight(i) > R(0)) { pos[i] = pos[i] + di; prev[i] = prev[i] + mul(di, R(0.06)); }
                if (weight(j) > R(0)) { pos[j] = pos[j] + dj; prev[j] = prev[j] + mul(dj, R(0.06)); }
            }
        }
        return;
    }
    const std::vector<V3<R>> P0 = pos;  // every bend reads the positions from before the pass
    for (int k = 0; k < cfg.bend_count; ++k) {
        V3<R> di, dj;
        if (!bend_one(bends[k], dt, P0, &di, &dj)) continue;
        const int i = bends[k].i, j = bends[k].j;
        if (weight(i) > R(0)) { pos[i] = pos[i] + di; prev[i] = prev[i] + mul(di, R(0.06)); }
        if (weight(j) > R(0)) { pos[j] = pos[j] + dj; prev[j] = prev[j] + mul(dj, R(0.06)); }
    }
}

// A41:3511-3590
template <class R> void Solver<R>::analytic(R dt) {
    const R margin = std::max(R(cfg.collision_margin), R(0));
    const R sdt = std::max(dt, R(kEpsF));
    for (int i = 0; i < V(); ++i) {
        if (inv_mass[i] <= 0.0f) continue;
        V3<R> p = pos[i], v = vel[i];
        bool hit = false;
        if (colliders.use_ground && p.z < R(colliders.ground_height) + margin) {
            R tz = R(colliders.ground_height) + margin;
            R pen = tz - p.z;
            V3<R> nn = mk<R>(0, 0, 1);
            R inward = std::max(-dot(v, nn), R(0));
            p.z = tz;
            contact_response(nn, inward + pen / sdt, &v);
            hit = true;
        }
        if (colliders.use_sphere && colliders.sphere_radius > 0.0f) {
            V3<R> c = arr3<R>(colliders.sphere_center);
            V3<R> d = p - c;
            R dist = std::sqrt(std::max(dot(d, d), R(kEpsF)));
            R target = R(colliders.sphere_radius) + margin;
            if (dist < target) {
                V3<R> nn = dist > R(kEpsF) ? mul(d, R(1) / dist) : mk<R>(0, 0, 1);
                R pen = target - dist;
                R inward = std::max(-dot(v, nn), R(0));
                p = c + mul(nn, target);
                contact_response(nn, inward + pen / sdt, &v);
                hit = true;
            }
        }
        if (colliders.use_wall) {
            V3<R> o = arr3<R>(colliders.wall_origin), nn = arr3<R>(colliders.wall_normal);
            R nl = std::sqrt(std::max(dot(nn, nn), R(kEpsF)));
            nn = mul(nn, R(1) / nl);
            R gap = dot(p - o, nn);
            if (gap < margin) {
                R pen = margin - gap;
                R inward = std::max(-dot(v, nn), R(0));
                p = p + mul(nn, margin - gap);
                contact_response(nn, inward + pen / sdt, &v);
                hit = true;
            }
        }
        pos[i] = p;
        if (hit) { vel[i] = v; prev[i] = p - mul(v, sdt); }
    }
}

// A41:1831-1867
template <class R> void Solver<R>::static_sdf_contacts(R dt) {
    const R margin = std::max(R(cfg.collision_margin), R(0));
    if (margin <= R(0)) return;
    for (int i = 0; i < V(); ++i) {
        if (inv_mass[i] <= 0.0f || pinned[i]) continue;
        V3<R> p = pos[i];
        R d = 0;
        if (!sdf_sample(p, &d) || d >= margin) continue;
        V3<R> nn = sdf_normal(p);
        if (!finite(nn)) continue;
        R pen = margin - d;
        if (!std::isfinite(pen) || pen <= R(0)) continue;
        V3<R> corrected = p + mul(nn, pen);
        V3<R> v = vel[i];
        R inward = std::max(-dot(v, nn), R(0));
        contact_response(nn, inward + pen / std::max(dt, R(kEpsF)), &v);
        pos[i] = corrected;
        vel[i] = v;
        prev[i] = corrected - mul(v, std::max(dt, R(kEpsF)));
        count(kCountStaticSdf);
    }
}

// A41:5846-5884
template <class R> void Solver<R>::velocity_update(R dt) {
    const R sdt = std::max(dt, R(kEpsF));
    const R max_fd = std::max(R(2.0e-4), std::max(R(cfg.cloth_thickness), R(0)) * R(0.35));
    const R max_fs = max_fd / sdt;
    for (int i = 0; i < V(); ++i) {
        if (inv_mass[i] <= 0.0f || pinned[i]) { vel[i] = mk<R>(0, 0, 0); continue; }
        V3<R> meas = mul(pos[i] - prev[i], R(1) / sdt);
        V3<R> base = vel[i];
        if (!finite(meas) || !finite(base)) { vel[i] = mk<R>(0, 0, 0); continue; }
        V3<R> pv = meas - base;
        R pdb = dot(pv, base);
        R fs = (std::isfinite(pdb) && pdb > R(0)) ? R(0.9) : R(1);
        V3<R> fb = mul(pv, fs);
        R speed = length(fb);
        if (fs < R(1) && std::isfinite(speed) && speed > max_fs && speed > R(kEpsF)) fb = mul(fb, max_fs / speed);
        if (!finite(fb)) fb = mk<R>(0, 0, 0);
        vel[i] = base + fb;
    }
}

// ------------------------------------------------------- self collision
// Buckets store (primitive) lists in insertion order; queries see the first 32.
struct HashTable {
    int buckets = 0;
    std::vector<std::vector<int>> lists;
    void reset(int b) { buckets = b; lists.assign(b, {}); }
};

// Fast-mode filters, A41:708-849: per-thread seen-id list and contact clusters.
struct SelfFilterState {
    int seen[kFilterLocalSlots];
    int nseen = 0;
    unsigned keys[kClusterMaxLocal];
    unsigned char owned[kClusterMaxLocal];
    int nclusters = 0;
};
inline int filter_normal_bucket(float v) { return v > 0.45f ? 2 : (v < -0.45f ? 0 : 1); }
// A41:742-760
template <class R> inline unsigned filter_cluster_key(int type, V3<R> point, V3<R> normal, R cell) {
    const R cs = std::max(cell, R(1.0e-3));
    unsigned key = hash_cell(cell_coord(point.x, cs), cell_coord(point.y, cs), cell_coord(point.z, cs), 1 << 20);
    const unsigned nk = static_cast<unsigned>(filter_normal_bucket(static_cast<float>(normal.x))) |
                        (static_cast<unsigned>(filter_normal_bucket(static_cast<float>(normal.y))) << 2) |
                        (static_cast<unsigned>(filter_normal_bucket(static_cast<float>(normal.z))) << 4);
    key ^= nk * 2654435761u;
    key ^= static_cast<unsigned>(type + 1) * 2246822519u;
    return key;
}

// A41:1574-1587; the sum is kept in 2^-32 fixed point so that its value does
// not depend on the order of the contributions.
template <class R> void Solver<R>::self_accumulate(int v, V3<R> d) {
    if (v < 0 || v >= V() || inv_mass[v] <= 0.0f || pinned[v]) return;
    const double s = 4294967296.0;
    acc_q[3 * v] += static_cast<int64_t>(std::llrint(double(d.x) * s));
    acc_q[3 * v + 1] += static_cast<int64_t>(std::llrint(double(d.y) * s));
    acc_q[3 * v + 2] += static_cast<int64_t>(std::llrint(double(d.z) * s));
    acc_w[v] += 1;
}

// A41:5226-5439: a vertex inside several nearby triangles (three or more
// penetrating contacts within the 27 triangle-hash cells around it) is pushed
// out along their penetration-weighted mean normal.
template <class R> template <class H> void Solver<R>::island_aggregate(const H& TH, R cs) {
    const int n = V();
    const R r = self_radius();
    const R onset = std::max(r * R(1.8), r + R(kEpsF));
    const R rest_skip = std::max(onset * R(1.25), r * R(2.5));
    if (r <= R(0)) return;
    const int max_nb = std::max(cfg.max_self_collision_neighbors, 1);
    const int max_contacts = std::max(std::min(max_nb * 2, kIslandMaxContacts), kIslandMinContacts);
    const int scan_limit = max_contacts * 10;
    for (int v = 0; v < n; ++v) {
        if (inv_mass[v] <= 0.0f || pinned[v]) continue;
        const V3<R> p = pos[v], pp = prev[v];
        if (!finite(p) || !finite(pp)) { count(kCountIslandGuarded); continue; }
        const int cx = cell_coord(p.x, cs), cy = cell_coord(p.y, cs), cz = cell_coord(p.z, cs);
        int accepted[kIslandMaxContacts];
        int scanned = 0, contacts = 0;
        R pen_sum = 0, sample_sum = 0, max_pen = 0;
        V3<R> axis = mk<R>(0, 0, 0), best = mk<R>(0, 0, 1);
        for (int dz = -1; dz <= 1 && contacts < max_contacts && scanned < scan_limit; ++dz)
            for (int dy = -1; dy <= 1 && contacts < max_contacts && scanned < scan_limit; ++dy)
                for (int dx = -1; dx <= 1 && contacts < max_contacts && scanned < scan_limit; ++dx) {
                    const auto& L = TH.lists[hash_cell(cx + dx, cy + dy, cz + dz, TH.buckets)];
                    const int lim = std::min(static_cast<int>(L.size()), kBucketSlots);
                    for (int k = 0; k < lim && contacts < max_contacts && scanned < scan_limit; ++k) {
                        ++scanned;
                        const int t = L[k];
                        bool dup = false;
                        for (int i = 0; i < contacts; ++i) if (accepted[i] == t) { dup = true; break; }
                        if (dup) continue;
                        const int ia = tris[3 * t], ib = tris[3 * t + 1], ic = tris[3 * t + 2];
                        if (ia >= n || ib >= n || ic >= n || ia == ib || ib == ic || ia == ic) continue;
                        if (v == ia || v == ib || v == ic) continue;
                        {
                            const V3<R> rd = rest[v] - closest_point_on_triangle(rest[v], rest[ia], rest[ib], rest[ic]);
                            if (dot(rd, rd) < rest_skip * rest_skip) continue;
                        }
                        const V3<R> a = pos[ia], b = pos[ib], c = pos[ic];
                        const V3<R> ap = prev[ia], bp = prev[ib], cp = prev[ic];
                        if (!finite(a) || !finite(b) || !finite(c) || !finite(ap) || !finite(bp) || !finite(cp)) {
                            count(kCountIslandGuarded);
                            continue;
                        }
                        const V3<R> q = closest_point_on_triangle(p, a, b, c);
                        V3<R> d = p - q;
                        const R dsq = dot(d, d);
                        R dist = std::sqrt(std::max(dsq, R(kEpsF)));
                        if (dist >= r) continue;
                        V3<R> nn = mk<R>(0, 0, 1);
                        if (dsq > R(kEpsF)) nn = mul(d, R(1) / dist);
                        else {
                            nn = cross(b - a, c - a);
                            const R nsq = dot(nn, nn);
                            if (nsq > R(kEpsF)) nn = mul(nn, rsqrt(nsq));
                            else {
                                const V3<R> rd = rest[v] - closest_point_on_triangle(rest[v], rest[ia], rest[ib], rest[ic]);
                                const R rsq = dot(rd, rd);
                                if (rsq > R(kEpsF)) nn = mul(rd, rsqrt(rsq));
                            }
                            d = mul(nn, dist);
                            dist = R(0);
                        }
                        if (!finite(nn)) { count(kCountIslandGuarded); continue; }
                        const V3<R> pd = pp - closest_point_on_triangle(pp, ap, bp, cp);
                        if (dot(pd, nn) > r * R(0.25) && dot(d, nn) > -r * R(0.10)) continue;
                        R wa = R(1) / R(3), wb = wa, wc = wa;
                        triangle_barycentric(q, a, b, c, &wa, &wb, &wc);
                        const R pen = r - dist;
                        if (pen <= R(0)) continue;
                        accepted[contacts++] = t;
                        pen_sum += pen;
                        sample_sum += R(inv_mass[ia]) * wa + R(inv_mass[ib]) * wb + R(inv_mass[ic]) * wc;
                        axis = axis + mul(nn, pen);
                        if (pen > max_pen) { max_pen = pen; best = nn; }
                    }
                }
        if (contacts <= 0) continue;
        counts[kCountIslandCandidates] += static_cast<unsigned long long>(contacts);
        if (contacts < kIslandMinContacts) { count(kCountIslandGuarded); continue; }
        const R avg = pen_sum / R(contacts);
        if (avg < r * R(kIslandMinAvgDepthScale)) { count(kCountIslandGuarded); continue; }
        const R axis_len = length(axis);
        const V3<R> agg = axis_len > R(kEpsF) ? mul(axis, R(1) / axis_len) : best;
        const R source_w = R(inv_mass[v]);
        const R sample_w = sample_sum / R(contacts);
        const R share = source_w / std::max(source_w + sample_w, R(kEpsF));
        const R max_delta = std::max(r * R(kIslandMaxDeltaScale), R(1.0e-5));
        const R dl = std::min(avg * R(kIslandCorrectionScale) * share, max_delta);
        if (!finite(agg) || !std::isfinite(dl) || dl <= R(kEpsF)) { count(kCountIslandGuarded); continue; }
        self_accumulate(v, mul(agg, dl));
        count(kCountIslandClusters);
        counts[kCountIslandVertexRefs] += static_cast<unsigned long long>(contacts);
        count(kCountIslandApplied);
        // A41:852-863
        const float fd = static_cast<float>(dl);
        if (std::isfinite(fd) && fd > 0.0f) {
            const unsigned long long micro = static_cast<unsigned long long>(std::min(fd, 1024.0f) * 1000000.0f);
            counts[kCountIslandMaxDeltaMicro] = std::max(counts[kCountIslandMaxDeltaMicro], micro);
        }
    }
}

template <class R> void Solver<R>::self_collision() {
    const int n = V();
    const bool fast = cfg.self_collision_mode == kSelfModeFast;
    const R r = self_radius();
    const R onset = std::max(r * R(1.8), r + R(kEpsF));
    const R rest_skip = std::max(onset * R(1.25), r * R(2.5));
    const R cs = std::max(r * R(1.8), R(1.0e-3));   // A41:6796, 6880, 6944
    const bool vhash = n >= kSelfHashMinCount;
    const bool thash = cfg.triangle_count >= kSelfTriangleHashMinCount;
    const bool ehash = cfg.edge_count >= kSelfEdgeHashMinCount;
    const R prim_margin = r * R(1.8);
    HashTable VH, TH, EH;
    auto tri_valid = [&](int t) {
        int a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        return a < n && b < n && c < n && a != b && b != c && a != c;
    };
    auto edge_valid = [&](int e) { int a = springs[e].i, b = springs[e].j; return a < n && b < n && a != b; };
    auto insert_box = [&](HashTable& H, int prim, V3<R> lo, V3<R> hi, V3<R> fb) {
        int x0 = cell_coord(lo.x, cs), y0 = cell_coord(lo.y, cs), z0 = cell_coord(lo.z, cs);
        int x1 = cell_coord(hi.x, cs), y1 = cell_coord(hi.y, cs), z1 = cell_coord(hi.z, cs);
        long long sx = (long long)x1 - x0 + 1, sy = (long long)y1 - y0 + 1, sz = (long long)z1 - z0 + 1;
        if (sx > 0 && sy > 0 && sz > 0 && sx * sy * sz > kSelfHashMaxPrimitiveCells) {  // A41:871-900
            x0 = x1 = cell_coord(fb.x, cs); y0 = y1 = cell_coord(fb.y, cs); z0 = z1 = cell_coord(fb.z, cs);
            count(kCountSelfOverflow);
        }
        for (int z = z0; z <= z1; ++z)
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    auto& L = H.lists[hash_cell(x, y, z, H.buckets)];
                    if (static_cast<int>(L.size()) >= kBucketSlots) count(kCountSelfOverflow);
                    L.push_back(prim);
                }
    };
    // hash builds, A41:4550-4575, 4632-4694, 5532-5593
    if (vhash) {
        VH.reset(next_pow2(std::max(kMinBuckets, n * 4)));
        for (int i = 0; i < n; ++i) {
            int cx = cell_coord(pos[i].x, cs), cy = cell_coord(pos[i].y, cs), cz = cell_coord(pos[i].z, cs);
            self_cells[3 * i] = cx; self_cells[3 * i + 1] = cy; self_cells[3 * i + 2] = cz;
            auto& L = VH.lists[hash_cell(cx, cy, cz, VH.buckets)];
            if (static_cast<int>(L.size()) >= kBucketSlots) count(kCountSelfOverflow);
            L.push_back(i);
        }
    }
    if (thash) {
        TH.reset(next_pow2(std::max(kMinBuckets, cfg.triangle_count * 4)));
        for (int t = 0; t < cfg.triangle_count; ++t) {
            if (!tri_valid(t)) continue;
            V3<R> a = pos[tris[3 * t]], b = pos[tris[3 * t + 1]], c = pos[tris[3 * t + 2]];
            V3<R> lo = mk(std::min(a.x, std::min(b.x, c.x)) - prim_margin, std::min(a.y, std::min(b.y, c.y)) - prim_margin,
                          std::min(a.z, std::min(b.z, c.z)) - prim_margin);
            V3<R> hi = mk(std::max(a.x, std::max(b.x, c.x)) + prim_margin, std::max(a.y, std::max(b.y, c.y)) + prim_margin,
                          std::max(a.z, std::max(b.z, c.z)) + prim_margin);
            insert_box(TH, t, lo, hi, mul((a + b) + c, R(1.0f / 3.0f)));
        }
    }
    if (ehash) {
        EH.reset(next_pow2(std::max(kMinBuckets, cfg.edge_count * 4)));
        for (int e = 0; e < cfg.edge_count; ++e) {
            if (!edge_valid(e)) continue;
            V3<R> a = pos[springs[e].i], b = pos[springs[e].j];
            V3<R> lo = mk(std::min(a.x, b.x) - prim_margin, std::min(a.y, b.y) - prim_margin, std::min(a.z, b.z) - prim_margin);
            V3<R> hi = mk(std::max(a.x, b.x) + prim_margin, std::max(a.y, b.y) + prim_margin, std::max(a.z, b.z) + prim_margin);
            insert_box(EH, e, lo, hi, mul(a + b, R(0.5)));
        }
    }
    const int passes = std::max(1, std::min(cfg.fast_self_collision_passes, 8));
    int max_nb = std::max(1, std::min(cfg.max_self_collision_neighbors, kBucketSlots));
    if (fast && max_nb > 16) max_nb = 16;  // A41:5133-5135, 5749-5751
    const R radius_half = r * R(0.5);
    const R thickness = radius_half + radius_half;
    auto valid_pair = [&](int s, int o) {  // A41:4696-4717
        if (s == o || r <= R(0) || onset <= r || inv_mass[s] <= 0.0f || pinned[s]) return false;
        V3<R> rd = rest[s] - rest[o];
        if (dot(rd, rd) < rest_skip * rest_skip) return false;
        V3<R> cd = pos[s] - pos[o], pd = prev[s] - prev[o];
        return dot(cd, cd) < onset * onset || dot(pd, pd) < onset * onset;
    };
    auto accumulate = [&](int v, V3<R> d) { self_accumulate(v, d); };
    // A41:762-791; a null state counts the visit only (the brute-force loops)
    auto try_note_seen = [&](SelfFilterState* st, int key) -> bool {
        if (!fast) return true;
        count(kCountFilterSeen);
        if (!st) return true;
        for (int i = 0; i < std::min(st->nseen, kFilterLocalSlots); ++i)
            if (st->seen[i] == key) { count(kCountFilterRejectedDuplicate); return false; }
        if (st->nseen >= kFilterLocalSlots) { count(kCountSelfOverflow); count(kCountFilterRejectedOwnership); return false; }
        st->seen[st->nseen++] = key;
        return true;
    };
    // A41:793-829
    auto try_local_accept = [&](SelfFilterState* st, int type, unsigned key) -> bool {
        if (!fast) return true;
        const int accept_slot = type == kPairVv ? kCountFilterAcceptedVv : (type == kPairVt ? kCountFilterAcceptedVt : kCountFilterAcceptedEe);
        for (int i = 0; i < std::min(st->nclusters, kClusterMaxLocal); ++i) {
            if (st->keys[i] != key) continue;
            if (st->owned[i] >= kClusterMaxContactsPerOwner) { count(kCountFilterRejectedOwnership); return false; }
            st->owned[i] += 1;
            count(kCountClusterOwnedContacts);
            count(accept_slot);
            return true;
        }
        if (st->nclusters >= kClusterMaxLocal) { count(kCountFilterRejectedOwnership); return false; }
        st->keys[st->nclusters] = key;
        st->owned[st->nclusters] = 1;
        st->nclusters += 1;
        count(kCountClusterCount);
        count(kCountClusterOwnedContacts);
        count(accept_slot);
        return true;
    };
    // neighbour table, A41:4738-4788 (hash) / 4790-4814 (brute force)
    auto build_neighbors = [&]() {
        for (int s = 0; s < n; ++s) {
            unsigned acc = 0;
            unsigned* out = &nb_index[static_cast<size_t>(s) * kNeighborSlots];
            if (vhash) {
                int cx = cell_coord(pos[s].x, cs), cy = cell_coord(pos[s].y, cs), cz = cell_coord(pos[s].z, cs);
                for (int dz = -1; dz <= 1 && acc < kNeighborSlots; ++dz)
                    for (int dy = -1; dy <= 1 && acc < kNeighborSlots; ++dy)
                        for (int dx = -1; dx <= 1 && acc < kNeighborSlots; ++dx) {
                            int qx = cx + dx, qy = cy + dy, qz = cz + dz;
                            const auto& L = VH.lists[hash_cell(qx, qy, qz, VH.buckets)];
                            int lim = std::min(static_cast<int>(L.size()), kBucketSlots);
                            for (int k = 0; k < lim && acc < kNeighborSlots; ++k) {
                                int o = L[k];
                                if (self_cells[3 * o] != qx || self_cells[3 * o + 1] != qy || self_cells[3 * o + 2] != qz || !valid_pair(s, o)) continue;
                                out[acc++] = static_cast<unsigned>(o);
                            }
                        }
            } else {
                for (int o = 0; o < n && acc < kNeighborSlots; ++o) if (valid_pair(s, o)) out[acc++] = static_cast<unsigned>(o);
            }
            if (acc >= kNeighborSlots) count(kCountSelfOverflow);
            nb_counts[s] = acc;
        }
    };
    // the fast mode builds the table once per self-collision substep and
    // reuses it on every pass (A41:9291-9301)
    if (fast) {
        build_neighbors();
        if (passes > 1) counts[kCountFilterCacheHits] += static_cast<unsigned long long>(n) * (passes - 1);
        counts[kCountFilterCacheMisses] += static_cast<unsigned long long>(n);
    }
    for (int pass = 0; pass < passes; ++pass) {
        std::fill(acc_q.begin(), acc_q.end(), 0);
        std::fill(acc_w.begin(), acc_w.end(), 0);
        snap_pos = pos;
        snap_prev = prev;
        if (!fast) build_neighbors();
        // vertex-vertex, A41:4816-4951
        for (int s = 0; s < n; ++s) {
            if (inv_mass[s] <= 0.0f || pinned[s] || nb_counts[s] == 0) continue;
            V3<R> pc = pos[s], pp = prev[s];
            SelfFilterState fs;
            for (unsigned k = 0; k < nb_counts[s]; ++k) {
                int o = static_cast<int>(nb_index[static_cast<size_t>(s) * kNeighborSlots + k]);
                if (o < 0 || o >= n || o == s) continue;
                count(fast ? kCountFilterSeen : kCountSelfCandidates);
                V3<R> qc = pos[o], qp = prev[o];
                if (thickness <= R(0)) continue;
                V3<R> resp = mk<R>(0, 0, 0);
                bool hit = false;
                V3<R> dref = pp - qp;
                R cco = dot(dref, dref) - thickness * thickness;
                V3<R> rel = (pc - pp) - (qc - qp);
                R aco = dot(rel, rel), bco = dot(dref, rel);
                const bool other_pinned = inv_mass[o] <= 0.0f || pinned[o];
                const R scale = other_pinned ? R(1) : R(0.5);
                if (cco >= R(0) && bco < R(0) && aco > R(1.0e-5)) {
                    R disc = bco * bco - aco * cco;
                    if (disc >= R(0)) {
                        R t = (-bco - std::sqrt(disc)) / aco;
                        if (t > R(0) && t <= R(1)) {
                            R sv = R(1) - t;
                            V3<R> ph = mul(pp, sv) + mul(pc, t), qh = mul(qp, sv) + mul(qc, t);
                            V3<R> nn = ph - qh;
                            R nsq = dot(nn, nn);
                            if (nsq > R(kEpsF)) {
                                nn = mul(nn, rsqrt(nsq));
                                R vn = dot((pc - ph) - (qc - qh), nn);
                                if (vn > R(0)) { resp = mul(nn, -vn * scale); hit = true; }
                            }
                        }
                    }
                }
                if (!hit) {
                    V3<R> d = pc - qc;
                    R dsq = dot(d, d);
                    if (dsq >= thickness * thickness) continue;
                    V3<R> nn = mk<R>(0, 0, 1);
                    R dist = std::sqrt(std::max(dsq, R(kEpsF)));
                    if (dsq > R(kEpsF)) nn = mul(d, R(1) / dist);
                    else {
                        V3<R> rd = rest[s] - rest[o];
                        R rsq = dot(rd, rd);
                        if (rsq > R(kEpsF)) nn = mul(rd, rsqrt(rsq));
                        dist = R(0);
                    }
                    resp = mul(nn, std::min((thickness - dist) * scale, thickness * R(0.35)));
                    hit = true;
                }
                if (hit && dot(resp, resp) > R(kEpsF)) {
                    R ml = std::max(thickness * R(0.35), R(1.0e-4));
                    R rl = length(resp);
                    if (rl > ml) resp = mul(resp, ml / rl);
                    if (fast) {
                        const V3<R> nn = mul(resp, R(1) / std::max(length(resp), R(kEpsF)));
                        const V3<R> mid = mul(pc + qc, R(0.5));
                        if (!try_local_accept(&fs, kPairVv, filter_cluster_key(kPairVv, mid, nn, std::max(thickness * R(2.5), R(1.0e-3)))))
                            continue;
                        count(kCountSelfCandidates);
                    }
                    accumulate(s, resp);
                    count(kCountExactImpulse);
                }
            }
        }
        // vertex-triangle, A41:4990-5107 pair, 5109-5182 hash, 5184-5224 brute
        auto vt_pair = [&](int v, int t, SelfFilterState* fs) -> bool {
            if (t < 0 || t >= cfg.triangle_count || r <= R(0) || onset <= r || !tri_valid(t)) return false;
            int ia = tris[3 * t], ib = tris[3 * t + 1], ic = tris[3 * t + 2];
            bool rest_near = v == ia || v == ib || v == ic;  // A41:4615-4630
            if (!rest_near) {
                V3<R> rq = closest_point_on_triangle(rest[v], rest[ia], rest[ib], rest[ic]);
                V3<R> rdd = rest[v] - rq;
                rest_near = dot(rdd, rdd) < rest_skip * rest_skip;
            }
            if (rest_near) {
                if (fast) count(kCountFilterRejectedRest);
                return false;
            }
            V3<R> p = pos[v], a = pos[ia], b = pos[ib], c = pos[ic];
            V3<R> q = closest_point_on_triangle(p, a, b, c);
            V3<R> d = p - q;
            R dsq = dot(d, d);
            R dist = std::sqrt(std::max(dsq, R(kEpsF)));
            if (dist >= onset) return false;
            V3<R> nn = mk<R>(0, 0, 1);
            if (dsq > R(kEpsF)) nn = mul(d, R(1) / dist);
            else {
                nn = cross(b - a, c - a);
                R nsq = dot(nn, nn);
                if (nsq > R(kEpsF)) nn = mul(nn, rsqrt(nsq));
                else {
                    V3<R> rdl = rest[v] - closest_point_on_triangle(rest[v], rest[ia], rest[ib], rest[ic]);
                    R rsq = dot(rdl, rdl);
                    if (rsq > R(kEpsF)) nn = mul(rdl, rsqrt(rsq));
                }
                dist = R(0);
            }
            if (fast && !try_local_accept(fs, kPairVt, filter_cluster_key(kPairVt, q, mk<R>(0, 0, 0), std::max(onset * R(12), R(1.0e-3)))))
                return true;
            count(kCountSelfCandidates);
            R wa = R(1) / R(3), wb = wa, wc = wa;
            triangle_barycentric(q, a, b, c, &wa, &wb, &wc);
            R pen = r - dist, push;
            if (pen > R(0)) push = std::min(pen * R(0.35) + r * R(0.01), r * R(0.25));
            else { R x = R(1) - std::min(dist / onset, R(1)); push = std::min(x * x * r * R(0.06), r * R(0.04)); }
            R wv = R(inv_mass[v]), wta = R(inv_mass[ia]) * wa, wtb = R(inv_mass[ib]) * wb, wtc = R(inv_mass[ic]) * wc;
            R total = wv + wta + wtb + wtc;
            if (push <= R(0) || total <= R(0)) return true;
            V3<R> delta = mul(nn, push / total);
            if (wv > R(0)) accumulate(v, mul(delta, wv));
            if (wta > R(0)) accumulate(ia, mul(delta, -wta));
            if (wtb > R(0)) accumulate(ib, mul(delta, -wtb));
            if (wtc > R(0)) accumulate(ic, mul(delta, -wtc));
            count(kCountSoft);
            return true;
        };
        if (cfg.triangle_count > 0) {
            for (int v = 0; v < n; ++v) {
                int accepted = 0;
                SelfFilterState fs;
                if (thash) {
                    int cx = cell_coord(pos[v].x, cs), cy = cell_coord(pos[v].y, cs), cz = cell_coord(pos[v].z, cs);
                    const auto& L = TH.lists[hash_cell(cx, cy, cz, TH.buckets)];
                    int lim = std::min(static_cast<int>(L.size()), kBucketSlots);
                    for (int k = 0; k < lim && accepted < max_nb; ++k) {
                        if (!try_note_seen(&fs, L[k])) continue;
                        if (vt_pair(v, L[k], &fs)) ++accepted;
                    }
                } else {
                    for (int t = 0; t < cfg.triangle_count && accepted < max_nb; ++t) {
                        if (!try_note_seen(nullptr, t)) continue;
                        if (vt_pair(v, t, &fs)) ++accepted;
                    }
                }
            }
        }
        // edge-edge, A41:5595-5722 pair, 5724-5802 hash, 5804-5844 brute
        auto ee_pair = [&](int ea, int eb, SelfFilterState* fs) -> bool {
            if (eb <= ea || ea < 0 || eb >= cfg.edge_count || r <= R(0) || onset <= r) return false;
            if (!edge_valid(ea) || !edge_valid(eb)) return false;
            int a0 = springs[ea].i, a1 = springs[ea].j, b0 = springs[eb].i, b1 = springs[eb].j;
            bool rest_near = a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1;  // A41:5507-5530
            if (!rest_near) {
                R s = 0, t = 0;
                closest_segment_parameters(rest[a0], rest[a1], rest[b0], rest[b1], &s, &t);
                V3<R> pa = rest[a0] + mul(rest[a1] - rest[a0], s), pb = rest[b0] + mul(rest[b1] - rest[b0], t);
                V3<R> dd = pa - pb;
                rest_near = dot(dd, dd) < rest_skip * rest_skip;
            }
            if (rest_near) {
                if (fast) count(kCountFilterRejectedRest);
                return false;
            }
            R s = 0, t = 0;
            closest_segment_parameters(pos[a0], pos[a1], pos[b0], pos[b1], &s, &t);
            V3<R> pa = pos[a0] + mul(pos[a1] - pos[a0], s), pb = pos[b0] + mul(pos[b1] - pos[b0], t);
            V3<R> d = pa - pb;
            R dsq = dot(d, d);
            R dist = std::sqrt(std::max(dsq, R(kEpsF)));
            if (dist >= onset) return false;
            V3<R> nn = mk<R>(0, 0, 1);
            if (dsq > R(kEpsF)) nn = mul(d, R(1) / dist);
            else {
                nn = cross(pos[a1] - pos[a0], pos[b1] - pos[b0]);
                R nsq = dot(nn, nn);
                if (nsq > R(kEpsF)) nn = mul(nn, rsqrt(nsq));
                else {
                    R rs = 0, rt = 0;
                    closest_segment_parameters(rest[a0], rest[a1], rest[b0], rest[b1], &rs, &rt);
                    V3<R> rd = (rest[a0] + mul(rest[a1] - rest[a0], rs)) - (rest[b0] + mul(rest[b1] - rest[b0], rt));
                    R rsq = dot(rd, rd);
                    if (rsq > R(kEpsF)) nn = mul(rd, rsqrt(rsq));
                }
                dist = R(0);
            }
            if (fast && !try_local_accept(fs, kPairEe, filter_cluster_key(kPairEe, mul(pa + pb, R(0.5)), mk<R>(0, 0, 0),
                                                                          std::max(onset * R(12), R(1.0e-3)))))
                return true;
            count(kCountSelfCandidates);
            R pen = r - dist, push;
            if (pen > R(0)) push = std::min(pen * R(0.35) + r * R(0.01), r * R(0.25));
            else { R x = R(1) - std::min(dist / onset, R(1)); push = std::min(x * x * r * R(0.06), r * R(0.04)); }
            R w0 = R(1) - s, w1 = s, w2 = R(1) - t, w3 = t;
            R ma0 = R(inv_mass[a0]), ma1 = R(inv_mass[a1]), mb0 = R(inv_mass[b0]), mb1 = R(inv_mass[b1]);
            R wa0 = ma0 * w0, wa1 = ma1 * w1, wb0 = mb0 * w2, wb1 = mb1 * w3;
            R total = ma0 * w0 * w0 + ma1 * w1 * w1 + mb0 * w2 * w2 + mb1 * w3 * w3;
            if (push <= R(0) || total <= R(0)) return true;
            V3<R> delta = mul(nn, push / total);
            if (wa0 > R(0)) accumulate(a0, mul(delta, wa0));
            if (wa1 > R(0)) accumulate(a1, mul(delta, wa1));
            if (wb0 > R(0)) accumulate(b0, mul(delta, -wb0));
            if (wb1 > R(0)) accumulate(b1, mul(delta, -wb1));
            count(kCountEdgeEdge);
            return true;
        };
        if (cfg.edge_count > 0) {
            for (int e = 0; e < cfg.edge_count; ++e) {
                if (ehash && !edge_valid(e)) continue;
                int accepted = 0;
                SelfFilterState fs;
                if (ehash) {
                    V3<R> mid = mul(pos[springs[e].i] + pos[springs[e].j], R(0.5));
                    int cx = cell_coord(mid.x, cs), cy = cell_coord(mid.y, cs), cz = cell_coord(mid.z, cs);
                    const auto& L = EH.lists[hash_cell(cx, cy, cz, EH.buckets)];
                    int lim = std::min(static_cast<int>(L.size()), kBucketSlots);
                    for (int k = 0; k < lim && accepted < max_nb; ++k) {
                        if (!try_note_seen(&fs, L[k])) continue;
                        if (ee_pair(e, L[k], &fs)) ++accepted;
                    }
                } else {
                    for (int o = e + 1; o < cfg.edge_count && accepted < max_nb; ++o) {
                        if (!try_note_seen(nullptr, o)) continue;
                        if (ee_pair(e, o, &fs)) ++accepted;
                    }
                }
            }
        }
        // overlap islands on the final fast pass, A41:5226-5439
        if (fast && pass == passes - 1 && cfg.triangle_count > 0 && thash) island_aggregate(TH, cs);
        // average + apply, A41:4953-4988
        const R max_delta = std::max(r * R(0.35), R(1.0e-4));
        for (int i = 0; i < n; ++i) {
            if (inv_mass[i] <= 0.0f || pinned[i] || acc_w[i] <= 0) continue;
            const double s = 1.0 / 4294967296.0;
            V3<R> sum = mk<R>(static_cast<R>(static_cast<double>(acc_q[3 * i]) * s), static_cast<R>(static_cast<double>(acc_q[3 * i + 1]) * s),
                              static_cast<R>(static_cast<double>(acc_q[3 * i + 2]) * s));
            V3<R> avg = mul(sum, R(1) / R(acc_w[i]));
            R dl = length(avg);
            if (dl > max_delta) avg = mul(avg, max_delta / dl);
            max_smoothed = std::max(max_smoothed, static_cast<float>(length(avg)));
            pos[i] = pos[i] + avg;
            prev[i] = prev[i] + mul(avg, R(0.55));
        }
    }
}

// ------------------------------------------------------ hard-stretch PCG
// Inner products are summed exactly: every float term is rounded once to a
// 2^-96 fixed-point integer and the 128-bit sum is order independent; the
// sum is then rounded to double.
template <class R> struct ExactSum {
    __int128 acc = 0;
    void add(R t) { acc += static_cast<__int128>(std::rint(static_cast<double>(t) * 79228162514264337593543950336.0)); }
    double value() const { return static_cast<double>(acc) * (1.0 / 79228162514264337593543950336.0); }
};

// A41:1158-1191 symmetric 3x3 inverse, diagonal fallback when singular.
template <class S, class R> S sym_inverse(const S& m, long long* guard) {
    const R a = m.m11, b = m.m12, c = m.m13, d = m.m22, e = m.m23, f = m.m33;
    const R df_ee = std::fma(d, f, -e * e), bf_ce = std::fma(b, f, -c * e), be_cd = std::fma(b, e, -c * d);
    const R det = std::fma(c, be_cd, std::fma(-b, bf_ce, a * df_ee));
    if (std::isfinite(det) && std::fabs(det) > R(1.0e-12)) {
        const R inv = R(1) / det;
        return S{df_ee * inv, std::fma(c, e, -b * f) * inv, be_cd * inv, std::fma(a, f, -c * c) * inv, std::fma(b, c, -a * e) * inv,
                 std::fma(a, d, -b * b) * inv};
    }
    ++*guard;
    return S{R(1) / std::max(std::fabs(a), R(1.0e-4)), R(0), R(0), R(1) / std::max(std::fabs(d), R(1.0e-4)), R(0),
             R(1) / std::max(std::fabs(f), R(1.0e-4))};
}

// A41:1193-1199
template <class S, class R> V3<R> sym_madd(const S& m, V3<R> v, V3<R> acc) {
    return mk(std::fma(m.m13, v.z, std::fma(m.m12, v.y, std::fma(m.m11, v.x, acc.x))),
              std::fma(m.m23, v.z, std::fma(m.m22, v.y, std::fma(m.m12, v.x, acc.y))),
              std::fma(m.m33, v.z, std::fma(m.m23, v.y, std::fma(m.m13, v.x, acc.z))));
}

template <class R> void Solver<R>::pcg_run(R dt) {
    (void)dt;
    const int n = V();
    const R strength_cfg = R(cfg.stretch_optimization_strength);
    const R thick = R(cfg.cloth_thickness);
    const bool hard = cfg.stretch_optimization_strength >= 0.95f && std::isfinite(cfg.cloth_thickness) && cfg.cloth_thickness > kEpsF;
    const Sym I{R(1), R(0), R(0), R(1), R(0), R(1)};
    // A41:3032-3069 reset
    for (int i = 0; i < n; ++i) {
        pcg_diag[i] = I; pcg_pinv[i] = I;
        pcg_rhs[i] = pcg_sol[i] = pcg_res[i] = pcg_z[i] = pcg_dir[i] = pcg_adir[i] = mk<R>(0, 0, 0);
    }
    std::fill(pcg_off.begin(), pcg_off.end(), Sym{});
    // A41:3085-3186 system build, rows summed in ascending edge order
    const R strength = clamp01(strength_cfg * R(2.5));
    std::vector<int> cursor(inc_off.begin(), inc_off.end() - 1);
    for (int e = 0; e < cfg.edge_count; ++e) {
        const Edge& E = springs[e];
        const int i = E.i, j = E.j;
        if (i >= n || j >= n || i == j) continue;
        const int eij = cursor[i]++, eji = cursor[j]++;
        const R wi = weight(i), wj = weight(j);
        const bool ai = wi > R(0), aj = wj > R(0);
        const R w = wi + wj;
        if (!std::isfinite(w) || w <= R(0)) continue;
        const R rest_l = R(E.rest);
        if (!std::isfinite(rest_l) || rest_l <= R(kEpsF)) continue;
        const V3<R> pi = pos[i], pj = pos[j];
        if (!finite(pi) || !finite(pj)) continue;
        const V3<R> delta = pj - pi;
        const R lsq = dot(delta, delta);
        if (!std::isfinite(lsq) || lsq <= R(kEpsF)) continue;
        const R len = std::sqrt(lsq);
        const R err = len - rest_l;
        if (!std::isfinite(err) || std::fabs(err) <= R(1.0e-6)) continue;
        if (strength <= R(0)) continue;
        R limit = std::max(R(1.0e-5), std::min(std::max(rest_l, thick) * R(0.5), R(0.25)));  // A41:1070-1089
        if (hard) limit = std::min(limit, std::max(R(1.0e-5), std::min(std::max(rest_l * R(0.18), thick * R(0.06)), R(0.0035))));
        R proj = err * strength;
        proj = std::max(-limit, std::min(proj, limit));
        if (!std::isfinite(proj) || std::fabs(proj) <= R(1.0e-7)) continue;
        const V3<R> nrm = mul(delta, R(1) / len);
        const V3<R> desired = mul(nrm, proj / w);
        if (ai) pcg_rhs[i] = pcg_rhs[i] + mul(desired, wi);
        if (aj) pcg_rhs[j] = pcg_rhs[j] + mul(desired, -wj);
        const R k = std::max(strength, R(1.0e-3));
        const Sym blk{nrm.x * nrm.x * k, nrm.x * nrm.y * k, nrm.x * nrm.z * k, nrm.y * nrm.y * k, nrm.y * nrm.z * k, nrm.z * nrm.z * k};
        auto addb = [&](Sym& d) { d.m11 += blk.m11; d.m12 += blk.m12; d.m13 += blk.m13; d.m22 += blk.m22; d.m23 += blk.m23; d.m33 += blk.m33; };
        if (ai) addb(pcg_diag[i]);
        if (aj) addb(pcg_diag[j]);
        if (ai && aj) {
            const Sym off{-blk.m11, -blk.m12, -blk.m13, -blk.m22, -blk.m23, -blk.m33};
            pcg_off[eij] = off;
            pcg_off[eji] = off;
        }
    }
    // A41:3188-3226 preconditioner and initial residual
    ExactSum<R> rz;
    for (int i = 0; i < n; ++i) {
        if (inv_mass[i] <= 0.0f || pinned[i]) {
            pcg_rhs[i] = pcg_res[i] = pcg_z[i] = pcg_dir[i] = pcg_sol[i] = mk<R>(0, 0, 0);
            pcg_pinv[i] = I;
            continue;
        }
        pcg_pinv[i] = sym_inverse<Sym, R>(pcg_diag[i], &pcg_guard);
        const V3<R> r = pcg_rhs[i];
        const V3<R> z = sym_madd(pcg_pinv[i], r, mk<R>(0, 0, 0));
        pcg_res[i] = r; pcg_z[i] = z; pcg_dir[i] = z; pcg_sol[i] = mk<R>(0, 0, 0);
        rz.add(dot(r, z));
    }
    pcg_rz = rz.value();
    pcg_initial = pcg_rz;
    pcg_dad = 0.0;
    pcg_rznext = 0.0;
    pcg_status = 0;
    const R wb_limit = hard ? std::max(R(2.5e-4), std::min(thick * R(0.035), R(9.0e-4))) : R(0.25);  // A41:1091-1105
    for (int it = 0; it < 8; ++it) {
        // A41:3228-3267 A*d
        if (pcg_status == 0) {
            ExactSum<R> dad;
            for (int i = 0; i < n; ++i) {
                if (inv_mass[i] <= 0.0f || pinned[i]) { pcg_adir[i] = mk<R>(0, 0, 0); continue; }
                V3<R> q = sym_madd(pcg_diag[i], pcg_dir[i], mk<R>(0, 0, 0));
                for (int c = inc_off[i]; c < inc_off[i + 1]; ++c) {
                    const Edge& E = springs[inc_edge[c]];
                    const int col = E.i == i ? E.j : E.i;
                    q = sym_madd(pcg_off[c], pcg_dir[col], q);
                }
                pcg_adir[i] = q;
                dad.add(dot(pcg_dir[i], q));
            }
            pcg_dad += dad.value();
        }
        // A41:3320-3405 alpha, update, write-back
        int status = 0;
        R alpha = 0;
        if (!std::isfinite(pcg_rz) || pcg_rz < 0.0) status = 1;
        else if (pcg_rz <= 1.0e-14) status = 2;
        else if (!std::isfinite(pcg_dad) || pcg_dad <= 1.0e-14) status = 3;
        else {
            alpha = static_cast<R>(pcg_rz / pcg_dad);
            if (!std::isfinite(alpha)) status = 4;
        }
        pcg_status = status;
        if (status != 0) {
            if (status != 2) ++pcg_guard;
            continue;
        }
        ExactSum<R> rzn;
        for (int i = 0; i < n; ++i) {
            if (inv_mass[i] <= 0.0f || pinned[i]) { pcg_sol[i] = pcg_res[i] = pcg_z[i] = mk<R>(0, 0, 0); continue; }
            const V3<R> d = pcg_dir[i], ad = pcg_adir[i];
            V3<R> sol = mk(std::fma(d.x, alpha, pcg_sol[i].x), std::fma(d.y, alpha, pcg_sol[i].y), std::fma(d.z, alpha, pcg_sol[i].z));
            V3<R> res = mk(std::fma(ad.x, -alpha, pcg_res[i].x), std::fma(ad.y, -alpha, pcg_res[i].y), std::fma(ad.z, -alpha, pcg_res[i].z));
            if (!finite(sol)) sol = mk<R>(0, 0, 0);
            if (!finite(res)) res = mk<R>(0, 0, 0);
            const V3<R> z = sym_madd(pcg_pinv[i], res, mk<R>(0, 0, 0));
            pcg_sol[i] = sol; pcg_res[i] = res; pcg_z[i] = z;
            rzn.add(dot(res, z));
            V3<R> delta = mul(d, alpha);
            const R l = length(delta);
            if (l > wb_limit) { delta = mul(delta, wb_limit / std::max(l, R(kEpsF))); count(kCountHardFallbacks); }
            pos[i] = pos[i] + delta;
            prev[i] = prev[i] + mul(delta, R(0.08));
            pcg_max_delta = std::max(pcg_max_delta, static_cast<float>(length(delta)));
        }
        pcg_rznext += rzn.value();
        if (it + 1 < 8) {
            // A41:3421-3450 search direction, A41:3452-3464 advance
            if (std::isfinite(pcg_rz) && std::isfinite(pcg_rznext) && pcg_rz > 1.0e-14 && pcg_rznext >= 0.0) {
                const R beta = static_cast<R>(pcg_rznext / pcg_rz);
                if (std::isfinite(beta)) {
                    for (int i = 0; i < n; ++i) {
                        if (inv_mass[i] <= 0.0f || pinned[i]) { pcg_dir[i] = mk<R>(0, 0, 0); continue; }
                        const V3<R> d = pcg_dir[i], z = pcg_z[i];
                        pcg_dir[i] = mk(std::fma(d.x, beta, z.x), std::fma(d.y, beta, z.y), std::fma(d.z, beta, z.z));
                    }
                }
            }
            if (std::isfinite(pcg_rznext) && pcg_rznext >= 0.0) {
                pcg_rz = pcg_rznext;
                pcg_dad = 0.0;
                pcg_rznext = 0.0;
            }
        }
    }
    ++pcg_runs;
}

// Edge passes of the hard-stretch polish and caps.  Kinds: 0 = normal cap
// (A41:2249-2362), 1 = tiny-edge cap (A41:2496-2601), 2 = extreme cap
// (A41:2603-2708), 3 = polish contribution (A41:1995-2080), 4 = tiny-edge
// contribution (A41:2399-2466).  Returns false when the edge does nothing;
// otherwise the position and previous-position deltas of both ends (a side
// with zero weight gets apply_* = false).
template <class R> struct EdgeOut { V3<R> di, qi, dj, qj; bool ai, aj; };

template <class R> bool edge_kind(int kind, V3<R> pi, V3<R> pj, V3<R> qi, V3<R> qj, R wi, R wj, R rest, R thick, R strength, R dt,
                                  EdgeOut<R>* out) {
    const R eps = R(kEpsF);
    const bool tiny_kind = kind == 1 || kind == 4;
    if (!std::isfinite(rest) || rest <= eps) return false;
    if (kind != 3 && (tiny_kind ? rest > R(7.5e-4) : rest <= R(7.5e-4))) return false;
    const R w = wi + wj;
    if (!std::isfinite(w) || w <= R(0)) return false;
    if (!finite(pi) || !finite(pj)) return false;
    if (kind != 4 && (!finite(qi) || !finite(qj))) return false;
    const V3<R> edge = pj - pi;
    const R lsq = dot(edge, edge);
    if (!std::isfinite(lsq) || lsq <= R(1.0e-14)) return false;
    const R len = std::sqrt(lsq);
    const bool high = strength >= R(0.95f);
    out->di = out->qi = out->dj = out->qj = mk<R>(0, 0, 0);
    out->ai = out->aj = false;
    if (kind >= 3) {
        const R target = kind == 3 ? rest * (high ? R(1.08f) : R(1.18f)) : rest * R(1.06f);
        const R excess = len - target;
        if (!std::isfinite(excess) || excess <= R(0)) return false;
        const V3<R> corr = mul(edge, (kind == 3 ? -R(0.90f) : -R(1.0f)) * excess / std::max(w * len, eps));
        if (!finite(corr)) return false;
        if (length(corr) <= R(1.0e-20)) return false;
        const R md = kind == 3 ? std::max(R(5.0e-4), std::min(std::max(rest * R(2), thick * R(0.25)), R(0.008)))
                               : std::min(std::max(rest * R(3), R(2.5e-4)), R(8.0e-4));
        if (wi > R(0)) {
            const V3<R> d = limit_delta(mul(corr, -wi), md);
            if (finite(d) && length(d) > R(0)) { out->di = d; out->ai = true; }
        }
        if (wj > R(0)) {
            const V3<R> d = limit_delta(mul(corr, wj), md);
            if (finite(d) && length(d) > R(0)) { out->dj = d; out->aj = true; }
        }
        return out->ai || out->aj;
    }
    R start, target, md, sync;
    if (kind == 0) {
        start = rest * (high ? R(1.12f) : R(1.32f)); target = rest * (high ? R(1.04f) : R(1.22f));
        md = std::min(std::max(rest * R(0.60f), thick * R(0.10f)), R(0.010f)); sync = R(0.08f);
    } else if (kind == 1) {
        start = rest * R(1.35f); target = rest * R(1.12f);
        md = std::min(std::max(rest * R(6.0f), R(5.0e-4)), R(1.5e-3)); sync = R(1.0f);
    } else {
        start = rest * R(1.24f); target = rest * R(1.10f);
        md = std::min(std::max(rest * R(1.25f), thick * R(0.10f)), R(0.006f)); sync = R(0.30f);
    }
    if (!std::isfinite(len) || len <= start) return false;
    const R excess = len - target;
    if (!std::isfinite(excess) || excess <= R(0)) return false;
    const V3<R> corr = mul(edge, -excess / std::max(w * len, eps));
    if (!finite(corr) || length(corr) <= R(1.0e-20)) return false;
    V3<R> pdi = mk<R>(0, 0, 0), pdj = mk<R>(0, 0, 0);
    const V3<R> nrm = mul(edge, R(1) / std::max(len, eps));
    const R sdt = kind == 0 ? std::max(dt, eps) : dt;
    if (std::isfinite(sdt) && sdt > eps) {
        const R inv = R(1) / sdt;
        const V3<R> vi = mul(pi - qi, inv), vj = mul(pj - qj, inv);
        const R sep = dot(vj - vi, nrm);
        if (std::isfinite(sep) && sep > R(0)) {
            const R trim = sep * sdt * R(1.0f) / w;
            if (wi > R(0)) pdi = pdi + mul(nrm, -wi * trim);
            if (wj > R(0)) pdj = pdj + mul(nrm, wj * trim);
        }
    }
    if (wi > R(0)) {
        const V3<R> d = limit_delta(mul(corr, -wi), md);
        if (finite(d) && length(d) > R(0)) { out->di = d; out->qi = pdi + mul(d, sync); out->ai = true; }
    }
    if (wj > R(0)) {
        const V3<R> d = limit_delta(mul(corr, wj), md);
        if (finite(d) && length(d) > R(0)) { out->dj = d; out->qj = pdj + mul(d, sync); out->aj = true; }
    }
    return out->ai || out->aj;
}

// One cap kind over edges [lo, hi) as a Jacobi sweep: every edge reads the
// positions from before the sweep (for a valid colour range this equals the
// reference's race-free per-colour launch).
template <class R> void Solver<R>::edge_jacobi(int kind, int lo, int hi, R dt) {
    const int n = V();
    const std::vector<V3<R>> P = pos, Q = prev;
    const R thick = R(cfg.cloth_thickness), st = R(cfg.stretch_optimization_strength);
    for (int e = lo; e < hi; ++e) {
        const Edge& E = springs[e];
        if (E.i >= n || E.j >= n || E.i == E.j) continue;
        EdgeOut<R> o;
        if (!edge_kind<R>(kind, P[E.i], P[E.j], Q[E.i], Q[E.j], weight(E.i), weight(E.j), R(E.rest), thick, st, dt, &o)) continue;
        if (o.ai) { pos[E.i] = pos[E.i] + o.di; prev[E.i] = prev[E.i] + o.qi; }
        if (o.aj) { pos[E.j] = pos[E.j] + o.dj; prev[E.j] = prev[E.j] + o.qj; }
    }
}

// Colour ranges forward (and then backward) `passes` times (A41:7761-7803);
// without a colouring, one range over all edges.
template <class R> void Solver<R>::cap_ranges(int kind, int passes, bool reverse_too, R dt) {
    if (cfg.stretch_optimization_strength < 0.95f || cfg.edge_count <= 0) return;
    for (int p = 0; p < passes; ++p) {
        if (edge_colors.empty()) { edge_jacobi(kind, 0, cfg.edge_count, dt); continue; }
        for (int c = 0; c < cfg.edge_color_count; ++c)
            edge_jacobi(kind, std::max(edge_colors[c], 0), std::min(edge_colors[c + 1], cfg.edge_count), dt);
        if (reverse_too)
            for (int c = cfg.edge_color_count - 1; c >= 0; --c)
                edge_jacobi(kind, std::max(edge_colors[c], 0), std::min(edge_colors[c + 1], cfg.edge_count), dt);
    }
}

// Polish / tiny-edge averaging: sum the edge contributions per vertex in
// ascending edge order, then apply the clamped average (A41:2082-2112,
// 2468-2494).
template <class R> void Solver<R>::gather_average(int kind, R dt) {
    const int n = V();
    const std::vector<V3<R>> P = pos, Q = prev;
    const R thick = R(cfg.cloth_thickness), st = R(cfg.stretch_optimization_strength);
    std::vector<V3<R>> sum(n, mk<R>(0, 0, 0));
    std::vector<int> cnt(n, 0);
    for (int e = 0; e < cfg.edge_count; ++e) {
        const Edge& E = springs[e];
        if (E.i >= n || E.j >= n || E.i == E.j) continue;
        EdgeOut<R> o;
        if (!edge_kind<R>(kind, P[E.i], P[E.j], Q[E.i], Q[E.j], weight(E.i), weight(E.j), R(E.rest), thick, st, dt, &o)) continue;
        if (o.ai) { sum[E.i] = sum[E.i] + o.di; cnt[E.i]++; }
        if (o.aj) { sum[E.j] = sum[E.j] + o.dj; cnt[E.j]++; }
    }
    const R md = kind == 3 ? std::max(R(5.0e-4), std::min(std::max(thick * R(0.20f), R(1.0e-3)), R(0.005f)))
                           : std::min(std::max(thick * R(0.015f), R(2.5e-4)), R(9.0e-4));
    const R sync = kind == 3 ? R(0.16f) : R(0.24f);
    for (int i = 0; i < n; ++i) {
        if (inv_mass[i] <= 0.0f || pinned[i] || cnt[i] <= 0) continue;
        V3<R> d = mul(sum[i], R(1) / R(cnt[i]));
        if (!finite(d)) { count(kCountHardFallbacks); continue; }
        d = limit_delta(d, md);
        if (!finite(d) || length(d) <= R(0)) continue;
        pos[i] = pos[i] + d;
        prev[i] = prev[i] + mul(d, sync);
    }
}

// A41:7823-7857
template <class R> void Solver<R>::polish(R dt) {
    if (cfg.edge_count <= 0 || cfg.stretch_optimization_strength < 0.95f) return;
    for (int p = 0; p < 2; ++p) {
        gather_average(3, dt);
        cap_ranges(0, 2, true, dt);
        for (int t = 0; t < 2; ++t) {
            gather_average(4, dt);
            cap_ranges(1, 3, false, dt);
        }
    }
}

// A41:7997-8027
template <class R> void Solver<R>::final_caps(R dt) {
    cap_ranges(0, 2, true, dt);
    cap_ranges(2, 3, false, dt);
    cap_ranges(1, 3, false, dt);
}

// A41:9087-9398
// A41:4251-4323 (triangles) then A41:4415-4511 (particles), unsigned
// contacts: the deepest contact of each kind pushes the vertex by half its
// depth (A41:129).  Only the vertex's own position is read and written.
template <class R> void Solver<R>::dynamic_contacts(R dt) {
    const DynamicSet& D = use_scene() ? scene->set : dyn;
    const int mode = use_scene() ? cross_mode : kCrossOff;
    auto enabled = [&](int src, int layer) {   // A41:3984-3997
        if (mode == kCrossOff) return true;
        if (target_src != 0 && src == target_src) return false;
        if (mode == kCrossLower) return layer < target_layer;
        return true;
    };
    const R eps = R(kEpsF);
    const R cm = std::max(R(cfg.collision_margin), R(0));
    if (D.nt > 0 && cm > R(0)) {
        for (int i = 0; i < V(); ++i) {
            if (w_zero(i)) continue;
            const Vec p = pos[i];
            Vec best_delta = mk<R>(0, 0, 0);
            R best_pen = R(0);
            unsigned long long cand = 0, rej = 0;
            auto consider = [&](int t) {   // A41:3999-4113
                if (t < 0 || t >= D.nt) return;
                if (!D.tsrc.empty() && !enabled(D.tsrc[t], D.tlayer.empty() ? 0 : D.tlayer[t])) return;
                ++cand;
                const Vec mn = mk<R>(D.lo[3 * t], D.lo[3 * t + 1], D.lo[3 * t + 2]);
                const Vec mx = mk<R>(D.hi[3 * t], D.hi[3 * t + 1], D.hi[3 * t + 2]);
                const Vec q0 = mk(std::min(std::max(p.x, mn.x), mx.x), std::min(std::max(p.y, mn.y), mx.y), std::min(std::max(p.z, mn.z), mx.z));
                const Vec da = p - q0;
                if (dot(da, da) > cm * cm) { ++rej; return; }
                const Vec cen = mk<R>(D.ctr[3 * t], D.ctr[3 * t + 1], D.ctr[3 * t + 2]);
                const R rr = R(D.rad[t]) + cm;
                const Vec dc = p - cen;
                if (dot(dc, dc) > rr * rr) { ++rej; return; }
                const float* T = &D.tv[9 * static_cast<size_t>(t)];
                const Vec a = mk<R>(T[0], T[1], T[2]), b = mk<R>(T[3], T[4], T[5]), c = mk<R>(T[6], T[7], T[8]);
                const Vec q = closest_point_on_triangle(p, a, b, c);
                const Vec d = p - q;
                const R dsq = dot(d, d);
                const R dist = std::sqrt(std::max(dsq, eps));
                const Vec tn = cross(b - a, c - a);
                const R tl = std::sqrt(std::max(dot(tn, tn), eps));
                const Vec tnorm = mul(tn, R(1) / tl);
                const Vec n = dsq > eps * R(4) ? mul(d, R(1) / std::max(dist, eps)) : tnorm;
                const R pen = cm - dist;
                if (dist >= cm) return;
                if (pen > best_pen) { best_pen = pen; best_delta = mul(n, pen); }
            };
            if (D.tri_hash) {   // A41:4115-4249, own cell only (unsigned query radius 0)
                const int cx = cell_coord<R>(p.x, R(D.tri_cell)), cy = cell_coord<R>(p.y, R(D.tri_cell)), cz = cell_coord<R>(p.z, R(D.tri_cell));
                const unsigned bk = hash_cell(cx, cy, cz, D.tri_buckets);
                for (int k = D.tri_off[bk]; k < D.tri_off[bk + 1]; ++k) {
                    const int t = D.tri_idx[k];
                    if (cx < D.cmin[3 * t] || cx > D.cmax[3 * t] || cy < D.cmin[3 * t + 1] || cy > D.cmax[3 * t + 1] ||
                        cz < D.cmin[3 * t + 2] || cz > D.cmax[3 * t + 2])
                        continue;   // A41:3934-3982: the query cell is not this triangle's canonical cell
                    consider(t);
                }
                for (int t : D.large) consider(t);
            } else {
                for (int t = 0; t < D.nt; ++t) consider(t);
            }
            count(kCountDynTriCandidates, cand);
            count(kCountDynTriAabbRejects, rej);
            if (best_pen > R(0)) {
                count(kCountTrianglePairs);
                const Vec applied = mul(best_delta, R(kNeighborImpulseScale));
                pos[i] = p + applied;
                const R vn = dot(vel[i], applied);
                if (vn < R(0)) {
                    const R lsq = std::max(dot(applied, applied), eps);
                    vel[i] = vel[i] - mul(applied, R(kNeighborImpulseScale) * vn / lsq);
                }
                prev[i] = pos[i] - mul(vel[i], std::max(dt, eps));
                count(kCountExactImpulse);
            }
        }
    }
    if (D.np > 0) {
        const R margin = std::max(R(cfg.collision_margin), R(0));
        for (int i = 0; i < V(); ++i) {
            if (w_zero(i)) continue;
            const Vec p = pos[i];
            Vec best_delta = mk<R>(0, 0, 0);
            R best_pen = R(0);
            auto consider = [&](int j) {   // A41:4354-4413
                if (j < 0 || j >= D.np) return;
                if (!D.psrc.empty() && !enabled(D.psrc[j], D.player.empty() ? 0 : D.player[j])) return;
                const Vec cen = mk<R>(D.pp[3 * j], D.pp[3 * j + 1], D.pp[3 * j + 2]);
                const R radius = R(std::fmax(D.pr[j], 0.0f)) + margin;
                if (radius <= R(0)) return;
                const Vec d = p - cen;
                const R dsq = dot(d, d);
                const R dist = std::sqrt(std::max(dsq, eps));
                if (dist >= radius) return;
                count(kCountDynParticleCandidates);
                const Vec n = dist > eps ? mul(d, R(1) / dist) : mk<R>(0, 0, 1);
                const R pen = radius - dist;
                if (pen > best_pen) { best_pen = pen; best_delta = mul(n, pen); }
            };
            if (D.part_hash) {
                const int cx = cell_coord<R>(p.x, R(D.part_cell)), cy = cell_coord<R>(p.y, R(D.part_cell)),
                          cz = cell_coord<R>(p.z, R(D.part_cell));
                for (int dz = -1; dz <= 1; ++dz)
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int qx = cx + dx, qy = cy + dy, qz = cz + dz;
                            const unsigned bk = hash_cell(qx, qy, qz, D.part_buckets);
                            const int lim = std::min(D.pcount[bk], kParticleBucketSlots);
                            for (int k = 0; k < lim; ++k) {
                                const int j = D.pidx[bk * kParticleBucketSlots + k];
                                if (D.pcell[3 * j] != qx || D.pcell[3 * j + 1] != qy || D.pcell[3 * j + 2] != qz) continue;
                                consider(j);
                            }
                        }
            } else {
                for (int j = 0; j < D.np; ++j) consider(j);
            }
            if (best_pen > R(0)) {
                pos[i] = p + mul(best_delta, R(kNeighborImpulseScale));
                count(kCountExactImpulse);
                count(kCountDynParticleContacts);
            }
        }
    }
}

// A41:995-1068: the signed contact response.  The push is a third of the
// (locally averaged) contact delta, clamped; the velocity keeps the
// collider's normal speed, loses the inward part, takes the static-contact
// friction and keeps an eighth of any outward speed.  The inside-only passes
// just remove the inward part.
template <class R>
void Solver<R>::signed_response(int i, V3<R> orig, V3<R> raw, V3<R> cvel, R proxy_depth, R dt, bool inside) {
    const R eps = R(kEpsF);
    if (i < 0 || i >= V() || !finite(orig) || !finite(raw)) return;
    const R raw_len = length(raw);
    if (!std::isfinite(raw_len) || raw_len <= eps) return;
    const V3<R> n = mul(raw, R(1) / raw_len);
    V3<R> applied = mul(raw, R(inside ? kSignedInsideRelaxation : kSignedRelaxation));
    {   // A41:951-968
        const R basis = std::max(std::max(R(cfg.cloth_thickness), R(0)), std::max(R(cfg.collision_margin), R(0)));
        const R max_len = inside ? std::max(R(kSignedInsideMaxDeltaFloor), basis * R(kSignedInsideMaxDeltaScale))
                                 : std::max(R(kDynamicMaxDeltaFloor), basis * R(kDynamicMaxDeltaScale));
        if (!finite(applied) || !std::isfinite(max_len) || max_len <= R(0)) applied = mk<R>(0, 0, 0);
        else {
            const R len = length(applied);
            if (!std::isfinite(len) || len <= eps) applied = mk<R>(0, 0, 0);
            else if (len > max_len) applied = mul(applied, max_len / len);
        }
    }
    const R applied_len = length(applied);
    if (!std::isfinite(applied_len) || applied_len <= eps) return;
    const R sdt = std::max(dt, eps);
    const V3<R> corrected = orig + applied;
    V3<R> v = vel[i];
    V3<R> base = mk<R>(0, 0, 0);
    if (inside) v = finite(v) ? v : mk<R>(0, 0, 0);
    if (finite(cvel)) {
        const R cn = dot(cvel, n);
        if (std::isfinite(cn)) base = mul(n, cn);
    }
    V3<R> rel = v - base;
    if (inside) {
        const R ns = dot(rel, n);
        if (std::isfinite(ns) && ns < R(0)) {
            rel = rel - mul(n, ns);
            v = base + rel;
        }
    } else {
        const R inward = std::max(-dot(rel, n), R(0));
        const R proxy = std::max(std::max(proxy_depth, raw_len), applied_len);
        contact_response(n, inward + proxy / sdt, &rel);
        const R outward = dot(rel, n);
        if (outward > R(0)) rel = rel - mul(n, outward * (R(1) - R(kSignedOutwardRetention)));
        v = base + rel;
    }
    pos[i] = corrected;
    vel[i] = finite(v) ? v : mk<R>(0, 0, 0);
    prev[i] = corrected - mul(vel[i], sdt);
}

// A41:3600-3633, 3999-4323, 4354-4511, 4520-4540: contacts with the cloth's
// own indexed (signed) triangles and its particles.  The triangle query
// visits every hash cell within the query radius (one cell, or up to four
// for the inside-only passes) and takes each triangle at its canonical cell.
// The contact kept is the nearest triangle (the first of equal distance)
// whose signed distance is under the contact margin (or, inside-only, under
// 1 mm behind the face); its push direction is the depth-weighted mean of all
// contacts scaled to its own depth.  The particle part (not inside-only) keeps
// the deepest particle the same way.
template <class R> void Solver<R>::signed_contacts(R dt, bool inside) {
    const DynamicSet& D = dyn;
    const R eps = R(kEpsF);
    const R cdt = (std::isfinite(cfg.dt) && cfg.dt > kEpsF) ? R(cfg.dt) : std::max(dt, eps);   // A41:4513-4518
    auto local_average = [&](V3<R> acc, R w, V3<R> fallback) {   // A41:978-993
        if (w > R(0) && finite(acc)) {
            const V3<R> avg = mul(acc, R(1) / w);
            const R al = length(avg), fl = length(fallback);
            if (finite(avg) && std::isfinite(al) && std::isfinite(fl) && al > eps && fl > eps) return mul(avg, fl / al);
        }
        return fallback;
    };
    const R cm = std::max(std::max(R(cfg.collision_margin), R(0)), std::max(R(cfg.cloth_thickness), R(0)) * R(kSignedContactMarginScale));
    const R qm = std::min(std::max(std::max(std::max(cm * R(2), R(kSignedQueryMargin)), std::max(R(D.tri_cell), R(0)) * R(1.25)), cm),
                          R(kSignedMaxQueryMargin));
    if (D.nt > 0 && cm > R(0) && qm > R(0)) {
        const R cs = std::max(R(D.tri_cell), R(1.0e-3));
        const int radius = inside ? std::min(std::max(static_cast<int>(std::ceil(std::max(qm, R(0)) / cs)), 1), 4) : 1;
        for (int i = 0; i < V(); ++i) {
            if (w_zero(i)) continue;
            const V3<R> p = pos[i];
            V3<R> best_delta = mk<R>(0, 0, 0), best_vel = mk<R>(0, 0, 0), acc = mk<R>(0, 0, 0);
            R best_pen = R(0), best_dist = R(3.402823466e+38f), accw = R(0);
            unsigned long long cand = 0, rej = 0;
            auto consider = [&](int t) {   // A41:3999-4113, signed branch
                if (t < 0 || t >= D.nt) return;
                ++cand;
                const V3<R> mn = mk<R>(D.lo[3 * t], D.lo[3 * t + 1], D.lo[3 * t + 2]);
                const V3<R> mx = mk<R>(D.hi[3 * t], D.hi[3 * t + 1], D.hi[3 * t + 2]);
                const V3<R> q0 = mk(std::min(std::max(p.x, mn.x), mx.x), std::min(std::max(p.y, mn.y), mx.y), std::min(std::max(p.z, mn.z), mx.z));
                const V3<R> da = p - q0;
                if (dot(da, da) > qm * qm) { ++rej; return; }
                const V3<R> cen = mk<R>(D.ctr[3 * t], D.ctr[3 * t + 1], D.ctr[3 * t + 2]);
                const R rr = R(D.rad[t]) + qm;
                const V3<R> dc = p - cen;
                if (dot(dc, dc) > rr * rr) { ++rej; return; }
                const float* T = &D.tv[9 * static_cast<size_t>(t)];
                const V3<R> a = mk<R>(T[0], T[1], T[2]), b = mk<R>(T[3], T[4], T[5]), c = mk<R>(T[6], T[7], T[8]);
                const V3<R> q = closest_point_on_triangle(p, a, b, c);
                const V3<R> d = p - q;
                const R dsq = dot(d, d);
                const R dist = std::sqrt(std::max(dsq, eps));
                const V3<R> tn = cross(b - a, c - a);
                const R tl = std::sqrt(std::max(dot(tn, tn), eps));
                const V3<R> tnorm = mul(tn, R(1) / tl);
                const R sd = dot(d, tnorm);
                R pen;
                if (inside) {
                    const R target = -R(kSignedInsideTargetMargin);
                    if (sd >= target) return;
                    pen = target - sd;
                } else {
                    pen = cm - sd;
                }
                if (pen <= R(0)) return;
                const V3<R> delta = mul(tnorm, pen);
                if (!inside && finite(delta)) {
                    const R w = std::max(pen, eps);
                    acc = acc + mul(delta, w);
                    accw += w;
                }
                if (dist < best_dist) {
                    V3<R> cv = mk<R>(0, 0, 0);
                    if (D.tri_motion && cdt > eps) {
                        const V3<R> c0 = mk<R>(D.prev_ctr[3 * t], D.prev_ctr[3 * t + 1], D.prev_ctr[3 * t + 2]);
                        cv = mul(cen - c0, R(1) / cdt);
                    }
                    best_dist = dist;
                    best_pen = pen;
                    best_delta = delta;
                    best_vel = cv;
                }
            };
            if (D.tri_hash) {   // A41:4115-4249
                const int cx = cell_coord<R>(p.x, R(D.tri_cell)), cy = cell_coord<R>(p.y, R(D.tri_cell)), cz = cell_coord<R>(p.z, R(D.tri_cell));
                for (int dz = -radius; dz <= radius; ++dz)
                    for (int dy = -radius; dy <= radius; ++dy)
                        for (int dx = -radius; dx <= radius; ++dx) {
                            const int qx = cx + dx, qy = cy + dy, qz = cz + dz;
                            const unsigned bk = hash_cell(qx, qy, qz, D.tri_buckets);
                            for (int k = D.tri_off[bk]; k < D.tri_off[bk + 1]; ++k) {
                                const int t = D.tri_idx[k];
                                if (t < 0 || t >= D.nt) continue;
                                // A41:3934-3982: the triangle's canonical cell for this vertex
                                const int kx = std::min(std::max(cx, D.cmin[3 * t]), D.cmax[3 * t]);
                                const int ky = std::min(std::max(cy, D.cmin[3 * t + 1]), D.cmax[3 * t + 1]);
                                const int kz = std::min(std::max(cz, D.cmin[3 * t + 2]), D.cmax[3 * t + 2]);
                                if (qx != kx || qy != ky || qz != kz) continue;
                                consider(t);
                            }
                        }
                for (int t : D.large) consider(t);
            } else {
                for (int t = 0; t < D.nt; ++t) consider(t);
            }
            count(kCountDynTriCandidates, cand);
            count(kCountDynTriAabbRejects, rej);
            if (best_pen > R(0)) {
                count(kCountTrianglePairs);
                signed_response(i, p, local_average(acc, accw, best_delta), best_vel, best_pen, dt, inside);
                count(kCountExactImpulse);
            }
        }
    }
    if (inside || D.np <= 0) return;
    const R margin = std::max(R(cfg.collision_margin), R(0));
    for (int i = 0; i < V(); ++i) {
        if (w_zero(i)) continue;
        const V3<R> p = pos[i];
        V3<R> best_delta = mk<R>(0, 0, 0), best_vel = mk<R>(0, 0, 0), acc = mk<R>(0, 0, 0);
        R best_pen = R(0), accw = R(0);
        auto consider = [&](int j) {   // A41:4354-4413, signed branch
            if (j < 0 || j >= D.np) return;
            const V3<R> cen = mk<R>(D.pp[3 * j], D.pp[3 * j + 1], D.pp[3 * j + 2]);
            const R radius = R(std::fmax(D.pr[j], 0.0f)) + margin;
            if (radius <= R(0)) return;
            const V3<R> d = p - cen;
            const R dsq = dot(d, d);
            const R dist = std::sqrt(std::max(dsq, eps));
            if (dist >= radius) return;
            count(kCountDynParticleCandidates);
            const V3<R> n = dist > eps ? mul(d, R(1) / dist) : mk<R>(0, 0, 1);
            const R pen = radius - dist;
            const V3<R> delta = mul(n, pen);
            if (finite(delta)) {
                const R w = std::max(pen, eps);
                acc = acc + mul(delta, w);
                accw += w;
            }
            if (pen > best_pen) {
                V3<R> pv = mk<R>(0, 0, 0);
                if (D.part_motion && cdt > eps) pv = mul(cen - mk<R>(D.pprev[3 * j], D.pprev[3 * j + 1], D.pprev[3 * j + 2]), R(1) / cdt);
                best_pen = pen;
                best_delta = delta;
                best_vel = pv;
            }
        };
        if (D.part_hash) {
            const int cx = cell_coord<R>(p.x, R(D.part_cell)), cy = cell_coord<R>(p.y, R(D.part_cell)), cz = cell_coord<R>(p.z, R(D.part_cell));
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int qx = cx + dx, qy = cy + dy, qz = cz + dz;
                        const unsigned bk = hash_cell(qx, qy, qz, D.part_buckets);
                        const int lim = std::min(D.pcount[bk], kParticleBucketSlots);
                        for (int k = 0; k < lim; ++k) {
                            const int j = D.pidx[bk * kParticleBucketSlots + k];
                            if (D.pcell[3 * j] != qx || D.pcell[3 * j + 1] != qy || D.pcell[3 * j + 2] != qz) continue;
                            consider(j);
                        }
                    }
        } else {
            for (int j = 0; j < D.np; ++j) consider(j);
        }
        if (best_pen > R(0)) {
            signed_response(i, p, local_average(acc, accw, best_delta), best_vel, best_pen, dt, false);
            count(kCountExactImpulse);
            count(kCountDynParticleContacts);
        }
    }
}

template <class R> bool Solver<R>::step(int substeps, int iterations, int fetch) {
    auto t0 = std::chrono::steady_clock::now();
    diag = SsblXpbdDiagnostics{};
    diag.finite_flag = 1;
    diag.static_triangle_count = cfg.static_triangle_count;
    diag.force_field_count = static_cast<long long>(fields.size());
    diag.unsupported_force_field_count = unsupported_fields;
    std::memset(counts, 0, sizeof(counts));
    diag.dynamic_triangle_count = use_scene() ? scene->set.nt : dyn.nt;   // A41:9103-9122
    diag.dynamic_particle_count = use_scene() ? scene->set.np : dyn.np;
    if (use_scene()) {
        diag.global_dynamic_particle_count = scene->set.np;
        diag.global_dynamic_triangle_count = scene->set.nt;
        diag.global_dynamic_hash_overflow = scene->hash_overflow;
    }
    if (!use_scene() || dyn.signed_set) dyn.build(counts);                  // A41:9129-9136
    // a signed own set takes the place of the unsigned contacts (A41:9240-9252)
    const bool dynamic_on = use_scene() ? (scene->set.nt > 0 || scene->set.np > 0) : (!dyn.signed_set && (dyn.nt > 0 || dyn.np > 0));
    const bool signed_on = dyn.signed_set && (dyn.nt > 0 || dyn.np > 0);
    max_smoothed = 0.0f;
    pcg_guard = 0;
    pcg_runs = 0;
    pcg_max_delta = 0.0f;
    pcg_rz = pcg_dad = pcg_rznext = pcg_initial = 0.0;
    pcg_status = 0;
    substeps = std::max(substeps, 1);
    iterations = std::max(iterations, 1);
    const R sub_dt = R(cfg.dt / static_cast<float>(substeps));
    const R exponent = R(1.0f / static_cast<float>(std::max(iterations + 1, 1)));
    const bool analytic_on = colliders.use_ground || colliders.use_wall || colliders.use_sphere;
    const R r = self_radius();
    for (int s = 0; s < substeps; ++s) {
        integrate(sub_dt);
        if (!pins.empty()) pin_project(exponent, sub_dt);
        for (int it = 0; it < iterations; ++it) {
            const bool last_it = it == iterations - 1;
            const bool pcg_on = cfg.stretch_optimization_enabled != 0;
            if (cfg.edge_count > 0 && !pcg_on) {
                if (stretch_mode == SSBL_EXT_STRETCH_COLORED) stretch_colored(sub_dt);
                else stretch_jacobi(sub_dt);
            }
            // A41:9158-9177: every cadence-th substep and always the last
            const int cadence = std::max(1, std::min(pcg_cadence, substeps));
            const bool run_pcg = pcg_on && last_it && (((s + 1) % cadence) == 0 || s == substeps - 1);
            if (run_pcg) {
                pcg_run(sub_dt);
                polish(sub_dt);
                cap_ranges(2, 3, false, sub_dt);   // extreme caps, A41:7974-7995
            }
            const bool final_c = s == substeps - 1 && last_it;
            if (final_c && cfg.lra_count > 0) lra_jacobi(sub_dt);
            if (final_c && cfg.bend_count > 0) bend_pass(sub_dt);
            if (!pins.empty()) pin_project(exponent, sub_dt);
            if (analytic_on) analytic(sub_dt);
            if (sdf_ready) static_sdf_contacts(sub_dt);
            if (last_it && dynamic_on) dynamic_contacts(sub_dt);                 // A41:9230-9252
            const int interval = std::max(cfg.self_collision_interval, 1);
            const bool self_it = cfg.self_collision && last_it;
            const bool scheduled = self_it && (((s + 1) % interval) == 0 || s == substeps - 1);
            const bool run_self = scheduled && V() > 0 && r > R(0);
            if ((scheduled && !run_self) || (self_it && !scheduled)) diag.self_collision_skipped_launches += 1;
            if (run_self) self_collision();
            if (last_it && signed_on) signed_contacts(sub_dt, false);            // A41:9322-9327
            if (run_pcg && final_c) final_caps(sub_dt);
            if (final_c && signed_on)                                             // A41:9351-9359
                for (int k = 0; k < kSignedInsidePasses; ++k) signed_contacts(sub_dt, true);
            // the reference's final pin polish also runs on every last iteration
            // because dynamic collision is scheduled there (A41:9230, 9360-9369)
            const bool polish = !pins.empty() && (analytic_on || sdf_ready || last_it || run_self);
            if (polish) pin_project(exponent, sub_dt);
        }
        velocity_update(sub_dt);
    }
    dyn.tri_motion = dyn.part_motion = false;   // A41:9389-9390
    if (fetch) {
        diag.abi41_soft_contact_count = static_cast<long long>(counts[kCountSoft]);
        diag.abi41_exact_impulse_contact_count = static_cast<long long>(counts[kCountExactImpulse]);
        diag.abi41_edge_edge_contact_count = static_cast<long long>(counts[kCountEdgeEdge]);
        diag.abi41_hard_projection_fallbacks = static_cast<long long>(counts[kCountHardFallbacks]);
        diag.abi41_max_smoothed_delta = max_smoothed;
        diag.fast_soft_repulsion_candidates = static_cast<long long>(counts[kCountSelfCandidates]);
        diag.fast_soft_repulsion_applied = static_cast<long long>(counts[kCountSoft]);
        diag.external_friction_corrections = static_cast<long long>(counts[kCountFriction]);
        diag.self_candidate_count = static_cast<long long>(counts[kCountSelfCandidates]);
        diag.static_sdf_contact_count = static_cast<long long>(counts[kCountStaticSdf]);
        diag.candidate_count = static_cast<long long>(counts[kCountDynParticleCandidates] + counts[kCountDynTriCandidates] +
                                                      counts[kCountSelfCandidates] + counts[kCountStaticSdf]);
        // A41:8078-8110
        diag.dynamic_particle_candidate_count = static_cast<long long>(counts[kCountDynParticleCandidates]);
        diag.dynamic_particle_contacts = static_cast<long long>(counts[kCountDynParticleContacts]);
        diag.dynamic_particle_overflow = 0;
        diag.dynamic_triangle_candidate_count = static_cast<long long>(counts[kCountDynTriCandidates]);
        diag.dynamic_triangle_bucket_overflow = static_cast<long long>(counts[kCountDynTriBucketOverflow]);
        diag.dynamic_triangle_large_primitive_count = static_cast<long long>(counts[kCountDynTriLarge]);
        diag.dynamic_triangle_aabb_reject_count = static_cast<long long>(counts[kCountDynTriAabbRejects]);
        diag.dynamic_triangle_max_bucket_occupancy = 0;
        if (dyn_tri_allocated) {
            diag.dynamic_triangle_large_primitive_count =
                std::max(diag.dynamic_triangle_large_primitive_count, static_cast<long long>(dyn.large.size()));
            diag.dynamic_triangle_max_bucket_occupancy = dyn.max_occupancy;
        }
        diag.fast_triangle_pair_candidates = diag.dynamic_triangle_candidate_count;
        diag.fast_triangle_pair_contacts = static_cast<long long>(counts[kCountTrianglePairs]);
        diag.fast_triangle_pair_skipped_rest = diag.dynamic_triangle_aabb_reject_count;
        diag.abi41_lra_tack_count = static_cast<long long>(counts[kCountLraTacks]);
        diag.abi41_bending_wing_count = counts[kCountBendingWings] > 0 ? static_cast<long long>(counts[kCountBendingWings]) : wing_count;
        diag.abi41_tack_jitter_guarded = static_cast<long long>(counts[kCountTackGuards]);
        diag.abi41_bending_guarded = static_cast<long long>(counts[kCountBendingGuards]);
        // A41:8128-8160
        diag.fast_overlap_island_candidates = static_cast<long long>(counts[kCountIslandCandidates]);
        diag.fast_overlap_island_clusters = static_cast<long long>(counts[kCountIslandClusters]);
        diag.fast_overlap_island_vertex_refs = static_cast<long long>(counts[kCountIslandVertexRefs]);
        diag.fast_overlap_island_applied_vertices = static_cast<long long>(counts[kCountIslandApplied]);
        diag.fast_overlap_island_guarded = static_cast<long long>(counts[kCountIslandGuarded]);
        diag.fast_overlap_island_max_delta = static_cast<float>(counts[kCountIslandMaxDeltaMicro]) * 1.0e-6f;
        diag.self_filter_seen = static_cast<long long>(counts[kCountFilterSeen]);
        diag.self_filter_accepted_vv = static_cast<long long>(counts[kCountFilterAcceptedVv]);
        diag.self_filter_accepted_vt = static_cast<long long>(counts[kCountFilterAcceptedVt]);
        diag.self_filter_accepted_ee = static_cast<long long>(counts[kCountFilterAcceptedEe]);
        diag.self_filter_rejected_rest = static_cast<long long>(counts[kCountFilterRejectedRest]);
        diag.self_filter_rejected_duplicate = static_cast<long long>(counts[kCountFilterRejectedDuplicate]);
        diag.self_filter_rejected_ownership = static_cast<long long>(counts[kCountFilterRejectedOwnership]);
        diag.self_filter_cache_hits = static_cast<long long>(counts[kCountFilterCacheHits]);
        diag.self_filter_cache_misses = static_cast<long long>(counts[kCountFilterCacheMisses]);
        diag.self_cluster_count = static_cast<long long>(counts[kCountClusterCount]);
        diag.self_cluster_owned_contacts = static_cast<long long>(counts[kCountClusterOwnedContacts]);
    }
    if (cfg.stretch_optimization_enabled) {  // A41:7586-7624, 8159-8176
        diag.abi41_pcg_csr_nnz = inc_off[cfg.vertex_count];
        diag.abi41_pcg_texture_ready = 1;
        diag.abi41_pcg_iterations = 8 * pcg_runs;
        if (fetch && pcg_runs > 0) {
            double init = pcg_initial;
            if ((!std::isfinite(init) || init <= 0.0) && std::isfinite(pcg_rz) && pcg_rz >= 0.0) init = pcg_rz;
            if (std::isfinite(init) && init >= 0.0) {
                diag.abi41_pcg_initial_residual = static_cast<float>(std::sqrt(std::max(init, 0.0)));
                diag.abi41_pcg_final_residual = diag.abi41_pcg_initial_residual;
            }
            const double fin = pcg_status == 2 ? pcg_rz : pcg_rznext;
            if ((pcg_status == 0 || pcg_status == 2) && std::isfinite(fin) && fin >= 0.0)
                diag.abi41_pcg_final_residual = static_cast<float>(std::sqrt(std::max(fin, 0.0)));
            else if (pcg_status != 0 && pcg_status != 2)
                diag.abi41_pcg_guarded += 1;
            diag.abi41_pcg_guarded += pcg_guard;
            diag.abi41_pcg_max_delta = pcg_max_delta;
        }
    }
    diag.abi41_bending_texture_ready = wing_count > 0 ? 1 : 0;
    populate_sdf_diag();
    diag.step_ms = static_cast<float>(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    diag.constraints_ms = diag.step_ms;
    diag.resolved_contacts = diag.abi41_soft_contact_count + diag.abi41_exact_impulse_contact_count +
                             diag.abi41_edge_edge_contact_count + diag.static_sdf_contact_count;
    return true;
}

// ----------------------------------------------------------- handle glue
struct Handle {
    int bits = 32;
    Solver<float>* f = nullptr;
    Solver<double>* d = nullptr;
    ~Handle() { delete f; delete d; }
};
template <class F> auto with(void* h, F fn) -> decltype(fn(*static_cast<Solver<float>*>(nullptr))) {
    Handle* H = static_cast<Handle*>(h);
    if (H->bits == 64) return fn(*H->d);
    return fn(*H->f);
}

template <class S> bool upload_pins(S& s, const int* idx, const float* p, const float* w, int count) {  // A41:6996-7049
    if (count < 0) return fail("invalid pin count");
    s.pins.clear();
    if (count == 0) return true;
    if (!idx || !p) return fail("pin arrays are required when pin count is nonzero");
    for (int k = 0; k < count; ++k) {
        Pin pin{idx[k], {p[3 * k], p[3 * k + 1], p[3 * k + 2]}, w ? w[k] : 1.0f};
        s.pins.push_back(pin);
    }
    return true;
}

template <class S> bool upload_fields(S& s, const SsblXpbdForceField* f, int count, int unsupported) {  // A41:5920-5964
    if (count < 0 || count > kMaxForceFields) return fail("force field count exceeds SSBL maximum");
    s.unsupported_fields = std::max(unsupported, 0);
    s.fields.clear();
    if (count == 0) return true;
    if (!f) return fail("missing force field data");
    s.fields.assign(f, f + count);
    return true;
}

template <class S> void copy_out(const S& s, const decltype(s.pos)& src, float* out) {
    for (int i = 0; i < s.cfg.vertex_count; ++i) {
        out[3 * i] = static_cast<float>(src[i].x);
        out[3 * i + 1] = static_cast<float>(src[i].y);
        out[3 * i + 2] = static_cast<float>(src[i].z);
    }
}

// A41:8338-8420 (soup) and 8575-8697 (particles) for the cloth's own colliders
template <class S> bool set_dynamic_triangles(S& s, const float* t, int count) {
    count = std::max(count, 0);
    if (count == 0) { s.dyn.clear_triangles(); return true; }
    if (!t) return fail("dynamic triangle data is required when triangle count is nonzero");
    s.dyn.set_triangles(t, count, nullptr, nullptr, s.cfg.collision_margin, s.cfg.cloth_thickness, false);
    s.dyn_tri_allocated = true;
    return true;
}
template <class S> bool set_dynamic_particles(S& s, const float* p, const float* r, int count) {
    count = std::max(count, 0);
    if (count == 0) { s.dyn.clear_particles(); return true; }
    if (!p || !r) return fail("dynamic particle position/radius data is required when particle count is nonzero");
    s.dyn.set_particles(p, nullptr, r, nullptr, nullptr, count, s.cfg.collision_margin, s.cfg.cloth_thickness);
    return true;
}

struct Batch { std::vector<void*> handles; };

}  // namespace

extern "C" {

// Test hook: 0 disables the dynamic collider hashes (brute-force queries over
// every triangle and particle, the reference's small-set path) for sets built
// afterwards.
SSBL_API int ssbl_oracle_dynamic_hash(int enable) {
    g_dynamic_hash = enable != 0;
    return 1;
}

SSBL_API int ssbl_oracle_set_precision(int bits) {
    if (bits != 32 && bits != 64) return fail("precision must be 32 or 64") ? 1 : 0;
    g_precision_bits = bits;
    return 1;
}

SSBL_API void* ssbl_create_solver(const SsblXpbdConfig* config, const SsblXpbdMesh* mesh) {
    g_error.clear();
    if (!config || config->vertex_count < 0 || config->edge_count < 0 || config->triangle_count < 0 || !std::isfinite(config->dt) ||
        config->dt <= 0.0f || !mesh || !mesh->positions || !mesh->inv_mass) {
        fail("invalid ABI41 solver create request");
        return nullptr;
    }
    Handle* h = new Handle();
    h->bits = g_precision_bits;
    bool ok;
    if (h->bits == 64) { h->d = new Solver<double>(); ok = h->d->create(config, mesh); }
    else { h->f = new Solver<float>(); ok = h->f->create(config, mesh); }
    if (!ok) { delete h; return nullptr; }
    return h;
}

SSBL_API int ssbl_destroy_solver(void* handle) { g_error.clear(); delete static_cast<Handle*>(handle); return 1; }

SSBL_API int ssbl_reset_solver(void* handle) {
    g_error.clear();
    if (!handle) return fail("invalid solver handle") ? 1 : 0;
    with(handle, [](auto& s) {
        s.pos = s.rest; s.prev = s.rest;
        for (auto& v : s.vel) v = {0, 0, 0};
        s.diag = SsblXpbdDiagnostics{}; s.diag.finite_flag = 1;
        return 0;
    });
    return 1;
}

SSBL_API int ssbl_update_pin_targets(void* handle, const int* indices, const float* positions, const float* weights, int count) {
    g_error.clear();
    if (!handle) return fail("invalid solver handle") ? 1 : 0;
    return with(handle, [&](auto& s) { return upload_pins(s, indices, positions, weights, count) ? 1 : 0; });
}

SSBL_API int ssbl_update_runtime_colliders(void* handle, const SsblXpbdRuntimeColliders* in) {
    g_error.clear();
    if (!handle || !in) return fail("invalid runtime collider update") ? 1 : 0;
    return with(handle, [&](auto& s) { s.colliders = *in; return 1; });
}

SSBL_API int ssbl_update_positions(void* handle, const float* p, int max_floats) {
    g_error.clear();
    if (!handle || !p) return fail("invalid position upload") ? 1 : 0;
    return with(handle, [&](auto& s) {
        if (max_floats < s.cfg.vertex_count * 3) return fail("invalid position upload") ? 1 : 0;
        for (int i = 0; i < s.cfg.vertex_count; ++i) {
            s.pos[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
            s.prev[i] = s.pos[i];
        }
        return 1;
    });
}

SSBL_API int ssbl_update_static_triangles(void* handle, const float* t, int count) {
    g_error.clear();
    if (!handle) return fail("invalid static triangle update") ? 1 : 0;
    count = std::max(count, 0);
    if (count > 0 && !t) return fail("static triangle data is required when triangle count is nonzero") ? 1 : 0;
    return with(handle, [&](auto& s) { return s.build_sdf(t, count) ? 1 : 0; });
}

SSBL_API int ssbl_update_dynamic_triangles(void* handle, const float* t, int count) {
    g_error.clear();
    if (!handle) return fail("invalid dynamic triangle update") ? 1 : 0;
    return with(handle, [&](auto& s) { return set_dynamic_triangles(s, t, count) ? 1 : 0; });
}

SSBL_API int ssbl_update_frame_inputs(void* handle, const SsblXpbdFrameInputs* in) {  // A41:9035-9085
    g_error.clear();
    if (!handle || !in) return fail("invalid frame input update") ? 1 : 0;
    return with(handle, [&](auto& s) {
        if (in->update_pin_targets && !upload_pins(s, in->pin_indices, in->pin_positions, in->pin_weights, in->pin_count)) return 0;
        if (in->update_runtime_colliders) s.colliders = in->runtime_colliders;
        if (in->update_dynamic_triangles && !set_dynamic_triangles(s, in->dynamic_triangles, in->dynamic_triangle_count)) return 0;
        if (in->update_dynamic_indexed_triangles) {   // A41:8433-8590
            const int nv = std::max(in->dynamic_triangle_vertex_count, 0), nt = std::max(in->dynamic_indexed_triangle_count, 0);
            if (nv > 0 && nt > 0 && (!in->dynamic_triangle_vertices || !in->dynamic_triangle_indices))
                return fail("dynamic indexed triangle positions and indices are required when count is nonzero") ? 1 : 0;
            s.dyn.set_indexed(in->dynamic_triangle_vertices, nv, in->dynamic_triangle_indices, nt, in->dynamic_triangle_topology_changed != 0,
                              s.cfg.collision_margin, s.cfg.cloth_thickness);
            if (nv > 0 && nt > 0) s.dyn_tri_allocated = true;
        }
        if (in->update_dynamic_particles &&
            !set_dynamic_particles(s, in->dynamic_particle_positions, in->dynamic_particle_radii, in->dynamic_particle_count))
            return 0;
        if (in->update_static_triangles) {
            int c = std::max(in->static_triangle_count, 0);
            if (c > 0 && !in->static_triangles) return fail("static triangle data is required when triangle count is nonzero") ? 1 : 0;
            if (!s.build_sdf(in->static_triangles, c)) return 0;
        }
        if (in->update_force_fields && !upload_fields(s, in->force_fields, in->force_field_count, in->unsupported_force_field_count)) return 0;
        return 1;
    });
}

SSBL_API int ssbl_step_solver_ex(void* handle, int substeps, int iterations, int fetch_diagnostics, int force_sync) {
    g_error.clear();
    (void)force_sync;
    if (!handle) return fail("invalid solver handle") ? 1 : 0;
    return with(handle, [&](auto& s) { return s.step(substeps, iterations, fetch_diagnostics) ? 1 : 0; });
}

SSBL_API int ssbl_step_solver(void* handle, int substeps, int iterations) { return ssbl_step_solver_ex(handle, substeps, iterations, 1, 1); }

// The allocation-free step variant of the GPU library; on the CPU it is the step itself.
SSBL_API int ssbl_ext_step_solver_ex_noalloc(void* handle, int substeps, int iterations, int fetch_diagnostics, int force_sync) {
    return ssbl_step_solver_ex(handle, substeps, iterations, fetch_diagnostics, force_sync);
}

SSBL_API int ssbl_download_positions(void* handle, float* out, int max_floats) {
    g_error.clear();
    if (!handle || !out) return fail("invalid position download") ? 1 : 0;
    return with(handle, [&](auto& s) {
        if (max_floats < s.cfg.vertex_count * 3) return fail("invalid position download") ? 1 : 0;
        copy_out(s, s.pos, out);
        return 1;
    });
}

SSBL_API int ssbl_get_diagnostics(void* handle, SsblXpbdDiagnostics* out) {
    g_error.clear();
    if (!handle || !out) return fail("invalid diagnostics request") ? 1 : 0;
    return with(handle, [&](auto& s) { *out = s.diag; return 1; });
}

SSBL_API unsigned int ssbl_capabilities(void) {
    return SSBL_CAP_STRETCH_OPTIMIZATION | SSBL_CAP_PIN_WEIGHTS | SSBL_CAP_GLOBAL_DYNAMIC_SCENE | SSBL_CAP_HIP_BATCHED_STEP;
}

SSBL_API void* ssbl_create_global_dynamic_scene(void) { g_error.clear(); return new Scene(); }
SSBL_API int ssbl_destroy_global_dynamic_scene(void* scene) {
    g_error.clear();
    if (!scene) return fail("invalid global dynamic scene handle") ? 1 : 0;
    delete static_cast<Scene*>(scene);
    return 1;
}
SSBL_API int ssbl_update_global_dynamic_scene(void* handle, const SsblXpbdGlobalDynamicSceneInputs* in) {  // A41:6597-6782
    g_error.clear();
    Scene* sc = static_cast<Scene*>(handle);
    if (!sc || !in) return fail("invalid global dynamic scene update") ? 1 : 0;
    const int np = std::max(in->particle_count, 0), nt = std::max(in->triangle_count, 0);
    if (np > 0 && (!in->particle_positions || !in->particle_radii || !in->particle_source_ids || !in->particle_collision_layers))
        return fail("global dynamic particle arrays are incomplete") ? 1 : 0;
    if (nt > 0 && (!in->triangle_vertices || !in->triangle_source_ids || !in->triangle_collision_layers))
        return fail("global dynamic triangle arrays are incomplete") ? 1 : 0;
    sc->margin = std::max(in->collision_margin, 0.0f);
    sc->thickness = std::max(in->cloth_thickness, 0.0f);
    std::memset(sc->counts, 0, sizeof(sc->counts));
    if (np == 0) sc->set.clear_particles();
    else
        sc->set.set_particles(in->particle_positions, in->particle_prev_positions ? in->particle_prev_positions : in->particle_positions,
                              in->particle_radii, in->particle_source_ids, in->particle_collision_layers, np, sc->margin, sc->thickness);
    if (nt == 0) sc->set.clear_triangles();
    else sc->set.set_triangles(in->triangle_vertices, nt, in->triangle_source_ids, in->triangle_collision_layers, sc->margin, sc->thickness, true);
    sc->set.build(sc->counts);
    sc->hash_overflow = static_cast<long long>(sc->counts[kCountDynParticleOverflow] + sc->counts[kCountDynTriBucketOverflow]);
    return 1;
}
SSBL_API int ssbl_attach_global_dynamic_scene(void* handle, void* scene, int target_source_id, int target_layer, int mode) {
    g_error.clear();
    if (!handle || !scene) return fail("invalid global dynamic scene attach") ? 1 : 0;
    return with(handle, [&](auto& s) {
        s.scene = static_cast<Scene*>(scene);
        s.target_src = target_source_id;
        s.target_layer = target_layer;
        s.cross_mode = mode;
        return 1;
    });
}
SSBL_API int ssbl_clear_global_dynamic_scene(void* handle) {
    g_error.clear();
    if (!handle) return fail("invalid global dynamic scene clear") ? 1 : 0;
    return with(handle, [&](auto& s) {
        s.scene = nullptr;
        s.target_src = 0;
        s.target_layer = 0;
        s.cross_mode = kCrossOff;
        return 1;
    });
}
SSBL_API const char* ssbl_last_error(void) { return g_error.c_str(); }

// ----------------------------------------------------------- extensions
SSBL_API int ssbl_ext_set_stretch_mode(void* handle, int mode) {
    g_error.clear();
    if (!handle || (mode != SSBL_EXT_STRETCH_JACOBI && mode != SSBL_EXT_STRETCH_COLORED)) return fail("invalid stretch mode") ? 1 : 0;
    return with(handle, [&](auto& s) {
        if (mode == SSBL_EXT_STRETCH_COLORED && !s.edge_coloring_valid)
            return fail("coloured stretch needs a valid edge colouring (edge color offsets, no vertex twice in one colour)") ? 1 : 0;
        s.stretch_mode = mode;
        return 1;
    });
}

SSBL_API int ssbl_ext_download_state(void* handle, float* p, float* q, float* v, int max_floats) {
    g_error.clear();
    if (!handle) return fail("invalid state download") ? 1 : 0;
    return with(handle, [&](auto& s) {
        if (max_floats < s.cfg.vertex_count * 3) return fail("invalid state download") ? 1 : 0;
        if (p) copy_out(s, s.pos, p);
        if (q) copy_out(s, s.prev, q);
        if (v) copy_out(s, s.vel, v);
        return 1;
    });
}

SSBL_API int ssbl_ext_self_hash_cells(void* handle, int* cells, int max_ints) {
    g_error.clear();
    if (!handle || !cells) return fail("invalid hash cell request") ? 1 : 0;
    return with(handle, [&](auto& s) {
        if (max_ints < 3 * s.cfg.vertex_count) return fail("invalid hash cell request") ? 1 : 0;
        std::copy(s.self_cells.begin(), s.self_cells.end(), cells);
        return 1;
    });
}

SSBL_API int ssbl_ext_self_neighbors(void* handle, unsigned* counts, unsigned* indices, int max_vertices) {
    g_error.clear();
    if (!handle || !counts || !indices) return fail("invalid neighbour request") ? 1 : 0;
    return with(handle, [&](auto& s) {
        if (max_vertices < s.cfg.vertex_count) return fail("invalid neighbour request") ? 1 : 0;
        std::copy(s.nb_counts.begin(), s.nb_counts.end(), counts);
        std::copy(s.nb_index.begin(), s.nb_index.end(), indices);
        return 1;
    });
}

// Oracle-only: the positions the last self-collision candidate table was
// built from (tests enumerate the same pairs by brute force).
SSBL_API int ssbl_oracle_self_snapshot(void* handle, float* p, float* q, int max_floats) {
    g_error.clear();
    if (!handle || !p || !q) return fail("invalid snapshot request") ? 1 : 0;
    return with(handle, [&](auto& s) {
        if (max_floats < 3 * s.cfg.vertex_count || s.snap_pos.empty()) return fail("no self-collision snapshot") ? 1 : 0;
        copy_out(s, s.snap_pos, p);
        copy_out(s, s.snap_prev, q);
        return 1;
    });
}

SSBL_API int ssbl_ext_static_sdf(void* handle, int* dims, float* ov, float* values, int max_values) {
    g_error.clear();
    if (!handle || !dims || !ov) return fail("invalid SDF request") ? 1 : 0;
    return with(handle, [&](auto& s) {
        if (!s.sdf_ready) return fail("no static SDF") ? 1 : 0;
        for (int k = 0; k < 3; ++k) dims[k] = s.sdf_dim[k];
        ov[0] = static_cast<float>(s.sdf_origin.x); ov[1] = static_cast<float>(s.sdf_origin.y);
        ov[2] = static_cast<float>(s.sdf_origin.z); ov[3] = static_cast<float>(s.sdf_voxel);
        if (values) {
            if (max_values < static_cast<int>(s.sdf.size())) return fail("SDF buffer too small") ? 1 : 0;
            for (size_t k = 0; k < s.sdf.size(); ++k) values[k] = static_cast<float>(s.sdf[k]);
        }
        return 1;
    });
}

SSBL_API void* ssbl_ext_create_batch(void* const* handles, int count) {
    g_error.clear();
    if (!handles || count <= 0) { fail("invalid batch"); return nullptr; }
    Batch* b = new Batch();
    b->handles.assign(handles, handles + count);
    return b;
}
SSBL_API int ssbl_ext_destroy_batch(void* batch) { delete static_cast<Batch*>(batch); return 1; }
SSBL_API int ssbl_ext_batch_set_members(void* batch, void* const* handles, int count) {
    g_error.clear();
    if (!batch || !handles || count <= 0) return fail("invalid batch member update") ? 1 : 0;
    static_cast<Batch*>(batch)->handles.assign(handles, handles + count);
    return 1;
}
SSBL_API int ssbl_ext_step_batch(void* batch, int substeps, int iterations, int force_sync) {
    g_error.clear();
    if (!batch) return fail("invalid batch") ? 1 : 0;
    for (void* h : static_cast<Batch*>(batch)->handles)
        if (!ssbl_step_solver_ex(h, substeps, iterations, 0, force_sync)) return 0;
    return 1;
}
SSBL_API long long ssbl_ext_global_scene_entries(void* scene) {
    g_error.clear();
    if (!scene) { fail("invalid global dynamic scene handle"); return -1; }
    const DynamicSet& d = static_cast<Scene*>(scene)->set;
    return d.tri_hash && !d.tri_off.empty() ? d.tri_off.back() : 0;
}
SSBL_API int ssbl_ext_step_batch_ex(void* batch, int substeps, int iterations, int fetch_diagnostics, int force_sync) {
    g_error.clear();
    if (!batch) return fail("invalid batch") ? 1 : 0;
    for (void* h : static_cast<Batch*>(batch)->handles)
        if (!ssbl_step_solver_ex(h, substeps, iterations, fetch_diagnostics, force_sync)) return 0;
    return 1;
}
SSBL_API int ssbl_ext_update_global_scene_from_handles(void* scene, void* const* handles, const int* source_ids, const int* layers,
                                                      const float* radii, int count, float margin, float thickness) {
    // the host packing of session_manager.py:3242-3300 from each handle's positions
    g_error.clear();
    if (!scene || count < 0 || (count > 0 && (!handles || !source_ids || !layers || !radii))) return fail("invalid global scene gather") ? 1 : 0;
    std::vector<float> pos, tv, pr;
    std::vector<int> psrc, play, tsrc, tlay;
    for (int k = 0; k < count; ++k) {
        if (!handles[k]) return fail("invalid solver handle in global scene gather") ? 1 : 0;
        with(handles[k], [&](auto& s) {
            const int n = s.cfg.vertex_count;
            for (int i = 0; i < n; ++i) {
                pos.push_back(static_cast<float>(s.pos[i].x));
                pos.push_back(static_cast<float>(s.pos[i].y));
                pos.push_back(static_cast<float>(s.pos[i].z));
                pr.push_back(radii[k]);
                psrc.push_back(source_ids[k]);
                play.push_back(layers[k]);
            }
            for (int t = 0; t < s.cfg.triangle_count; ++t) {
                const int ia = s.tris[3 * t], ib = s.tris[3 * t + 1], ic = s.tris[3 * t + 2];
                const bool ok = ia >= 0 && ib >= 0 && ic >= 0 && ia < n && ib < n && ic < n;
                for (int v : {ia, ib, ic}) {
                    tv.push_back(ok ? static_cast<float>(s.pos[v].x) : 0.0f);
                    tv.push_back(ok ? static_cast<float>(s.pos[v].y) : 0.0f);
                    tv.push_back(ok ? static_cast<float>(s.pos[v].z) : 0.0f);
                }
                tsrc.push_back(source_ids[k]);
                tlay.push_back(layers[k]);
            }
            return 0;
        });
    }
    SsblXpbdGlobalDynamicSceneInputs in{};
    in.particle_positions = pos.data();
    in.particle_prev_positions = nullptr;
    in.particle_radii = pr.data();
    in.particle_source_ids = psrc.data();
    in.particle_collision_layers = play.data();
    in.particle_count = static_cast<int>(pr.size());
    in.triangle_vertices = tv.data();
    in.triangle_source_ids = tsrc.data();
    in.triangle_collision_layers = tlay.data();
    in.triangle_count = static_cast<int>(tsrc.size());
    in.collision_margin = margin;
    in.cloth_thickness = thickness;
    return ssbl_update_global_dynamic_scene(scene, &in);
}
SSBL_API int ssbl_ext_set_static_sdf(void* handle, const int* dims, const float* ov, const float* values, int count) {
    g_error.clear();
    if (!handle) return fail("invalid solver handle") ? 1 : 0;
    return with(handle, [&](auto& s) { return s.set_sdf_grid(dims, ov, values, count) ? 1 : 0; });
}
SSBL_API long long ssbl_ext_device_bytes(void*) { fail("oracle: no device memory"); return -1; }
SSBL_API int ssbl_ext_time_batch_pass(void*, int, float*) { return fail("oracle: no device timing") ? 1 : 0; }
SSBL_API int ssbl_ext_time_batch_iteration(void*, int, float*) { return fail("oracle: no device timing") ? 1 : 0; }
SSBL_API int ssbl_ext_synchronize(void) { return 1; }
// The oracle launches nothing, so it has no graphs to replay or launches to tune.
SSBL_API int ssbl_ext_set_graph_mode(int) { return 0; }
SSBL_API int ssbl_ext_set_autotune(int) { return 0; }
SSBL_API int ssbl_ext_set_batch_parts(int) { return 1; }
SSBL_API int ssbl_ext_tuned_pool_limit(void*, void*) { return -1; }
SSBL_API int ssbl_ext_graph_stats(void*, void*, long long* captures, long long* replays) {
    if (!captures || !replays) return fail("invalid graph stats request") ? 1 : 0;
    *captures = 0;
    *replays = 0;
    return 1;
}
SSBL_API int ssbl_ext_kernel_timing(int) { return 1; }
SSBL_API int ssbl_ext_kernel_stats(char* out, int cap) {
    if (out && cap > 2) std::snprintf(out, static_cast<size_t>(cap), "{}");
    return 1;
}
SSBL_API const char* ssbl_ext_backend_name(void) { return "oracle cpu"; }

}  // extern "C"
