// xpbd_math.hpp -- float3 arithmetic for the gfx950 kernels.
//
// The operation order of every helper is fixed so results are reproducible:
// dot products are fused multiply-adds in x, y, z order (the reference's
// fma_rn helpers, ssbl_xpbd_cuda_abi41.cu:622-636), 1/sqrt is computed as a
// correctly rounded division by a correctly rounded square root, and the
// library is built with -ffp-contract=off so no other multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>

namespace ssbl {

constexpr float kEps = 1.0e-7f;

struct F3 { float x, y, z; };

__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 f3(float4 v) { return F3{v.x, v.y, v.z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 mul(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot(F3 a, F3 b) {
    return __builtin_fmaf(a.z, b.z, __builtin_fmaf(a.y, b.y, __builtin_fmaf(a.x, b.x, 0.0f)));
}
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return f3(__builtin_fmaf(a.y, b.z, -(a.z * b.y)), __builtin_fmaf(a.z, b.x, -(a.x * b.z)),
              __builtin_fmaf(a.x, b.y, -(a.y * b.x)));
}
__device__ __forceinline__ float len3(F3 a) { return sqrtf(fmaxf(dot(a, a), 0.0f)); }
__device__ __forceinline__ bool finite3(F3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
__device__ __forceinline__ float clamp01f(float v) { return fminf(fmaxf(v, 0.0f), 1.0f); }
__device__ __forceinline__ float rsqrt_exact(float v) { return 1.0f / sqrtf(v); }

// sqrtf(fmaxf(x, kEps)), bit for bit, in fewer instructions.  The correctly
// rounded root of a float y >= 2^-96 is the hardware root s (within 1 ulp)
// moved to a neighbour s' = s -+ 1 ulp when the residual fma(-s', s, y) says
// the root lies beyond the midpoint: the sequence the compiler emits for
// sqrtf, minus its rescaling of inputs below 2^-96 and its pass-through of
// +-0 and +inf, neither of which a value >= kEps needs (+inf comes out as
// +inf: both residuals are NaN).  NaN x gives the root of kEps, as fmaxf does.
__device__ __forceinline__ float sqrt_eps(float x) {
    const float y = fmaxf(x, kEps);
    const float s = __builtin_amdgcn_sqrtf(y);
    const float down = __int_as_float(__float_as_int(s) - 1), up = __int_as_float(__float_as_int(s) + 1);
    float r = __builtin_fmaf(-down, s, y) <= 0.0f ? down : s;
    r = __builtin_fmaf(-up, s, y) > 0.0f ? up : r;
    return r;
}

// sqrtf(x) > m for a float x >= 0 (or NaN or +inf) and a float m in
// [2^-60, 2^60], without taking the root.  sqrtf(x) > m exactly when sqrt(x)
// lies above the midpoint h between m and the next float; h has 25
// significant bits, so h and h*h (at most 50 bits) are exact in double, and
// sqrt(x) can never equal h (h*h needs more than the 24 bits of a float).
// Most x are far from h*h, which lies in (m^2, m^2 (1 + 2^-23 + 2^-48)]:
// with t = m*m rounded, t (1 + 2^-20) rounded is at least
// m^2 (1 + 7 * 2^-23) and t (1 - 2^-20) rounded is below m^2, so x above the
// first is greater, x at or below the second is not, and only x between them
// (NaN included) takes the comparison in double.  That branch is taken by the
// wave as a whole, only when one of its lanes needs it.
__device__ __forceinline__ bool sqrt_greater(float x, float m) {
    const float t = m * m;
    const bool above = x > t * 1.00000095367431640625f;
    const bool band = !above && !(x <= t * 0.99999904632568359375f);
    bool r = above;
    if (__ballot(band) != 0ull) {
        const double h = (static_cast<double>(m) + static_cast<double>(__int_as_float(__float_as_int(m) + 1))) * 0.5;
        if (band) r = static_cast<double>(x) > h * h;
    }
    return r;
}
__device__ __forceinline__ float4 with_w(F3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// Loads and stores through the global address space.  The kernels take
// their array pointers from a table in device memory, so the compiler sees
// generic pointers and emits flat accesses, which also count against the
// scalar-load counter and make every scalar wait drain the vector loads in
// flight; these helpers emit global_load / global_store instead.
#define SSBL_GLOBAL __attribute__((address_space(1)))
// A table that no kernel writes (a collider set's view), read through the
// constant address space: its uniform fields come in with scalar loads.
#define SSBL_CONST __attribute__((address_space(4)))
typedef float gv4f __attribute__((ext_vector_type(4)));
typedef int gv2i __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float gload(const float* p) { return *(const SSBL_GLOBAL float*)(p); }
__device__ __forceinline__ int2 gload(const int2* p) {
    const gv2i x = *(const SSBL_GLOBAL gv2i*)(p);
    return make_int2(x.x, x.y);
}
__device__ __forceinline__ float4 gload(const float4* p) {
    const gv4f x = *(const SSBL_GLOBAL gv4f*)(p);
    return make_float4(x.x, x.y, x.z, x.w);
}
typedef float gv3f __attribute__((ext_vector_type(3)));
// x, y, z of a float4 array element with one 12-byte load: a 16-byte load
// whose .w is never read leaves a register the compiler reuses, and it then
// waits for the load to land before the reuse.
__device__ __forceinline__ F3 gload_xyz(const float4* p) {
    const gv3f x = *(const SSBL_GLOBAL gv3f*)(p);
    return f3(x.x, x.y, x.z);
}
typedef int gv4i __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int gload(const int* p) { return *(const SSBL_GLOBAL int*)(p); }
__device__ __forceinline__ short gload(const short* p) { return *(const SSBL_GLOBAL short*)(p); }
__device__ __forceinline__ int4 gload(const int4* p) {
    const gv4i x = *(const SSBL_GLOBAL gv4i*)(p);
    return make_int4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ unsigned long long gload(const unsigned long long* p) { return *(const SSBL_GLOBAL unsigned long long*)(p); }
__device__ __forceinline__ void gstore(unsigned long long* p, unsigned long long v) { *(SSBL_GLOBAL unsigned long long*)(p) = v; }
__device__ __forceinline__ void gstore(float4* p, float4 v) {
    gv4f x;
    x.x = v.x; x.y = v.y; x.z = v.z; x.w = v.w;
    *(SSBL_GLOBAL gv4f*)(p) = x;
}

__device__ __forceinline__ F3 limit_delta(F3 d, float max_len) {
    if (!finite3(d) || !isfinite(max_len) || max_len <= 0.0f) return f3(0.0f, 0.0f, 0.0f);
    const float len = len3(d);
    if (!isfinite(len) || len <= kEps) return f3(0.0f, 0.0f, 0.0f);
    return len > max_len ? mul(d, max_len / len) : d;
}

__device__ __forceinline__ F3 normalize_or(F3 v, F3 fb) {
    const float len = len3(v);
    if (!isfinite(len) || len <= kEps) return fb;
    return mul(v, 1.0f / len);
}

// Closest point on triangle abc to p (Ericson, Real-Time Collision Detection
// 5.1.5; the region order of ssbl_xpbd_cuda_abi41.cu:1589-1635).
__device__ inline F3 closest_point_on_triangle(F3 p, F3 a, F3 b, F3 c) {
    const F3 ab = b - a, ac = c - a, ap = p - a;
    const float d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    const F3 bp = p - b;
    const float d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return a + mul(ab, d1 / fmaxf(d1 - d3, kEps));
    const F3 cp = p - c;
    const float d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return a + mul(ac, d2 / fmaxf(d2 - d6, kEps));
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float w = (d4 - d3) / fmaxf((d4 - d3) + (d5 - d6), kEps);
        return b + mul(c - b, w);
    }
    const float denom = 1.0f / fmaxf(va + vb + vc, kEps);
    const float v = vb * denom, w = vc * denom;
    return a + (mul(ab, v) + mul(ac, w));
}

__device__ __forceinline__ int cell_coord(float value, float cell) {
    return static_cast<int>(floorf(value / fmaxf(cell, kEps)));
}

// Spatial hash of a cell (ssbl_xpbd_cuda_abi41.cu:658-664), bit-exact.
__device__ __forceinline__ unsigned hash_cell(int x, int y, int z, int buckets) {
    const unsigned h = static_cast<unsigned>(x) * 73856093u ^ static_cast<unsigned>(y) * 19349663u ^
                       static_cast<unsigned>(z) * 83492791u;
    return h & static_cast<unsigned>(buckets - 1);
}

// Wave64-aggregated counter increment: one atomic per wavefront.
__device__ __forceinline__ void wave_count(unsigned long long* counts, int slot, bool cond) {
    const unsigned long long active = __ballot(1);
    const unsigned long long hits = __ballot(cond);
    const int leader = __ffsll(static_cast<long long>(active)) - 1;
    if (static_cast<int>(threadIdx.x & 63u) == leader && hits != 0ull) {
        atomicAdd(&counts[slot], static_cast<unsigned long long>(__popcll(hits)));
    }
}

// Slot of bucket b in a block's LDS table of kSlots slots (a power of two;
// linear probing from b's low bits), inserting b and appending the slot to
// s_list when it is new; -1 when kProbes probes find no slot.  Slots are
// never freed, so a bucket inserted once is found again at the same place,
// and one that found no slot never finds one later in the launch.  s_bucket
// starts filled with -1 and *s_n with 0.
template <int kSlots, int kProbes>
__device__ __forceinline__ int lds_bucket_slot(int* s_bucket, int* s_list, int* s_n, int b) {
    int h = b & (kSlots - 1);
    int old = atomicCAS(&s_bucket[h], -1, b);
    for (int k = 1; old != -1 && old != b; ++k) {
        if (k == kProbes) return -1;
        h = (h + 1) & (kSlots - 1);
        old = atomicCAS(&s_bucket[h], -1, b);
    }
    if (old == -1) s_list[atomicAdd(s_n, 1)] = h;
    return h;
}

// Orders a wave's LDS writes before its later LDS reads of other lanes'
// slots (a wave's LDS operations complete in issue order; this keeps the
// compiler from moving them across, and waits for those in flight).
// s_waitcnt operand of gfx9: lgkmcnt (bits 11:8) 0, vmcnt (bits 3:0, 15:14)
// and expcnt (bits 6:4) at their maxima, i.e. left alone.
constexpr unsigned kWaitLgkmZero = 0xc07fu;
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_s_waitcnt(kWaitLgkmZero);
    __builtin_amdgcn_wave_barrier();
}

// Wave64 sum of a per-lane count added to counts[slot] by one lane.  Every
// lane of the wave must call it.
__device__ __forceinline__ void wave_add(unsigned long long* counts, int slot, unsigned v) {
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63u) == 0u && v != 0u) atomicAdd(&counts[slot], static_cast<unsigned long long>(v));
}

// Wave64 maximum of non-negative floats into *dst (compared as their bit
// patterns, which order non-negative floats), one atomic per wavefront.
// Every lane of the wave must call it.
__device__ __forceinline__ void wave_max_nonneg(float* dst, float v) {
    int m = __float_as_int(fmaxf(v, 0.0f));
    for (int d = 32; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63u) == 0u && m > 0) atomicMax(reinterpret_cast<int*>(dst), m);
}

}  // namespace ssbl
