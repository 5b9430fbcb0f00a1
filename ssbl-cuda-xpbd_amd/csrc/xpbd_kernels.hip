// xpbd_kernels.hip -- per-vertex and per-colour kernels of the XPBD substep
// for gfx950 (wave64, 256-thread workgroups, one 256-vertex tile per group).
//
// Reference path (native/src/ssbl_xpbd_cuda_abi41.cu, cited A41:<line>):
// integrate A41:1869-1925, pins A41:3466-3509, direct stretch A41:1927-1980,
// bend A41:2886-2970, LRA A41:2710-2803, analytic contacts A41:3511-3590 with
// the response of A41:910-949, static SDF contact A41:1831-1867, velocity
// update A41:5846-5884, SDF build A41:1787-1829.
//
// The reference runs each of these as its own launch over all vertices (or
// edges, with float atomics).  Here every vertex-local step of one solver
// iteration is fused into a single pass: the stretch projection is gathered per
// vertex from an incidence list (a Jacobi sweep: neighbours are read from the
// ping-pong source buffer), then pins, analytic and SDF contacts, the second
// pin projection and, on the last iteration, the velocity update run on the
// vertex's registers before one coalesced write.  A batch of cloths is a table
// of ClothView + a tile table; every launch covers all tiles of all cloths.
#include "xpbd_device.hpp"
#include "xpbd_dynamic.hpp"
#include "xpbd_math.hpp"
#include "xpbd_timing.hpp"

namespace ssbl {

// Tile index for this workgroup: consecutive tiles go to the same XCD (blocks
// are dealt round-robin over the 8 XCDs), so a cloth's neighbouring tiles
// share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int ntiles) {
    const int b = static_cast<int>(blockIdx.x);
    const int xcd = b & 7, local = b >> 3;
    const int per = ntiles >> 3, extra = ntiles & 7;
    return xcd * per + min(xcd, extra) + local;
}

// ------------------------------------------------------------ contacts
// A41:910-949 velocity response of a static contact.
__device__ inline void contact_response(const ClothView& c, F3 n, float proxy, F3& v, bool count) {
    if (!finite3(n) || !finite3(v)) return;
    const float vn = dot(v, n);
    if (vn < 0.0f) v = v - mul(n, vn);
    bool corrected = false;
    F3 tangent = v - mul(n, dot(v, n));
    float tl = len3(tangent);
    const float damping = clamp01f(c.tangent_damping);
    if (damping > 0.0f && tl > 1.0e-6f) { v = v - mul(tangent, damping); corrected = true; }
    tangent = v - mul(n, dot(v, n));
    tl = len3(tangent);
    const float friction = fmaxf(c.friction, 0.0f);
    if (friction > 0.0f && proxy > 0.0f && tl > 1.0e-6f) {
        const float drop = fminf(tl, friction * proxy);
        if (drop > 1.0e-7f) { v = v - mul(tangent, drop / fmaxf(tl, kEps)); corrected = true; }
    }
    if (count && corrected) atomicAdd(&c.counts[kCountFriction], 1ull);
}

// Registers of one vertex during a pass.  The previous position and the
// velocity are read from HBM only when a step of the pass needs them and
// written back only when a step changed them: a constraint iteration with no
// pin or contact on a vertex moves neither array.
struct VertexState {
    const ClothView* c;
    int v;
    F3 p;
    F3 q_, vel_;
    bool q_have, q_dirty, v_have, v_dirty;
    __device__ F3 q() {
        if (!q_have) { q_ = gload_xyz(c->prev + v); q_have = true; }
        return q_;
    }
    __device__ void set_q(F3 x) { q_ = x; q_have = true; q_dirty = true; }
    __device__ F3 vel() {
        if (!v_have) { vel_ = gload_xyz(c->vel + v); v_have = true; }
        return vel_;
    }
    __device__ void set_vel(F3 x) { vel_ = x; v_have = true; v_dirty = true; }
};

// A41:3466-3509: every pin of vertex v in ascending pin order, records
// b .. e of the vertex-ordered pin records (pin_off[v], pin_off[v + 1]);
// `first` is record b, already loaded.
__device__ inline void apply_pins(const ClothView& c, VertexState& s, bool pinned, float dt, int b, int e, float4 first) {
    for (int k = b; k < e; ++k) {
        const float4 pin = k == b ? first : gload(c.pins + k);
        const float relax = pin.w;
        if (relax < 0.0f) continue;                       // weight < 0.05 or not finite
        const F3 target = f3(pin.x, pin.y, pin.z);
        if (relax >= 2.0f || pinned) {                    // hard pin
            s.p = target;
            s.set_q(target);
            s.set_vel(f3(0.0f, 0.0f, 0.0f));
            continue;
        }
        if (relax <= 0.0f) continue;
        const F3 corrected = s.p + mul(target - s.p, relax);
        const F3 damped = mul(s.vel(), fmaxf(1.0f - relax, 0.0f));
        s.p = corrected;
        s.set_q(corrected - mul(damped, fmaxf(dt, kEps)));
        s.set_vel(damped);
    }
}
__device__ inline void apply_pins(const ClothView& c, VertexState& s, bool pinned, float dt) {
    const int b = c.pin_off[s.v], e = c.pin_off[s.v + 1];
    apply_pins(c, s, pinned, dt, b, e, b < e ? gload(c.pins + b) : make_float4(0.0f, 0.0f, 0.0f, -1.0f));
}

// A vertex's pins as a pass applies them: the range, requested with the
// vertex's other loads at the start of a pass whose `ops` apply pins
// ({0, 0} otherwise).  The contact pass also requests the first record with
// them (request_first); the vertex passes load it where it is applied (held
// through their gather and contact work, it would not fit their register
// budget).
struct VertexPins {
    int b, e;
    float4 first;
    bool have_first;
    __device__ __forceinline__ void range(const ClothView& c, int v, unsigned ops) {
        have_first = false;
        if (!(ops & (kOpPin | kOpPinPolish | kOpIntegratePin))) { b = e = 0; return; }
        const int* off = c.pin_off;
        b = gload(off + v);
        e = gload(off + v + 1);
    }
    __device__ __forceinline__ void request_first(const ClothView& c) {
        first = b < e ? gload(c.pins + b) : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        have_first = true;
    }
    __device__ __forceinline__ void apply(const ClothView& c, VertexState& s, bool pinned, float dt) const {
        if (b >= e) return;
        apply_pins(c, s, pinned, dt, b, e, have_first ? first : gload(c.pins + b));
    }
    // The pin projection after a pass's contact work, for a vertex whose
    // range was not empty (`had`): the range is read again rather than held
    // through the contact work, whose registers it would need.  After a
    // first application in the same pass (`again`) a vertex the reference
    // flags as pinned is skipped: every pin of such a vertex sets position,
    // previous position and velocity from its record alone, and the contact
    // steps between the two applications skip pinned vertices, so the
    // second application would write what the first wrote.
    __device__ __forceinline__ void apply_after_contacts(const ClothView& c, VertexState& s, bool pinned, float dt, bool had,
                                                         bool again) const {
        if (!had || (again && pinned)) return;
        int v = s.v;
        asm volatile("" : "+v"(v));   // a load made again, not the first one kept
        const int rb = gload(c.pin_off + v), re = gload(c.pin_off + v + 1);
        apply_pins(c, s, pinned, dt, rb, re, gload(c.pins + rb));
    }
};

// A41:3511-3590
__device__ inline void analytic_contacts(const ClothView& c, VertexState& s, float dt, bool count) {
    const float margin = fmaxf(c.collision_margin, 0.0f);
    const float sdt = fmaxf(dt, kEps);
    const SsblXpbdRuntimeColliders& col = c.colliders;
    F3 p = s.p;
    F3 v = f3(0.0f, 0.0f, 0.0f);
    bool hit = false;
    if (col.use_ground && p.z < col.ground_height + margin) {
        v = s.vel();
        const float tz = col.ground_height + margin;
        const float pen = tz - p.z;
        const F3 n = f3(0.0f, 0.0f, 1.0f);
        const float inward = fmaxf(-dot(v, n), 0.0f);
        p.z = tz;
        contact_response(c, n, inward + pen / sdt, v, count);
        hit = true;
    }
    if (col.use_sphere && col.sphere_radius > 0.0f) {
        const F3 center = f3(col.sphere_center[0], col.sphere_center[1], col.sphere_center[2]);
        const F3 d = p - center;
        const float dist = sqrt_eps(dot(d, d));
        const float target = col.sphere_radius + margin;
        if (dist < target) {
            if (!hit) v = s.vel();
            const F3 n = dist > kEps ? mul(d, 1.0f / dist) : f3(0.0f, 0.0f, 1.0f);
            const float pen = target - dist;
            const float inward = fmaxf(-dot(v, n), 0.0f);
            p = center + mul(n, target);
            contact_response(c, n, inward + pen / sdt, v, count);
            hit = true;
        }
    }
    if (col.use_wall) {
        const F3 o = f3(col.wall_origin[0], col.wall_origin[1], col.wall_origin[2]);
        F3 n = f3(col.wall_normal[0], col.wall_normal[1], col.wall_normal[2]);
        const float nl = sqrt_eps(dot(n, n));
        n = mul(n, 1.0f / nl);
        const float gap = dot(p - o, n);
        if (gap < margin) {
            if (!hit) v = s.vel();
            const float pen = margin - gap;
            const float inward = fmaxf(-dot(v, n), 0.0f);
            p = p + mul(n, margin - gap);
            contact_response(c, n, inward + pen / sdt, v, count);
            hit = true;
        }
    }
    if (hit) {
        s.p = p;
        s.set_vel(v);
        s.set_q(p - mul(v, sdt));
    }
}

// One trilinear sample (A41:1645-1691, A41:1693-1723) split in two halves:
// corners() issues the eight loads of the cell containing grid coordinates
// (gx, gy, gz) inside the grid and keeps the weights; blend() interpolates
// once they arrive.  Callers with several samples issue all their loads
// first.  Element offsets are 32-bit: a grid holds at most INT_MAX voxels.
struct SdfCell {
    float k[8];
    float tx, ty, tz;
    __device__ __forceinline__ void corners(const ClothView& c, float gx, float gy, float gz) {
        const int dx = c.sdf_dim[0], dy = c.sdf_dim[1], dz = c.sdf_dim[2];
        const int x0 = min(max(static_cast<int>(floorf(gx)), 0), dx - 1), y0 = min(max(static_cast<int>(floorf(gy)), 0), dy - 1),
                  z0 = min(max(static_cast<int>(floorf(gz)), 0), dz - 1);
        const int x1 = min(x0 + 1, dx - 1), y1 = min(y0 + 1, dy - 1), z1 = min(z0 + 1, dz - 1);
        tx = gx - static_cast<float>(x0);
        ty = gy - static_cast<float>(y0);
        tz = gz - static_cast<float>(z0);
        const unsigned r00 = static_cast<unsigned>((z0 * dy + y0) * dx), r10 = static_cast<unsigned>((z0 * dy + y1) * dx),
                       r01 = static_cast<unsigned>((z1 * dy + y0) * dx), r11 = static_cast<unsigned>((z1 * dy + y1) * dx);
        const float* sdf = c.sdf;
        k[0] = gload(sdf + (r00 + x0)); k[1] = gload(sdf + (r00 + x1));
        k[2] = gload(sdf + (r10 + x0)); k[3] = gload(sdf + (r10 + x1));
        k[4] = gload(sdf + (r01 + x0)); k[5] = gload(sdf + (r01 + x1));
        k[6] = gload(sdf + (r11 + x0)); k[7] = gload(sdf + (r11 + x1));
    }
    __device__ __forceinline__ float blend() const {
        auto lerp = [](float a, float b, float t) { return a + (b - a) * t; };
        const float c00 = lerp(k[0], k[1], tx), c10 = lerp(k[2], k[3], tx);
        const float c01 = lerp(k[4], k[5], tx), c11 = lerp(k[6], k[7], tx);
        return lerp(lerp(c00, c10, ty), lerp(c01, c11, ty), tz);
    }
};

// A41:1645-1691 trilinear sample; false outside the grid.
__device__ inline bool sdf_sample(const ClothView& c, F3 p, float* out) {
    if (c.sdf_dim[0] < 2 || c.sdf_dim[1] < 2 || c.sdf_dim[2] < 2 || c.sdf_voxel <= 0.0f) return false;
    const float inv = c.sdf_inv;
    const float gx = (p.x - c.sdf_origin[0]) * inv, gy = (p.y - c.sdf_origin[1]) * inv, gz = (p.z - c.sdf_origin[2]) * inv;
    if (gx < 0.0f || gy < 0.0f || gz < 0.0f || gx > c.sdf_top[0] || gy > c.sdf_top[1] || gz > c.sdf_top[2]) return false;
    SdfCell cell;
    cell.corners(c, gx, gy, gz);
    *out = cell.blend();
    return isfinite(*out);
}

// A41:1693-1723, the loads only (blend() gives the sample).
__device__ __forceinline__ void sdf_clamped_corners(const ClothView& c, F3 p, SdfCell& cell) {
    const float inv = c.sdf_inv_clamped;
    const float gx = fminf(fmaxf((p.x - c.sdf_origin[0]) * inv, 0.0f), c.sdf_top[0]);
    const float gy = fminf(fmaxf((p.y - c.sdf_origin[1]) * inv, 0.0f), c.sdf_top[1]);
    const float gz = fminf(fmaxf((p.z - c.sdf_origin[2]) * inv, 0.0f), c.sdf_top[2]);
    cell.corners(c, gx, gy, gz);
}

// The corner loads of the clamped sample at p (A41:1693-1723).
__device__ __forceinline__ void sdf_clamped_loads(const ClothView& c, F3 p, float (&k)[8]) {
    SdfCell cell;
    sdf_clamped_corners(c, p, cell);
#pragma unroll
    for (int j = 0; j < 8; ++j) k[j] = cell.k[j];
}

// The clamped sample at p from its corners k: the cell and weights are
// computed again, exactly as for the loads.
__device__ __forceinline__ float sdf_clamped_blend(const ClothView& c, F3 p, const float (&k)[8]) {
    const float inv = c.sdf_inv_clamped;
    const float gx = fminf(fmaxf((p.x - c.sdf_origin[0]) * inv, 0.0f), c.sdf_top[0]);
    const float gy = fminf(fmaxf((p.y - c.sdf_origin[1]) * inv, 0.0f), c.sdf_top[1]);
    const float gz = fminf(fmaxf((p.z - c.sdf_origin[2]) * inv, 0.0f), c.sdf_top[2]);
    SdfCell cell;
    cell.tx = gx - static_cast<float>(min(max(static_cast<int>(floorf(gx)), 0), c.sdf_dim[0] - 1));
    cell.ty = gy - static_cast<float>(min(max(static_cast<int>(floorf(gy)), 0), c.sdf_dim[1] - 1));
    cell.tz = gz - static_cast<float>(min(max(static_cast<int>(floorf(gz)), 0), c.sdf_dim[2] - 1));
#pragma unroll
    for (int j = 0; j < 8; ++j) cell.k[j] = k[j];
    return cell.blend();
}

// Hides x's value from the compiler, so that arithmetic on it is done again
// rather than kept in registers from an earlier identical computation.
__device__ __forceinline__ F3 opaque(F3 x) {
    asm volatile("" : "+v"(x.x), "+v"(x.y), "+v"(x.z));
    return x;
}

// A41:1771-1785, falling back to the nearest static triangle (A41:1749-1769).
// The six clamped samples of the central differences load in two rounds:
// the x and y pairs' 32 corners together, then the z pair's with the
// x, y, z of the caller's `extra` element (returned in *extra_out).  The
// weights are computed
// after the loads return, so that only the corner values are held while
// they are in flight.
__device__ inline F3 sdf_normal(const ClothView& c, F3 p, const float4* extra, F3* extra_out) {
    const float h = fmaxf(c.sdf_voxel, 1.0e-5f);
    float kxh[8], kxl[8], kyh[8], kyl[8];
    sdf_clamped_loads(c, p + f3(h, 0.0f, 0.0f), kxh);
    sdf_clamped_loads(c, p + f3(-h, 0.0f, 0.0f), kxl);
    sdf_clamped_loads(c, p + f3(0.0f, h, 0.0f), kyh);
    sdf_clamped_loads(c, p + f3(0.0f, -h, 0.0f), kyl);
    __builtin_amdgcn_sched_barrier(0);
    const F3 pa = opaque(p);
    const float dx = sdf_clamped_blend(c, pa + f3(h, 0.0f, 0.0f), kxh) - sdf_clamped_blend(c, pa + f3(-h, 0.0f, 0.0f), kxl);
    const float dy = sdf_clamped_blend(c, pa + f3(0.0f, h, 0.0f), kyh) - sdf_clamped_blend(c, pa + f3(0.0f, -h, 0.0f), kyl);
    __builtin_amdgcn_sched_barrier(0);
    float kzh[8], kzl[8];
    sdf_clamped_loads(c, pa + f3(0.0f, 0.0f, h), kzh);
    sdf_clamped_loads(c, pa + f3(0.0f, 0.0f, -h), kzl);
    *extra_out = gload_xyz(extra);
    __builtin_amdgcn_sched_barrier(0);
    const F3 pb = opaque(pa);
    const float dz = sdf_clamped_blend(c, pb + f3(0.0f, 0.0f, h), kzh) - sdf_clamped_blend(c, pb + f3(0.0f, 0.0f, -h), kzl);
    const F3 n = f3(dx, dy, dz);
    const float lsq = dot(n, n);
    if (isfinite(lsq) && lsq > 1.0e-12f) return mul(n, rsqrt_exact(lsq));
    F3 best_n = f3(0.0f, 0.0f, 1.0f);
    float best = 3.402823466e+38f;
    for (int t = 0; t < c.nstatic; ++t) {
        const float* T = c.static_tris + 9 * t;
        const F3 a = f3(T[0], T[1], T[2]), b = f3(T[3], T[4], T[5]), cc = f3(T[6], T[7], T[8]);
        const F3 d = p - closest_point_on_triangle(p, a, b, cc);
        const float dsq = dot(d, d);
        if (dsq < best) {
            best = dsq;
            best_n = dsq > kEps * kEps ? normalize_or(d, best_n) : normalize_or(cross(b - a, cc - a), best_n);
        }
    }
    return normalize_or(best_n, f3(0.0f, 0.0f, 1.0f));
}

// A41:1831-1867 (the caller skips pinned vertices).
__device__ inline void sdf_contact(const ClothView& c, VertexState& s, float dt, bool count) {
    const float margin = fmaxf(c.collision_margin, 0.0f);
    if (margin <= 0.0f) return;
    float d = 0.0f;
    if (!sdf_sample(c, s.p, &d) || d >= margin) return;
    // (the depth is formed before the normal, and kept from being moved
    // after it: one register held through its loads instead of two)
    float pen = margin - d;
    asm volatile("" : "+v"(pen));
    // the velocity load goes out with the normal's last round
    F3 vraw;
    const F3 n = sdf_normal(c, s.p, c.vel + s.v, &vraw);
    F3 v = s.v_have ? s.vel_ : vraw;
    if (!finite3(n)) return;
    if (!isfinite(pen) || pen <= 0.0f) return;
    const F3 corrected = s.p + mul(n, pen);
    const float inward = fmaxf(-dot(v, n), 0.0f);
    contact_response(c, n, inward + pen / fmaxf(dt, kEps), v, count);
    s.p = corrected;
    s.set_vel(v);
    s.set_q(corrected - mul(v, fmaxf(dt, kEps)));
    if (count) atomicAdd(&c.counts[kCountStaticSdf], 1ull);
}

// Dynamic colliders, unsigned contacts: the deepest triangle contact
// (A41:3999-4323) then the deepest particle contact (A41:4354-4511) push the
// vertex by half their depth (A41:129).  Only the vertex's own position is
// read and written, so running both here equals the reference's two
// vertex-parallel launches.
__device__ inline bool dyn_enabled(const ClothView& c, int src, int layer) {   // A41:3984-3997
    if (c.dyn_mode == 0) return true;
    if (c.dyn_target_src != 0 && src == c.dyn_target_src) return false;
    if (c.dyn_mode == 2) return layer < c.dyn_target_layer;
    return true;
}

// True when the filter rejects every entry a bucket bin covers
// ({offset, count, min layer, the one source of all entries | -1 | -2 empty}).
__device__ __forceinline__ bool dyn_bucket_rejected(const ClothView& c, int4 bin) {
    if (bin.w == -2) return true;
    if (c.dyn_mode == 0) return false;
    if (c.dyn_target_src != 0 && bin.w == c.dyn_target_src) return true;
    return c.dyn_mode == 2 && bin.z >= c.dyn_target_layer;
}

// One dynamic triangle against point p (A41:3999-4323 per candidate): counts
// the candidate, the AABB and bounding-sphere rejections, and returns the
// contact depth (0 when out of reach) with the push along the contact normal.
template <typename Corners>
__device__ __forceinline__ float dyn_triangle_depth_with(const TriBox& bx, Corners corners, F3 p, float cm, unsigned& cand, unsigned& rej,
                                                         F3& delta) {
    ++cand;
    const float4 mn = box_lo(bx), mx = box_hi(bx);
    const F3 q0 = f3(fminf(fmaxf(p.x, mn.x), mx.x), fminf(fmaxf(p.y, mn.y), mx.y), fminf(fmaxf(p.z, mn.z), mx.z));
    const F3 da = p - q0;
    if (dot(da, da) > cm * cm) { ++rej; return 0.0f; }
    const float4 ce = bx.s;
    const float rr = ce.w + cm;
    const F3 dc = p - f3(ce);
    if (dot(dc, dc) > rr * rr) { ++rej; return 0.0f; }
    F3 a, b, cc;
    corners(a, b, cc);
    const F3 q = closest_point_on_triangle(p, a, b, cc);
    const F3 d = p - q;
    const float dsq = dot(d, d);
    const float dist = sqrt_eps(dsq);
    const F3 tn = cross(b - a, cc - a);
    const float tl = sqrt_eps(dot(tn, tn));
    const F3 tnorm = mul(tn, 1.0f / tl);
    const F3 n = dsq > kEps * 4.0f ? mul(d, 1.0f / fmaxf(dist, kEps)) : tnorm;
    if (dist >= cm) return 0.0f;
    const float pen = cm - dist;
    delta = mul(n, pen);
    return pen;
}

// The corners are read only once the box and sphere tests pass.
__device__ __forceinline__ float dyn_triangle_depth(const SSBL_CONST DynView& D, int t, const TriBox& bx, F3 p, float cm, unsigned& cand,
                                                    unsigned& rej, F3& delta) {
    return dyn_triangle_depth_with(
        bx,
        [&](F3& a, F3& b, F3& c) {
            a = f3(gload(D.tv + 3 * t));
            b = f3(gload(D.tv + 3 * t + 1));
            c = f3(gload(D.tv + 3 * t + 2));
        },
        p, cm, cand, rej, delta);
}

// One dynamic particle (position and radius pj) against point p
// (A41:4354-4511 per candidate).
__device__ __forceinline__ float dyn_particle_depth_at(float4 pj, F3 p, float margin, unsigned& cand, F3& delta) {
    const float radius = fmaxf(pj.w, 0.0f) + margin;
    if (radius <= 0.0f) return 0.0f;
    const F3 d = p - f3(pj);
    const float dsq = dot(d, d);
    const float dist = sqrt_eps(dsq);
    if (dist >= radius) return 0.0f;
    ++cand;
    const F3 n = dist > kEps ? mul(d, 1.0f / dist) : f3(0.0f, 0.0f, 1.0f);
    const float pen = radius - dist;
    delta = mul(n, pen);
    return pen;
}

__device__ __forceinline__ float dyn_particle_depth(const SSBL_CONST DynView& D, int j, F3 p, float margin, unsigned& cand, F3& delta) {
    return dyn_particle_depth_at(gload(D.pp + j), p, margin, cand, delta);
}

// False when no point of neighbour cell (cx+dx, cy+dy, cz+dz) of p's cell
// lies within `reach` of p (with slack for the rounding of the cell bounds
// and of the distance itself): particles there can neither touch p nor count
// as candidates, so the search may pass the cell by.
__device__ __forceinline__ bool cell_in_reach(F3 p, int cx, int cy, int cz, int dx, int dy, int dz, float cs, float reach) {
    const float gx = dx < 0 ? p.x - static_cast<float>(cx) * cs : (dx > 0 ? static_cast<float>(cx + 1) * cs - p.x : 0.0f);
    const float gy = dy < 0 ? p.y - static_cast<float>(cy) * cs : (dy > 0 ? static_cast<float>(cy + 1) * cs - p.y : 0.0f);
    const float gz = dz < 0 ? p.z - static_cast<float>(cz) * cs : (dz > 0 ? static_cast<float>(cz + 1) * cs - p.z : 0.0f);
    const float r = reach * 1.001f + (fabsf(p.x) + fabsf(p.y) + fabsf(p.z)) * 1.0e-6f + 1.0e-7f;
    return !(gx * gx + gy * gy + gz * gz >= r * r);
}

// The deepest contact of a sequence visited in order is the first one of
// greatest depth.  Depths are positive, so (depth bits, ~order) compares as
// one unsigned 64-bit key whose maximum names that contact.
__device__ __forceinline__ unsigned long long depth_key(float pen, unsigned order) {
    return pen > 0.0f ? (static_cast<unsigned long long>(__float_as_uint(pen)) << 32) | (0xffffffffu - order) : 0ull;
}


__device__ __forceinline__ F3 shfl3(F3 a, int lane) { return f3(__shfl(a.x, lane, 64), __shfl(a.y, lane, 64), __shfl(a.z, lane, 64)); }

// Dynamic colliders, unsigned contacts: the deepest triangle contact
// (A41:3999-4323) then the deepest particle contact (A41:4354-4511) push the
// vertex by half their depth (A41:129).  Only the vertex's own position is
// read and written, so running both here equals the reference's two
// vertex-parallel launches.
//
// A vertex's hash probe is a chain of dependent loads, its length set by
// the vertex's bucket; vertices of one wave have buckets of very different
// sizes (most are turned away by a bucket bin or the coarse summary with one
// load).  So a wave pools its probes into one list of items -- (vertex,
// bucket entry) for triangles, (vertex, neighbour cell) for particles -- and
// its lanes take 64 items at a time.  Each item's (depth, ~visiting order)
// goes into its vertex's maximum in LDS: the deepest contact, the first in
// the reference's visiting order among equals.  The vertex then evaluates
// its winner once more for the push.  The entries of an unsigned triangle
// bucket are stored in no particular order, so there the visiting order is
// the key (triangle * 64 + cell index), the order of a walk of the bucket in
// ascending key order.  Every lane of the wave must call these (`active`
// false for lanes without a vertex).
// Above `pool_max` probing vertices a wave's particle searches run one per
// lane (27 cells each, a z plane of bins loaded together): the pooled list
// pays its owner lookups on every item, and with most lanes probing the
// per-lane walks are already balanced.  The launch takes the limit
// (kDefaultPoolMax unless tuned); every value gives the same contacts.

// The diagnostic weight of a triangle found by a query, for counters equal
// to those of one table over the whole scene: that table (hash_cell over
// `buckets`) lists the triangle once per cell of its span [lo, hi] hashed to
// the query cell (x, y, z)'s bucket, and A41:3934-3982 takes each listing as
// a candidate.  The search evaluates only the listing under the query's own
// cell, which carries this count.  Spans of hashed triangles hold at most 64
// cells.
__device__ inline int one_table_listings(int4 lo, int4 hi, int x, int y, int z, int buckets) {
    const unsigned h = hash_cell(x, y, z, buckets);
    int m = 0;
    for (int qz = lo.z; qz <= hi.z; ++qz)
        for (int qy = lo.y; qy <= hi.y; ++qy)
            for (int qx = lo.x; qx <= hi.x; ++qx) m += hash_cell(qx, qy, qz, buckets) == h ? 1 : 0;
    return m;
}

// The pooled triangle search of one wave (vertices at p; lanes without a
// vertex pass active false): the wave's (vertex, bucket entry) items, the
// rounds of 64 items `slice`, slice + nslices, slice + 2 nslices, ... of them
// (all rounds with one slice), each item's (depth, ~key) folded into its
// vertex's maximum in LDS.  Returns the lane's vertex's maximum over those
// rounds (0: no contact among them); `total` gets the wave's item count.
// The maximum over all slices names the same contact as one pass over all
// rounds: the key orders contacts exactly.  A hash with layer ranges is
// searched range by range (the rounds numbered on across the ranges), and
// in the lower-layers mode only the ranges below the query's layer.
__device__ inline unsigned long long dynamic_triangle_pool(const ClothView& c, F3 p, bool active, int slice, int nslices, bool count,
                                                           int& total_out) {
    const SSBL_CONST DynView& D = *(const SSBL_CONST DynView*)(c.dyn);
    const int lane = static_cast<int>(__lane_id());
    const float cm = fmaxf(c.collision_margin, 0.0f);
    const int nt = D.nt;
    const int nsl = max(nslices, 1);
    total_out = 0;
    // own cell only: the unsigned query radius is 0 (A41:3623-3633)
    int cx = 0, cy = 0, cz = 0;
    if (active) { cx = cell_coord(p.x, D.tri_cell); cy = cell_coord(p.y, D.tri_cell); cz = cell_coord(p.z, D.tri_cell); }
    unsigned cand = 0, rej = 0;
    const int qtag = active ? cell_tag(cx, cy, cz) : 0;
    __shared__ unsigned long long s_best[kTile];
    unsigned long long* best = s_best + (threadIdx.x & ~63u);
    best[lane] = 0ull;
    wave_lds_sync();
    const int classes = max(D.tri_classes, 1);
    int rounds_before = 0;   // rounds of the ranges searched so far (uniform)
    for (int cls = 0; cls < classes; ++cls) {
        // a range of layers at or above the query's holds nothing the
        // lower-layers filter accepts (uniform: one cloth per wave)
        if (classes > 1 && c.dyn_mode == 2 && D.tri_class_layer[cls] >= c.dyn_target_layer) continue;
        int4 bin = make_int4(0, 0, 0, -2);
        if (active) {
            bin = gload(D.tri_bin + tri_bucket(classes, D.tri_buckets, D.tri_class_base[cls], D.tri_class_base[cls + 1], cx, cy, cz));
            if (classes > 1) {
                // one layer per range (its filter is the range skip above);
                // the bin's third field is its entries' one cell tag
                if (bin.w == -2 || (c.dyn_mode != 0 && c.dyn_target_src != 0 && bin.w == c.dyn_target_src) ||
                    (bin.z != 0 && bin.z != cell_tag(cx, cy, cz)))
                    bin.y = 0;
            } else if (dyn_bucket_rejected(c, bin)) {
                bin.y = 0;
            }
        }
        // The wave's probes as one list of (vertex, entry) items, 64 at a
        // time: lane i owns items [start_i, start_i + n_i); item w belongs to
        // the last lane whose start is at most w.
        const int n_i = bin.y;
        int incl = n_i;
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        const int start = incl - n_i;
        const int total = __shfl(incl, 63, 64);
        total_out += total;
        // this slice's first round of the range: rounds are dealt to the
        // slices by their number counted across the ranges
        const int r0 = ((slice - rounds_before) % nsl + nsl) % nsl;
        rounds_before += (total + 63) / 64;
        if (64 * r0 >= total) continue;   // uniform over the wave
        const int first = bin.x - start;   // entry of item w is first + w in the owner's bucket
        // owner lane of item w: the last lane whose start is at most w
        auto owner_of = [&](int w) {
            int own = 0;
            for (int step = 32; step > 0; step >>= 1) {
                const int probe = own + step;
                const int sp = __shfl(start, min(probe, 63), 64);
                if (probe < 64 && sp <= w) own = probe;
            }
            return own;
        };
        auto entry_of = [&](int w, int own) {
            const int k = __shfl(first, own, 64) + w;
            return w < total ? gload(D.tri_ent + k) : make_int4(-1, 0, 0, 0);
        };
        // Items flow through three stages: a round's entry records are
        // requested two rounds ahead and its triangles' bounds and corners
        // one round ahead, so a round issues no load it waits for itself.
        // The source filter (in the entry) runs before the bounds are read;
        // it and the canonical-cell test (A41:3934-3982) are both side-effect
        // free.
        auto accepted = [&](int4 e) {
            return e.x >= 0 && e.x / kDynMaxHashCells < nt && (!D.tsrc || dyn_enabled(c, e.y, e.z));
        };
        // a triangle's bounds and corners, requested together one round
        // ahead (the corners whether or not the bounds pass)
        struct TriRec {
            TriBox bx;
            float4 a, b, c;
        };
        // only an entry listed under the query's own cell (its tag first) can
        // be the query's listing of its triangle: the others are skipped
        // before their records are requested
        auto record_of = [&](int4 e, int tag) {
            TriRec r{};
            if (accepted(e) && e.w == tag) {
                const int t = e.x / kDynMaxHashCells;
                r.bx = gload(D.box + t);
                r.a = gload(D.tv + 3 * t);
                r.b = gload(D.tv + 3 * t + 1);
                r.c = gload(D.tv + 3 * t + 2);
            }
            return r;
        };
        const int stride = 64 * nsl;
        const int b0 = 64 * r0;
        int own = owner_of(b0 + lane);
        int4 en = entry_of(b0 + lane, own);
        int own1 = owner_of(b0 + stride + lane);
        int4 en1 = entry_of(b0 + stride + lane, own1);
        TriRec rc = record_of(en, __shfl(qtag, own, 64));
        for (int base = b0; base < total; base += stride) {
            const int w = base + lane;
            const int ox = __shfl(cx, own, 64), oy = __shfl(cy, own, 64), oz = __shfl(cz, own, 64);
            const F3 op = shfl3(p, own);
            const int own2 = owner_of(w + 2 * stride);
            const int4 en2 = entry_of(w + 2 * stride, own2);
            const TriRec rc1 = record_of(en1, __shfl(qtag, own1, 64));
            if (w < total && accepted(en) && en.w == cell_tag(ox, oy, oz)) {
                const int4 lo = box_cmin(rc.bx), hi = box_cmax(rc.bx);
                // A triangle whose span holds the query's cell is listed under
                // that cell, in this bucket: that listing alone is evaluated
                // (its other listings here name the same triangle, so the
                // deepest contact and its triangle are the same)
                const int sx = hi.x - lo.x + 1, sy = hi.y - lo.y + 1;
                const bool own_listing = ((oz - lo.z) * sy + (oy - lo.y)) * sx + (ox - lo.x) == en.x % kDynMaxHashCells;
                if (ox >= lo.x && ox <= hi.x && oy >= lo.y && oy <= hi.y && oz >= lo.z && oz <= hi.z && own_listing) {
                    F3 d;
                    const unsigned rej0 = rej;
                    const float pen = dyn_triangle_depth_with(
                        rc.bx,
                        [&](F3& a, F3& b, F3& c) {
                            a = f3(rc.a);
                            b = f3(rc.b);
                            c = f3(rc.c);
                        },
                        op, cm, cand, rej, d);
                    if (count) {
                        // counters of one table over the whole scene, which
                        // lists the triangle once for each of its span cells
                        // hashed to the query's bucket (the contact is the
                        // same: every listing names the same triangle)
                        const unsigned m = static_cast<unsigned>(one_table_listings(lo, hi, ox, oy, oz, D.tri_buckets));
                        cand += m - 1u;
                        if (rej != rej0) rej += m - 1u;
                    }
                    const unsigned long long kk = depth_key(pen, static_cast<unsigned>(en.x));
                    if (kk != 0ull) atomicMax(best + own, kk);
                }
            }
            own = own1;
            en = en1;
            rc = rc1;
            own1 = own2;
            en1 = en2;
        }
    }
    wave_lds_sync();
    const unsigned long long top = best[lane];
    if (count && cand) atomicAdd(&c.counts[kCountDynTriCandidates], static_cast<unsigned long long>(cand));
    if (count && rej) atomicAdd(&c.counts[kCountDynTriAabbRejects], static_cast<unsigned long long>(rej));
    return top;
}

// The triangle contact of a vertex (A41:3999-4323) from the maximum key of
// its pooled search (`top`, 0: none): the winning entry evaluated again for
// its push, then the large triangles (walked by every vertex), then the
// response.  Without a triangle hash (small sets) the vertex walks every
// triangle here instead.
__device__ inline void dynamic_triangle_contact(const ClothView& c, VertexState& s, bool active, float dt, bool count,
                                                unsigned long long top) {
    const SSBL_CONST DynView& D = *(const SSBL_CONST DynView*)(c.dyn);
    const float cm = fmaxf(c.collision_margin, 0.0f);
    const int nt = D.nt;
    if (!active || nt <= 0 || !(cm > 0.0f)) return;
    float best_pen = 0.0f;
    F3 best_delta = f3(0.0f, 0.0f, 0.0f);
    unsigned cand = 0, rej = 0;
    if (D.tri_hash) {
        if (top != 0ull) {
            // the winning entry again, for its push (same inputs, same result)
            const int t = static_cast<int>(0xffffffffu - static_cast<unsigned>(top)) / kDynMaxHashCells;
            unsigned dc = 0, dr = 0;
            F3 d;
            best_pen = dyn_triangle_depth(D, t, gload(D.box + t), s.p, cm, dc, dr, d);
            best_delta = d;
        }
        const int nl = gload(D.nlarge);
        for (int k = 0; k < nl; ++k) {
            const int t = gload(D.large + k);
            if (t < 0 || t >= nt) continue;
            if (D.tsrc && !dyn_enabled(c, D.tsrc[t], D.tlayer ? D.tlayer[t] : 0)) continue;
            F3 d;
            const float pen = dyn_triangle_depth(D, t, D.box[t], s.p, cm, cand, rej, d);
            if (pen > best_pen) { best_pen = pen; best_delta = d; }
        }
    } else {
        for (int t = 0; t < nt; ++t) {
            if (D.tsrc && !dyn_enabled(c, D.tsrc[t], D.tlayer ? D.tlayer[t] : 0)) continue;
            F3 d;
            const float pen = dyn_triangle_depth(D, t, D.box[t], s.p, cm, cand, rej, d);
            if (pen > best_pen) { best_pen = pen; best_delta = d; }
        }
    }
    if (count && cand) atomicAdd(&c.counts[kCountDynTriCandidates], static_cast<unsigned long long>(cand));
    if (count && rej) atomicAdd(&c.counts[kCountDynTriAabbRejects], static_cast<unsigned long long>(rej));
    if (best_pen > 0.0f) {
        const F3 applied = mul(best_delta, 0.5f);
        const F3 np = s.p + applied;
        F3 v = s.vel();
        const float vn = dot(v, applied);
        if (vn < 0.0f) {
            const float lsq = fmaxf(dot(applied, applied), kEps);
            v = v - mul(applied, 0.5f * vn / lsq);
            s.set_vel(v);
        }
        s.p = np;
        s.set_q(np - mul(v, fmaxf(dt, kEps)));
        if (count) {
            atomicAdd(&c.counts[kCountTrianglePairs], 1ull);
            atomicAdd(&c.counts[kCountExactImpulse], 1ull);
        }
    }
}

// Calls f(entry, k) for the first `lim` entries of a particle bucket at
// `base`, in bucket order, loading them four at a time (an entry's cell,
// source and layer, and position: the loads of a batch are independent, so
// a batch costs one round trip instead of four).
template <class F>
__device__ __forceinline__ void for_bucket_entries(const SSBL_CONST DynView& D, int base, int lim, F f) {
    for (int k0 = 0; k0 < lim; k0 += 4) {
        int4 ck[4];
        int2 sl[4];
        float4 pp[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const bool in = k0 + u < lim;
            const PartEntry* e = D.pent + base + min(k0 + u, lim - 1);
            ck[u] = in ? gload(&e->ck) : make_int4(0, 0, 0, -1);
            sl[u] = in ? gload(reinterpret_cast<const int2*>(&e->sl)) : make_int2(0, 0);
            pp[u] = in ? gload(&e->p) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (k0 + u < lim) f(ck[u], sl[u], pp[u], k0 + u);
    }
}

__device__ inline int dynamic_particle_contacts(const ClothView& c, VertexState& s, bool active, bool count, int pool_max) {
    const SSBL_CONST DynView& D = *(const SSBL_CONST DynView*)(c.dyn);
    const int lane = static_cast<int>(__lane_id());
    const int np_count = D.np;
    int work = 0;   // the wave's cell searches
    if (np_count > 0) {
        const float margin = fmaxf(c.collision_margin, 0.0f);
        const float reach = D.max_radius + margin;   // no particle reaches further
        float best_pen = 0.0f;
        F3 best_delta = f3(0.0f, 0.0f, 0.0f);
        unsigned cand = 0;
        if (D.part_hash) {
            // the coarse summary holds for this query when its reach box covers
            // the query's reach (the largest radius plus this cloth's margin)
            bool probe = active;
            if (active && D.coarse && D.max_radius + margin <= D.coarse_reach && finite3(s.p) && *D.coarse_off == 0) {
                const float cc = D.coarse_cell;
                const int2 e = gload(D.coarse + hash_cell(cell_coord(s.p.x, cc), cell_coord(s.p.y, cc), cell_coord(s.p.z, cc), D.coarse_buckets));
                probe = !(e.y == -2 || (c.dyn_mode != 0 && ((c.dyn_target_src != 0 && e.y == c.dyn_target_src) ||
                                                            (c.dyn_mode == 2 && e.x >= c.dyn_target_layer))));
            }
            // Otherwise the wave's searches as one list of (vertex, cell)
            // items, 27 per probing vertex, 64 at a time (the owner lookup of
            // the triangle query).  Cell c = 9 (dz + 1) + 3 (dy + 1) + (dx + 1) is the
            // reference's visiting order, x innermost; an item's entries keep
            // their bucket order k, so (depth, ~(16 c + k)) orders contacts as
            // the serial search meets them.
            const unsigned long long probes = __ballot(probe);
            work = 27 * __popcll(probes);
            if (__popcll(probes) > pool_max) {
                // many probes in this wave: each lane searches its own 27 cells
                if (probe) {
                    const int cx = cell_coord(s.p.x, D.part_cell), cy = cell_coord(s.p.y, D.part_cell), cz = cell_coord(s.p.z, D.part_cell);
                    // the bins of a z plane are loaded together; their
                    // entries are then taken as one list in the reference's
                    // order (z, then y, then x innermost; bucket order
                    // within a cell), loaded four at a time
                    for (int dz = -1; dz <= 1; ++dz) {
                        int skip[9], pre[10];   // entry j of bin k is pent[j + skip[k]] for pre[k] <= j < pre[k + 1]
                        pre[0] = 0;
#pragma unroll
                        for (int k = 0; k < 9; ++k) {
                            const int4 bin = cell_in_reach(s.p, cx, cy, cz, k % 3 - 1, k / 3 - 1, dz, D.part_cell, reach)
                                                 ? gload(D.pbin + hash_cell(cx + k % 3 - 1, cy + k / 3 - 1, cz + dz, D.part_buckets))
                                                 : make_int4(0, 0, 0, -2);
                            const int n = dyn_bucket_rejected(c, bin) ? 0 : min(bin.y, kDynParticleSlots);
                            skip[k] = bin.x - pre[k];
                            pre[k + 1] = pre[k] + n;
                        }
                        const int n_plane = pre[9];
                        for (int j0 = 0; j0 < n_plane; j0 += 4) {
                            int4 ck[4];
                            int2 sl[4];
                            float4 pp[4];
                            int kb[4];
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                const int j = min(j0 + u, n_plane - 1);
                                int k = 0, at = j + skip[0];
#pragma unroll
                                for (int kk = 1; kk < 9; ++kk)
                                    if (pre[kk] <= j) { k = kk; at = j + skip[kk]; }
                                kb[u] = k;
                                const PartEntry* e = D.pent + at;
                                ck[u] = gload(&e->ck);
                                sl[u] = gload(reinterpret_cast<const int2*>(&e->sl));
                                pp[u] = gload(&e->p);
                            }
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                if (j0 + u >= n_plane) break;
                                const int k9 = kb[u];
                                const int qx = cx + k9 % 3 - 1, qy = cy + k9 / 3 - 1, qz = cz + dz;
                                if (ck[u].x != qx || ck[u].y != qy || ck[u].z != qz) continue;
                                if (D.psrc && !dyn_enabled(c, sl[u].x, sl[u].y)) continue;
                                if (ck[u].w < 0 || ck[u].w >= np_count) continue;
                                F3 d;
                                const float pen = dyn_particle_depth_at(pp[u], s.p, margin, cand, d);
                                if (pen > best_pen) { best_pen = pen; best_delta = d; }
                            }
                        }
                    }
                }
            }
            const int n_i = probe && __popcll(probes) <= pool_max ? 27 : 0;
            int incl = n_i;
            for (int d = 1; d < 64; d <<= 1) {
                const int o = __shfl_up(incl, d, 64);
                if (lane >= d) incl += o;
            }
            const int start = incl - n_i;
            const int total = __shfl(incl, 63, 64);
            if (total > 0) {   // uniform over the wave
                __shared__ unsigned long long s_best[kTile];
                unsigned long long* best = s_best + (threadIdx.x & ~63u);
                best[lane] = 0ull;
                wave_lds_sync();
                int cx = 0, cy = 0, cz = 0;
                if (probe) { cx = cell_coord(s.p.x, D.part_cell); cy = cell_coord(s.p.y, D.part_cell); cz = cell_coord(s.p.z, D.part_cell); }
                for (int base = 0; base < total; base += 64) {
                    const int w = base + lane;
                    int own = 0;
                    for (int step = 32; step > 0; step >>= 1) {
                        const int probe_lane = own + step;
                        const int sp = __shfl(start, min(probe_lane, 63), 64);
                        if (probe_lane < 64 && sp <= w) own = probe_lane;
                    }
                    const int ostart = __shfl(start, own, 64);
                    const int ox = __shfl(cx, own, 64), oy = __shfl(cy, own, 64), oz = __shfl(cz, own, 64);
                    const F3 op = shfl3(s.p, own);
                    if (w < total) {
                        const int cell = w - ostart;
                        const int dx = cell % 3 - 1, dy = (cell / 3) % 3 - 1, dz = cell / 9 - 1;
                        if (cell_in_reach(op, ox, oy, oz, dx, dy, dz, D.part_cell, reach)) {
                            const int qx = ox + dx, qy = oy + dy, qz = oz + dz;
                            const int4 bin = gload(D.pbin + hash_cell(qx, qy, qz, D.part_buckets));
                            if (!dyn_bucket_rejected(c, bin)) {
                                unsigned long long key = 0ull;
                                const int lim = min(bin.y, kDynParticleSlots);
                                for_bucket_entries(D, bin.x, lim, [&](int4 ck, int2 sl, float4 pp, int k) {
                                    if (ck.x != qx || ck.y != qy || ck.z != qz) return;
                                    if (D.psrc && !dyn_enabled(c, sl.x, sl.y)) return;
                                    if (ck.w < 0 || ck.w >= np_count) return;
                                    F3 d;
                                    const unsigned long long kk = depth_key(dyn_particle_depth_at(pp, op, margin, cand, d),
                                                                            static_cast<unsigned>(cell * kDynParticleSlots + k));
                                    key = kk > key ? kk : key;
                                });
                                if (key != 0ull) atomicMax(best + own, key);
                            }
                        }
                    }
                }
                wave_lds_sync();
                const unsigned long long top = best[lane];
                if (top != 0ull) {
                    // the winning particle again, for its push (same inputs, same result)
                    const unsigned order = 0xffffffffu - static_cast<unsigned>(top);
                    const int cell = static_cast<int>(order) / kDynParticleSlots, k = static_cast<int>(order) % kDynParticleSlots;
                    const int qx = cx + cell % 3 - 1, qy = cy + (cell / 3) % 3 - 1, qz = cz + cell / 9 - 1;
                    const int4 bin = gload(D.pbin + hash_cell(qx, qy, qz, D.part_buckets));
                    unsigned dc = 0;
                    F3 d;
                    best_pen = dyn_particle_depth_at(gload(D.pent + bin.x + k).p, s.p, margin, dc, d);
                    best_delta = d;
                }
            }
        } else if (active) {
            for (int j = 0; j < np_count; ++j) {
                if (D.psrc && !dyn_enabled(c, D.psrc[j], D.player ? D.player[j] : 0)) continue;
                F3 d;
                const float pen = dyn_particle_depth(D, j, s.p, margin, cand, d);
                if (pen > best_pen) { best_pen = pen; best_delta = d; }
            }
        }
        if (count && cand) atomicAdd(&c.counts[kCountDynParticleCandidates], static_cast<unsigned long long>(cand));
        if (active && best_pen > 0.0f) {
            s.p = s.p + mul(best_delta, 0.5f);
            if (count) {
                atomicAdd(&c.counts[kCountExactImpulse], 1ull);
                atomicAdd(&c.counts[kCountDynParticleContacts], 1ull);
            }
        }
    }
    return work;
}

// A41:5846-5884
__device__ inline void velocity_update(const ClothView& c, VertexState& s, float dt) {
    const float sdt = fmaxf(dt, kEps);
    const F3 meas = mul(s.p - s.q(), 1.0f / sdt);
    const F3 base = s.vel();
    if (!finite3(meas) || !finite3(base)) { s.set_vel(f3(0.0f, 0.0f, 0.0f)); return; }
    const F3 pv = meas - base;
    const float pdb = dot(pv, base);
    const float fs = (isfinite(pdb) && pdb > 0.0f) ? 0.9f : 1.0f;
    const float max_fd = fmaxf(2.0e-4f, fmaxf(c.thickness, 0.0f) * 0.35f);
    const float max_fs = max_fd / sdt;
    F3 fb = mul(pv, fs);
    const float speed = len3(fb);
    if (fs < 1.0f && isfinite(speed) && speed > max_fs && speed > kEps) fb = mul(fb, max_fs / speed);
    if (!finite3(fb)) fb = f3(0.0f, 0.0f, 0.0f);
    s.set_vel(base + fb);
}

// ---------------------------------------------------------- force fields
// A41:1263-1531 (Blender-compatible improved Perlin noise for turbulence).
__constant__ unsigned char kNoisePerm[256] = {
    0xA2,0xA0,0x19,0x3B,0xF8,0xEB,0xAA,0xEE,0xF3,0x1C,0x67,0x28,0x1D,0xED,0x0,0xDE,0x95,0x2E,0xDC,0x3F,0x3A,0x82,0x35,0x4D,0x6C,0xBA,0x36,0xD0,0xF6,0xC,0x79,0x32,
    0xD1,0x59,0xF4,0x8,0x8B,0x63,0x89,0x2F,0xB8,0xB4,0x97,0x83,0xF2,0x8F,0x18,0xC7,0x51,0x14,0x65,0x87,0x48,0x20,0x42,0xA8,0x80,0xB5,0x40,0x13,0xB2,0x22,0x7E,0x57,
    0xBC,0x7F,0x6B,0x9D,0x86,0x4C,0xC8,0xDB,0x7C,0xD5,0x25,0x4E,0x5A,0x55,0x74,0x50,0xCD,0xB3,0x7A,0xBB,0xC3,0xCB,0xB6,0xE2,0xE4,0xEC,0xFD,0x98,0xB,0x96,0xD3,0x9E,
    0x5C,0xA1,0x64,0xF1,0x81,0x61,0xE1,0xC4,0x24,0x72,0x49,0x8C,0x90,0x4B,0x84,0x34,0x38,0xAB,0x78,0xCA,0x1F,0x1,0xD7,0x93,0x11,0xC1,0x58,0xA9,0x31,0xF9,0x44,0x6D,
    0xBF,0x33,0x9C,0x5F,0x9,0x94,0xA3,0x85,0x6,0xC6,0x9A,0x1E,0x7B,0x46,0x15,0x30,0x27,0x2B,0x1B,0x71,0x3C,0x5B,0xD6,0x6F,0x62,0xAC,0x4F,0xC2,0xC0,0xE,0xB1,0x23,
    0xA7,0xDF,0x47,0xB0,0x77,0x69,0x5,0xE9,0xE6,0xE7,0x76,0x73,0xF,0xFE,0x6E,0x9B,0x56,0xEF,0x12,0xA5,0x37,0xFC,0xAE,0xD9,0x3,0x8E,0xDD,0x10,0xB9,0xCE,0xC9,0x8D,
    0xDA,0x2A,0xBD,0x68,0x17,0x9F,0xBE,0xD4,0xA,0xCC,0xD2,0xE8,0x43,0x3D,0x70,0xB7,0x2,0x7D,0x99,0xD8,0xD,0x60,0x8A,0x4,0x2C,0x3E,0x92,0xE5,0xAF,0x53,0x7,0xE0,
    0x29,0xA6,0xC5,0xE3,0xF5,0xF7,0x4A,0x41,0x26,0x6A,0x16,0x5E,0x52,0x2D,0x21,0xAD,0xF0,0x91,0xFF,0xEA,0x54,0xFA,0x66,0x1A,0x45,0x39,0xCF,0x75,0xA4,0x88,0xFB,0x5D};

__device__ __forceinline__ int perm(int i) { return kNoisePerm[i & 255]; }
__device__ __forceinline__ float fade(float t) { return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f); }
__device__ __forceinline__ float nlerp(float t, float a, float b) { return a + t * (b - a); }
__device__ __forceinline__ float grad(int h, float x, float y, float z) {
    h &= 15;
    const float u = h < 8 ? x : y;
    const float v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) == 0 ? u : -u) + ((h & 2) == 0 ? v : -v);
}
__device__ inline float perlin(float x, float y, float z) {
    const float fx = floorf(x), fy = floorf(y), fz = floorf(z);
    const int X = static_cast<int>(fx) & 255, Y = static_cast<int>(fy) & 255, Z = static_cast<int>(fz) & 255;
    x -= fx; y -= fy; z -= fz;
    const float u = fade(x), v = fade(y), w = fade(z);
    // the 512-entry table of the reference is the 256-entry permutation twice
    const int A = perm(X) + Y, AA = perm(A) + Z, AB = perm(A + 1) + Z;
    const int B = perm(X + 1) + Y, BA = perm(B) + Z, BB = perm(B + 1) + Z;
    return nlerp(w,
                 nlerp(v, nlerp(u, grad(perm(AA), x, y, z), grad(perm(BA), x - 1.0f, y, z)),
                       nlerp(u, grad(perm(AB), x, y - 1.0f, z), grad(perm(BB), x - 1.0f, y - 1.0f, z))),
                 nlerp(v, nlerp(u, grad(perm(AA + 1), x, y, z - 1.0f), grad(perm(BA + 1), x - 1.0f, y, z - 1.0f)),
                       nlerp(u, grad(perm(AB + 1), x, y - 1.0f, z - 1.0f), grad(perm(BB + 1), x - 1.0f, y - 1.0f, z - 1.0f))));
}
__device__ inline float turbulence(float size, float x, float y, float z) {
    if (size != 0.0f && isfinite(size)) { const float inv = 1.0f / size; x *= inv; y *= inv; z *= inv; }
    float sum = 0.0f, amp = 1.0f, fs = 1.0f;
    for (int i = 0; i <= 2; ++i, amp *= 0.5f, fs *= 2.0f) sum += (0.5f + 0.5f * perlin(fs * x, fs * y, fs * z)) * amp;
    return sum * (4.0f / 7.0f);
}
// sin and pow are evaluated in double and rounded once: the hash noise
// multiplies sin by 43758.5, so a one-ulp difference between maths libraries
// would change the noise value entirely.
__device__ __forceinline__ float sin_r(float x) { return static_cast<float>(sin(static_cast<double>(x))); }
__device__ __forceinline__ float pow_r(float x, float y) {
    return static_cast<float>(pow(static_cast<double>(x), static_cast<double>(y)));
}
__device__ inline float field_noise(F3 p, int seed, float salt) {
    const float value = sin_r(p.x * 12.9898f + p.y * 78.233f + p.z * 37.719f + (static_cast<float>(seed) + salt) * 19.191f) * 43758.5453f;
    return (value - floorf(value)) * 2.0f - 1.0f;
}
__device__ inline float falloff(const SsblXpbdForceField& f, float distance, float radial) {
    if (!isfinite(distance)) return 0.0f;
    const float mn = fmaxf(f.distance_min, 0.0f), mx = fmaxf(f.distance_max, 0.0f);
    const float rmn = fmaxf(f.radial_min, 0.0f), rmx = fmaxf(f.radial_max, 0.0f);
    if (f.use_max_distance && mx > 0.0f && distance >= mx) return 0.0f;
    if (f.use_radial_max && rmx > 0.0f && radial >= rmx) return 0.0f;
    float att = 1.0f;
    if (f.use_max_distance && mx > mn && distance > mn) att *= clamp01f((mx - distance) / fmaxf(mx - mn, kEps));
    if (f.use_min_distance && f.falloff_power > 0.0f && distance > mn) {
        const float ref = fmaxf(mn, 1.0e-4f);
        att *= pow_r(ref / fmaxf(distance, ref), f.falloff_power);
    }
    if (f.use_radial_max && rmx > rmn && radial > rmn) att *= clamp01f((rmx - radial) / fmaxf(rmx - rmn, kEps));
    if (f.use_radial_min && f.radial_falloff > 0.0f && radial > rmn) {
        const float ref = fmaxf(rmn, 1.0e-4f);
        att *= pow_r(ref / fmaxf(radial, ref), f.radial_falloff);
    }
    return isfinite(att) ? att : 0.0f;
}
__device__ inline F3 limit_accel(F3 a) {
    const float len = len3(a);
    if (!isfinite(len) || len <= 5000.0f) return isfinite(len) ? a : f3(0.0f, 0.0f, 0.0f);
    return mul(a, 5000.0f / fmaxf(len, kEps));
}
__device__ inline F3 with_flow(const SsblXpbdForceField& f, F3 force, F3 vel, float fo) {
    if (f.type != 6 && f.type != 9) {
        const float flow = isfinite(f.flow) ? f.flow : 0.0f;
        if (fabsf(flow) > 1.0e-6f && fo != 0.0f) force = force + mul(vel, -flow * fo);
    }
    return limit_accel(force);
}
__device__ inline F3 eval_field(const SsblXpbdForceField& f, F3 p, F3 vel, F3 n, bool has_n) {
    const F3 zero = f3(0.0f, 0.0f, 0.0f);
    if (!isfinite(f.strength) || f.strength == 0.0f || f.apply_to_location == 0) return zero;
    const F3 delta = p - f3(f.origin[0], f.origin[1], f.origin[2]);
    const F3 axis = normalize_or(f3(f.axis[0], f.axis[1], f.axis[2]), f3(0.0f, 0.0f, 1.0f));
    const F3 radial_delta = f.use_2d_force ? delta - mul(axis, dot(delta, axis)) : delta;
    const float distance = len3(delta), radial_distance = len3(radial_delta);
    const float fo = falloff(f, distance, radial_distance);
    const float strength = f.strength * fo;
    if (strength == 0.0f || !isfinite(strength)) return zero;
    switch (f.type) {
    case 1: {
        if (!has_n) return zero;
        const F3 dir = normalize_or(f3(f.direction[0], f.direction[1], f.direction[2]), zero);
        const float proj = dot(dir, n);
        if (!isfinite(proj) || proj <= 0.0f) return zero;
        float mag = strength * proj;
        if (f.noise > 0.0f) mag *= fmaxf(0.0f, 1.0f + f.noise * field_noise(p, f.seed, 0.0f));
        return with_flow(f, mul(dir, mag), vel, fo);
    }
    case 2: case 5: {
        const F3 sd = f.use_2d_force ? radial_delta : delta;
        const float sdist = f.use_2d_force ? radial_distance : distance;
        if (sdist <= 1.0e-6f) return zero;
        float scale = strength / fmaxf(sdist, kEps);
        if (f.type == 5) scale /= fmaxf(sdist * sdist, 0.01f);
        return with_flow(f, mul(sd, scale), vel, fo);
    }
    case 3: {
        const F3 radial = delta - mul(axis, dot(delta, axis));
        if (dot(radial, radial) <= 1.0e-10f) return zero;
        return with_flow(f, mul(normalize_or(cross(axis, radial), zero), strength), vel, fo);
    }
    case 6: {
        const float sdist = f.use_2d_force ? radial_distance : distance;
        const F3 sd = f.use_2d_force ? radial_delta : delta;
        if (sdist <= 1.0e-6f) return zero;
        const F3 dir = mul(sd, 1.0f / fmaxf(sdist, kEps));
        const float spring = -strength * (sdist - fmaxf(f.rest_length, 0.0f));
        const float damping = -dot(vel, dir) * fmaxf(f.harmonic_damping, 0.0f) * fabsf(strength);
        return with_flow(f, mul(dir, spring + damping), vel, fo);
    }
    case 7: {
        if (distance <= 1.0e-6f) return zero;
        const F3 dir = mul(delta, 1.0f / fmaxf(distance, kEps));
        const float radius = fmaxf(f.rest_length, fmaxf(f.size, 0.1f));
        const float ratio = fminf(radius / fmaxf(distance, 1.0e-3f), 6.0f);
        const float r2 = ratio * ratio, r6 = r2 * r2 * r2;
        return with_flow(f, mul(dir, strength * (r6 * r6 - r6)), vel, fo);
    }
    case 8: {
        const F3 ax = normalize_or(f3(f.direction[0], f.direction[1], f.direction[2]), zero);
        return with_flow(f, mul(cross(vel, ax), strength), vel, fo);
    }
    case 9: {
        const float speed = len3(vel);
        if (speed <= 1.0e-6f) return zero;
        float lin = fmaxf(f.linear_drag, 0.0f);
        const float quad = fmaxf(f.quadratic_drag, 0.0f);
        if (lin <= 0.0f && quad <= 0.0f) lin = fabsf(strength);
        return with_flow(f, mul(vel, -(lin + quad * speed)), vel, fo);
    }
    case 4: {
        const F3 q = f.use_global_coords ? p : delta + axis;
        const F3 nv = f3(-1.0f + 2.0f * turbulence(f.size, q.x, q.y, q.z), -1.0f + 2.0f * turbulence(f.size, q.y, q.z, q.x),
                         -1.0f + 2.0f * turbulence(f.size, q.z, q.x, q.y));
        return with_flow(f, mul(nv, strength), vel, fo);
    }
    case 10: {
        float freq = fmaxf(fmaxf(f.noise, f.texture_nabla), 0.25f);
        if (f.size > 1.0e-6f) freq = fmaxf(freq, 1.0f / f.size);
        const F3 q = mul(delta, freq);
        const F3 nv = f3(field_noise(q, f.seed, 0.0f), field_noise(q, f.seed, 7.0f), field_noise(q, f.seed, 13.0f));
        return with_flow(f, mul(nv, strength), vel, fo);
    }
    default:
        return zero;
    }
}

// ------------------------------------------------------------- integrate
// A41:1213-1261 area-free vertex normal over the incident triangles, read
// from the pass's source positions.
__device__ inline F3 vertex_normal(const ClothView& c, const float4* src, int v, bool* has) {
    *has = false;
    const int s = c.tri_off[v], e = c.tri_off[v + 1];
    if (e <= s) return f3(0.0f, 0.0f, 0.0f);
    F3 acc = f3(0.0f, 0.0f, 0.0f);
    for (int k = s; k < e; ++k) {
        const int4 t = c.tris[c.tri_inc[k]];
        const F3 a = f3(src[t.x]), b = f3(src[t.y]), cc = f3(src[t.z]);
        const F3 n = cross(b - a, cc - a);
        const float lsq = dot(n, n);
        if (!isfinite(lsq) || lsq <= 1.0e-10f) continue;
        acc = acc + mul(n, rsqrt_exact(lsq));
    }
    const float asq = dot(acc, acc);
    if (!isfinite(asq) || asq <= 1.0e-10f) return f3(0.0f, 0.0f, 0.0f);
    *has = true;
    return mul(acc, rsqrt_exact(asq));
}

// A41:1869-1925 without forces, for a vertex whose state s.p / velocity are
// in registers: the previous position becomes s.p, a pinned vertex goes to
// its rest position with zero velocity, a free one moves by its damped,
// gravity-accelerated velocity.  Leaves s ready to be stored, all dirty.
__device__ __forceinline__ void integrate_plain(const ClothView& c, VertexState& s, bool pinned, F3 vel, float dt) {
    s.q_ = s.p;
    s.q_have = s.q_dirty = true;
    s.v_have = s.v_dirty = true;
    if (pinned) {
        s.p = f3(gload(c.rest + s.v));
        s.vel_ = f3(0.0f, 0.0f, 0.0f);
        return;
    }
    const F3 acc = f3(c.gravity[0], c.gravity[1], c.gravity[2]);
    F3 nv = vel + mul(acc, dt);
    nv = mul(nv, c.damping);
    s.vel_ = nv;
    s.p = s.p + mul(nv, dt);
}

// kFull: normals, overpressure and force fields are compiled in (the plain
// variant keeps the register footprint, and so the occupancy, of the common
// gravity-only case).
template <bool kFull>
__global__ __launch_bounds__(kTile) void k_integrate(const ClothView* __restrict__ views, const Tile* __restrict__ tiles,
                                                     int ntiles, int parity, unsigned ops) {
    const Tile tile = tiles[xcd_tile(ntiles)];
    const int v = tile.begin + static_cast<int>(threadIdx.x);
    if (v >= tile.end) return;
    const ClothView& c = views[tile.cloth];
    const float4* src = parity ? c.pos1 : c.pos0;
    float4* dst = parity ? c.pos0 : c.pos1;
    const float dt = c.dt_sub;
    const float4 p4 = src[v];
    const bool pinned = p4.w == 0.0f;
    // the previous position becomes the pass's source position
    VertexState s{&c, v, f3(p4), f3(p4), f3(0.0f, 0.0f, 0.0f), true, true, true, true};
    if (pinned) {
        s.p = f3(c.rest[v]);
    } else {
        const F3 vel = f3(c.vel[v]);
        if constexpr (kFull) {
            F3 acc = f3(c.gravity[0], c.gravity[1], c.gravity[2]);
            F3 nv;
            const F3 p = s.p;
            bool has = false;
            const F3 n = c.need_normal ? vertex_normal(c, src, v, &has) : f3(0.0f, 0.0f, 0.0f);
            const float im = p4.w;
            const bool pressure = c.use_pressure && has;
            if (pressure) {
                const float pf = fmaxf(c.pressure_strength, 0.0f);
                const float share = c.area_share[v];
                if (isfinite(share) && share > 0.0f) acc = acc + mul(n, pf * c.pressure_accel_scale * share * im);
            }
            if (c.nfield > 0) {
                F3 fa = f3(0.0f, 0.0f, 0.0f);
                const int cnt = min(c.nfield, kMaxForceFields);
                for (int k = 0; k < cnt; ++k) fa = fa + eval_field(c.fields[k], p, vel, n, has);
                acc = acc + limit_accel(fa);
            }
            nv = vel + mul(acc, dt);
            nv = mul(nv, c.damping);
            if (pressure) {
                const float local_rest = c.rest_scale[v];
                if (isfinite(local_rest) && local_rest > kEps && isfinite(dt) && dt > kEps) {
                    const float pa = fmaxf(c.pressure_strength, 0.0f) * fmaxf(c.pressure_accel_scale, 0.0f) *
                                     fmaxf(c.area_share[v], 0.0f) * fmaxf(im, 0.0f);
                    const float max_pd = fmaxf(local_rest * fmaxf(c.pressure_impulse_rest_scale, 0.0f), 1.0e-7f);
                    const float pd = pa * dt * dt;
                    if (isfinite(pd) && pd > max_pd) nv = nv - mul(n, (pd - max_pd) / dt);
                    const float vr = fmaxf(local_rest, fmaxf(c.pressure_velocity_length_scale, 0.0f));
                    const float max_out = fmaxf(vr * fmaxf(c.pressure_velocity_rest_scale, 0.0f) / dt, 1.0e-6f);
                    const float out = dot(nv, n);
                    if (isfinite(out) && out > max_out) nv = nv - mul(n, out - max_out);
                }
            }
            s.vel_ = nv;
            s.p = s.p + mul(nv, dt);
        } else {
            integrate_plain(c, s, false, vel, dt);
        }
    }
    if (ops & kOpPin) apply_pins(c, s, pinned, dt);
    dst[v] = with_w(s.p, p4.w);
    c.prev[v] = with_w(s.q_, 0.0f);
    c.vel[v] = with_w(s.vel_, 0.0f);
}

// Dynamic collider contacts of the last iteration (A41:9230-9252), in two
// launches: the pooled triangle searches (k_dynamic_triangles), then a pass
// of their own (the hash probes want registers the constraint passes do
// not): each vertex's triangle contact from the searches' result, then its
// particle contacts on the updated position, as the reference's two
// vertex-parallel launches do; only the vertex's own state is read and
// written.  `tail_ops` are the iteration's closing vertex-local steps when
// nothing else runs before them: the second pin projection and the velocity
// update, exactly as the in-place vertex pass would apply them, for every
// vertex of every cloth in the launch.
//
// The triangle searches of a wave can take tens of rounds where the cloths
// pile up, and one launch lasts as long as its slowest wave.  So they run as
// units of (tile, slice, slices): the unit's waves take every slices-th
// round of their lists, starting at round `slice`, and fold their vertices'
// maxima into dyn_best (a device atomic maximum per vertex with a contact).
// Heavy tiles get up to kDynMaxSlices units, so their rounds spread over
// that many workgroups.  The order kernel below deals the units out from the
// step before's work.  The contact pass reads each vertex's key and clears it.
__global__ __launch_bounds__(kTile) void k_dynamic_triangles(const ClothView* __restrict__ views, const Tile* __restrict__ tiles,
                                                             int parity, unsigned ops, const int* __restrict__ units,
                                                             const int* __restrict__ nunits, int* tri_work) {
    if (static_cast<int>(blockIdx.x) >= *nunits) return;   // uniform over the block
    const int u = units[blockIdx.x];
    const int ti = u & kUnitTileMask, slice = (u >> kUnitSliceShift) & 15, nsl = ((u >> kUnitSlicesShift) & 15) + 1;
    const Tile tile = tiles[ti];
    const ClothView& c = views[tile.cloth];
    if (c.dyn == nullptr || !c.dyn_best) return;
    {
        const SSBL_CONST DynView& D = *(const SSBL_CONST DynView*)(c.dyn);
        if (!D.tri_hash || D.nt <= 0 || !(fmaxf(c.collision_margin, 0.0f) > 0.0f)) return;
    }
    const int v = tile.begin + static_cast<int>(threadIdx.x);
    const float4* pos = parity ? c.pos1 : c.pos0;
    const bool in_tile = v < tile.end;
    const float4 p4 = in_tile ? gload(pos + v) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const bool active = in_tile && p4.w != 0.0f;
    int total = 0;
    const unsigned long long top = dynamic_triangle_pool(c, f3(p4), active, slice, nsl, (ops & kOpCount) != 0u, total);
    if (active && top != 0ull) atomicMax(c.dyn_best + v, top);
    // the tile's heaviest wave list, for the next step's units
    if (slice == 0 && tri_work && (threadIdx.x & 63u) == 0u && total > 0) atomicMax(&tri_work[ti], total);
}

__global__ __launch_bounds__(kTile) void k_dynamic_pass(const ClothView* __restrict__ views, const Tile* __restrict__ tiles,
                                                        int ntiles, int parity, unsigned ops, unsigned tail_ops,
                                                        const int* __restrict__ tile_order, int* tile_work) {
    // with an order, the tiles whose particle searches took longest in the
    // step before come first: the launch lasts as long as its slowest waves,
    // so those start at once instead of in a later round of waves
    const int ti = tile_order ? tile_order[blockIdx.x] : xcd_tile(ntiles);
    const Tile tile = tiles[ti];
    const ClothView& c = views[tile.cloth];
    const bool has_dyn = c.dyn != nullptr;
    const bool tail = (tail_ops & kOpVelocity) != 0u;
    if (!has_dyn && !tail) return;   // uniform over the block
    const int v = tile.begin + static_cast<int>(threadIdx.x);
    // every lane stays for the wave-wide probes; lanes past the tile or on a
    // pinned vertex take part without a vertex of their own
    float4* pos = parity ? c.pos1 : c.pos0;
    const bool in_tile = v < tile.end;
    const float4 p4 = in_tile ? gload(pos + v) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    unsigned long long top = (has_dyn && in_tile && c.dyn_best) ? gload(c.dyn_best + v) : 0ull;
    const bool active = in_tile && p4.w != 0.0f;
    VertexState s{&c, v, f3(p4), f3(0.0f, 0.0f, 0.0f), f3(0.0f, 0.0f, 0.0f), false, false, false, false};
    // the closing steps' reads (pin range, previous position, velocity) go
    // out with the vertex's first loads instead of after the searches
    VertexPins pins;
    pins.range(c, v, 0u);
    if (tail && in_tile) {
        pins.range(c, v, tail_ops);
        s.q_ = gload_xyz(c.prev + v);
        s.vel_ = gload_xyz(c.vel + v);
        s.q_have = s.v_have = true;
        pins.request_first(c);
    }
    if (has_dyn) {
        if (top != 0ull) gstore(c.dyn_best + v, 0ull);   // ready for the next searches
        if (!active) top = 0ull;
        const bool count = (ops & kOpCount) != 0u;
        dynamic_triangle_contact(c, s, active, c.dt_sub, count, top);
        const int work = dynamic_particle_contacts(c, s, active, count, static_cast<int>((ops >> kOpPoolShift) & 0x7fu));
        if (tile_work && (threadIdx.x & 63u) == 0u && work > 0) atomicAdd(&tile_work[ti], work);
    }
    if (!in_tile) return;
    if (tail) {
        const float dt = c.dt_sub;
        const bool pinned = p4.w == 0.0f;
        if ((tail_ops & kOpPinPolish) && ((tail_ops & kOpLastIter) || c.contacts_on)) pins.apply(c, s, pinned, dt);
        if (pinned) s.set_vel(f3(0.0f, 0.0f, 0.0f));
        else velocity_update(c, s, dt);
        if (tail_ops & kOpIntegrate) {
            // the next substep's integration, as k_integrate<false> would read
            // this state back: the substep's final position stays in this
            // buffer, the integrated one goes to the other
            gstore(pos + v, with_w(s.p, p4.w));
            integrate_plain(c, s, pinned, s.vel_, dt);
            if (tail_ops & kOpIntegratePin) pins.apply(c, s, pinned, dt);
            gstore((parity ? c.pos0 : c.pos1) + v, with_w(s.p, p4.w));
            gstore(c.prev + v, with_w(s.q_, 0.0f));
            gstore(c.vel + v, with_w(s.vel_, 0.0f));
            return;
        }
    } else if (!active) {
        return;
    }
    if (s.p.x != p4.x || s.p.y != p4.y || s.p.z != p4.z) gstore(pos + v, with_w(s.p, p4.w));
    if (s.q_dirty) gstore(c.prev + v, with_w(s.q_, 0.0f));
    if (s.v_dirty) gstore(c.vel + v, with_w(s.vel_, 0.0f));
}

// The next step's launch plan of the dynamic contacts, from the work of the
// step before (then cleared), in one workgroup:
//  - the contact pass's tile order: tiles by their particle-search work
//    (work[0, n)), most first, counting-sorted on the bit length of the work;
//  - the triangle search units: tiles by the longest item list of one of
//    their waves (work[n, 2n)), most first, each tile with
//    ceil(list / kDynSliceItems) slices (1 to kDynMaxSlices), extra slices
//    granted down that order while the launch's n spare units last.
// Layout of `order` (dynamic_order_ints): [0, n) the pass's order, [n, 2n)
// the tiles sorted by triangle work, [2n] the unit count, [2n + 1, 4n + 1)
// the units, tile | slice << kUnitSliceShift | (slices - 1) << kUnitSlicesShift.
__global__ __launch_bounds__(1024) void k_order_tiles(int ntiles, int* work, int* order) {
    __shared__ int hist[33];
    __shared__ int sums[1024];
    const int tid = static_cast<int>(threadIdx.x);
    auto key = [](int w) { return 32 - (w > 0 ? 32 - __clz(w) : 0); };   // more work, smaller key
    auto sort = [&](const int* w, int* out) {
        if (tid < 33) hist[tid] = 0;
        __syncthreads();
        for (int t = tid; t < ntiles; t += 1024) atomicAdd(&hist[key(w[t])], 1);
        __syncthreads();
        if (tid == 0) {
            int run = 0;
            for (int k = 0; k < 33; ++k) {
                const int n = hist[k];
                hist[k] = run;
                run += n;
            }
        }
        __syncthreads();
        for (int t = tid; t < ntiles; t += 1024) out[atomicAdd(&hist[key(w[t])], 1)] = t;
        __syncthreads();
    };
    sort(work, order);
    int* sorted = order + ntiles;
    sort(work + ntiles, sorted);
    // wanted extra slices of sorted position q: d_q; granted down the order
    // while the spare units (ntiles) last, so position q's units start at
    // q + min(E_q, ntiles) with E_q = d_0 + ... + d_{q-1}
    const int chunk = (ntiles + 1023) / 1024;
    const int lo = min(tid * chunk, ntiles), hi = min(lo + chunk, ntiles);
    auto extra = [&](int q) {
        const int w = work[ntiles + sorted[q]];
        return w <= kDynSliceItems ? 0 : min((w - 1) / kDynSliceItems + 1, kDynMaxSlices) - 1;
    };
    int mine = 0;
    for (int q = lo; q < hi; ++q) mine += extra(q);
    sums[tid] = mine;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {   // inclusive scan of the per-thread sums
        const int o = tid >= d ? sums[tid - d] : 0;
        __syncthreads();
        sums[tid] += o;
        __syncthreads();
    }
    int e = sums[tid] - mine;   // E at position lo
    int* units = order + 2 * ntiles + 1;
    for (int q = lo; q < hi; ++q) {
        const int dq = extra(q);
        const int first = q + min(e, ntiles);
        const int slices = 1 + min(e + dq, ntiles) - min(e, ntiles);
        for (int k = 0; k < slices; ++k)
            units[first + k] = sorted[q] | (k << kUnitSliceShift) | ((slices - 1) << kUnitSlicesShift);
        e += dq;
    }
    if (tid == 1023) order[2 * ntiles] = ntiles + min(sums[1023], ntiles);
    __syncthreads();
    for (int t = tid; t < 2 * ntiles; t += 1024) work[t] = 0;
}

// ------------------------------------------------- signed dynamic contacts
// A41:80-83, 130-137
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

// A41:978-993: the depth-weighted mean direction of all contacts at the
// length of the kept one.
__device__ inline F3 local_average_delta(F3 acc, float w, F3 fallback) {
    if (w > 0.0f && finite3(acc)) {
        const F3 avg = mul(acc, 1.0f / w);
        const float al = len3(avg), fl = len3(fallback);
        if (finite3(avg) && isfinite(al) && isfinite(fl) && al > kEps && fl > kEps) return mul(avg, fl / al);
    }
    return fallback;
}

// A41:995-1068
__device__ inline void signed_response(const ClothView& c, VertexState& s, F3 orig, F3 raw, F3 cvel, float proxy_depth, float dt,
                                       bool inside, bool count) {
    if (!finite3(orig) || !finite3(raw)) return;
    const float raw_len = len3(raw);
    if (!isfinite(raw_len) || raw_len <= kEps) return;
    const F3 n = mul(raw, 1.0f / raw_len);
    F3 applied = mul(raw, inside ? kSignedInsideRelaxation : kSignedRelaxation);
    {   // A41:951-968
        const float basis = fmaxf(fmaxf(c.thickness, 0.0f), fmaxf(c.collision_margin, 0.0f));
        const float max_len = inside ? fmaxf(kSignedInsideMaxDeltaFloor, basis * kSignedInsideMaxDeltaScale)
                                     : fmaxf(kDynamicMaxDeltaFloor, basis * kDynamicMaxDeltaScale);
        if (!finite3(applied) || !isfinite(max_len) || max_len <= 0.0f) applied = f3(0.0f, 0.0f, 0.0f);
        else {
            const float len = len3(applied);
            if (!isfinite(len) || len <= kEps) applied = f3(0.0f, 0.0f, 0.0f);
            else if (len > max_len) applied = mul(applied, max_len / len);
        }
    }
    const float applied_len = len3(applied);
    if (!isfinite(applied_len) || applied_len <= kEps) return;
    const float sdt = fmaxf(dt, kEps);
    const F3 corrected = orig + applied;
    F3 v = s.vel();
    F3 base = f3(0.0f, 0.0f, 0.0f);
    if (inside) v = finite3(v) ? v : f3(0.0f, 0.0f, 0.0f);
    if (finite3(cvel)) {
        const float cn = dot(cvel, n);
        if (isfinite(cn)) base = mul(n, cn);
    }
    F3 rel = v - base;
    if (inside) {
        const float ns = dot(rel, n);
        if (isfinite(ns) && ns < 0.0f) {
            rel = rel - mul(n, ns);
            v = base + rel;
        }
    } else {
        const float inward = fmaxf(-dot(rel, n), 0.0f);
        const float proxy = fmaxf(fmaxf(proxy_depth, raw_len), applied_len);
        contact_response(c, n, inward + proxy / sdt, rel, count);
        const float outward = dot(rel, n);
        if (outward > 0.0f) rel = rel - mul(n, outward * (1.0f - kSignedOutwardRetention));
        v = base + rel;
    }
    s.p = corrected;
    const F3 vf = finite3(v) ? v : f3(0.0f, 0.0f, 0.0f);
    s.set_vel(vf);
    s.set_q(corrected - mul(vf, sdt));
}

// A41:3600-3633, 3999-4323, 4354-4511, 4520-4540.  The triangle query visits
// every hash cell within the query radius and takes each triangle at its
// canonical cell; it keeps the nearest triangle (the first of equal distance)
// whose signed distance is under the contact margin, or, inside-only, more
// than 1 mm behind the face.  The particle part keeps the deepest particle.
// Both push along the depth-weighted mean of all their contacts.  These are
// the cloth's own colliders: no source filter.
__device__ inline void signed_contacts(const ClothView& c, VertexState& s, bool inside, bool count) {
    const SSBL_CONST DynView& D = *(const SSBL_CONST DynView*)(c.sdyn);
    const float dt = c.dt_sub, cdt = c.collider_dt;
    const float cm = fmaxf(fmaxf(c.collision_margin, 0.0f), fmaxf(c.thickness, 0.0f) * kSignedContactMarginScale);
    const float qm = fminf(fmaxf(fmaxf(fmaxf(cm * 2.0f, kSignedQueryMargin), fmaxf(D.tri_cell, 0.0f) * 1.25f), cm), kSignedMaxQueryMargin);
    const int nt = D.nt;
    if (nt > 0 && cm > 0.0f && qm > 0.0f) {
        const F3 p = s.p;
        F3 best_delta = f3(0.0f, 0.0f, 0.0f), best_vel = f3(0.0f, 0.0f, 0.0f), acc = f3(0.0f, 0.0f, 0.0f);
        float best_pen = 0.0f, best_dist = 3.402823466e+38f, accw = 0.0f;
        unsigned cand = 0, rej = 0;
        auto consider = [&](int t, const TriBox& bx) {
            if (t < 0 || t >= nt) return;
            ++cand;
            const float4 mn = box_lo(bx), mx = box_hi(bx);
            const F3 q0 = f3(fminf(fmaxf(p.x, mn.x), mx.x), fminf(fmaxf(p.y, mn.y), mx.y), fminf(fmaxf(p.z, mn.z), mx.z));
            const F3 da = p - q0;
            if (dot(da, da) > qm * qm) { ++rej; return; }
            const float4 ce = bx.s;
            const float rr = ce.w + qm;
            const F3 dc = p - f3(ce);
            if (dot(dc, dc) > rr * rr) { ++rej; return; }
            const F3 a = f3(D.tv[3 * t]), b = f3(D.tv[3 * t + 1]), cc = f3(D.tv[3 * t + 2]);
            const F3 q = closest_point_on_triangle(p, a, b, cc);
            const F3 d = p - q;
            const float dsq = dot(d, d);
            const float dist = sqrt_eps(dsq);
            const F3 tn = cross(b - a, cc - a);
            const float tl = sqrt_eps(dot(tn, tn));
            const F3 tnorm = mul(tn, 1.0f / tl);
            const float sd = dot(d, tnorm);
            float pen;
            if (inside) {
                const float target = -kSignedInsideTargetMargin;
                if (sd >= target) return;
                pen = target - sd;
            } else {
                pen = cm - sd;
            }
            if (pen <= 0.0f) return;
            const F3 delta = mul(tnorm, pen);
            if (!inside && finite3(delta)) {
                const float w = fmaxf(pen, kEps);
                acc = acc + mul(delta, w);
                accw += w;
            }
            if (dist < best_dist) {
                F3 cv = f3(0.0f, 0.0f, 0.0f);
                if (D.tri_motion && cdt > kEps) cv = mul(f3(ce) - f3(D.pctr[t]), 1.0f / cdt);
                best_dist = dist;
                best_pen = pen;
                best_delta = delta;
                best_vel = cv;
            }
        };
        if (D.tri_hash) {
            const float cs = fmaxf(D.tri_cell, 1.0e-3f);
            const int radius = inside ? min(max(static_cast<int>(ceilf(fmaxf(qm, 0.0f) / cs)), 1), 4) : 1;
            const int cx = cell_coord(p.x, D.tri_cell), cy = cell_coord(p.y, D.tri_cell), cz = cell_coord(p.z, D.tri_cell);
            for (int dz = -radius; dz <= radius; ++dz)
                for (int dy = -radius; dy <= radius; ++dy)
                    for (int dx = -radius; dx <= radius; ++dx) {
                        const int qx = cx + dx, qy = cy + dy, qz = cz + dz;
                        const int4 bin = D.tri_bin[hash_cell(qx, qy, qz, D.tri_buckets)];
                        for (int k = bin.x; k < bin.x + bin.y; ++k) {
                            const int t = D.tri_keys[k] / kDynMaxHashCells;
                            if (t < 0 || t >= nt) continue;
                            // A41:3934-3982: the triangle's canonical cell for this vertex
                            const TriBox bx = D.box[t];
                            const int4 lo = box_cmin(bx), hi = box_cmax(bx);
                            if (qx != min(max(cx, lo.x), hi.x) || qy != min(max(cy, lo.y), hi.y) || qz != min(max(cz, lo.z), hi.z)) continue;
                            consider(t, bx);
                        }
                    }
            const int nl = *D.nlarge;
            for (int k = 0; k < nl; ++k) {
                const int t = D.large[k];
                if (t >= 0 && t < nt) consider(t, D.box[t]);
            }
        } else {
            for (int t = 0; t < nt; ++t) consider(t, D.box[t]);
        }
        if (count && cand) atomicAdd(&c.counts[kCountDynTriCandidates], static_cast<unsigned long long>(cand));
        if (count && rej) atomicAdd(&c.counts[kCountDynTriAabbRejects], static_cast<unsigned long long>(rej));
        if (best_pen > 0.0f) {
            if (count) atomicAdd(&c.counts[kCountTrianglePairs], 1ull);
            signed_response(c, s, p, local_average_delta(acc, accw, best_delta), best_vel, best_pen, dt, inside, count);
            if (count) atomicAdd(&c.counts[kCountExactImpulse], 1ull);
        }
    }
    const int np_count = D.np;
    if (inside || np_count <= 0) return;
    const float margin = fmaxf(c.collision_margin, 0.0f);
    const F3 p = s.p;
    F3 best_delta = f3(0.0f, 0.0f, 0.0f), best_vel = f3(0.0f, 0.0f, 0.0f), acc = f3(0.0f, 0.0f, 0.0f);
    float best_pen = 0.0f, accw = 0.0f;
    unsigned cand = 0;
    auto consider = [&](int j) {
        if (j < 0 || j >= np_count) return;
        const float4 pj = D.pp[j];
        const float radius = fmaxf(pj.w, 0.0f) + margin;
        if (radius <= 0.0f) return;
        const F3 d = p - f3(pj);
        const float dsq = dot(d, d);
        const float dist = sqrt_eps(dsq);
        if (dist >= radius) return;
        ++cand;
        const F3 n = dist > kEps ? mul(d, 1.0f / dist) : f3(0.0f, 0.0f, 1.0f);
        const float pen = radius - dist;
        const F3 delta = mul(n, pen);
        if (finite3(delta)) {
            const float w = fmaxf(pen, kEps);
            acc = acc + mul(delta, w);
            accw += w;
        }
        if (pen > best_pen) {
            F3 pv = f3(0.0f, 0.0f, 0.0f);
            if (D.part_motion && cdt > kEps) pv = mul(f3(pj) - f3(D.pprev[j]), 1.0f / cdt);
            best_pen = pen;
            best_delta = delta;
            best_vel = pv;
        }
    };
    if (D.part_hash) {
        const int cx = cell_coord(p.x, D.part_cell), cy = cell_coord(p.y, D.part_cell), cz = cell_coord(p.z, D.part_cell);
        for (int k27 = 0; k27 < 27; ++k27) {   // z, then y, then x innermost
            const int qx = cx + k27 % 3 - 1, qy = cy + (k27 / 3) % 3 - 1, qz = cz + k27 / 9 - 1;
            const int4 bin = gload(D.pbin + hash_cell(qx, qy, qz, D.part_buckets));
            const int lim = min(bin.y, kDynParticleSlots);
            for (int k = 0; k < lim; ++k) {
                const int4 ck = gload(&D.pent[bin.x + k].ck);
                if (ck.x != qx || ck.y != qy || ck.z != qz) continue;
                consider(ck.w);
            }
        }
    } else {
        for (int j = 0; j < np_count; ++j) consider(j);
    }
    if (count && cand) atomicAdd(&c.counts[kCountDynParticleCandidates], static_cast<unsigned long long>(cand));
    if (best_pen > 0.0f) {
        signed_response(c, s, p, local_average_delta(acc, accw, best_delta), best_vel, best_pen, dt, false, count);
        if (count) {
            atomicAdd(&c.counts[kCountExactImpulse], 1ull);
            atomicAdd(&c.counts[kCountDynParticleContacts], 1ull);
        }
    }
}

template <bool kInside>
__global__ __launch_bounds__(kTile) void k_signed_pass(const ClothView* __restrict__ views, const Tile* __restrict__ tiles, int ntiles,
                                                       int parity, unsigned ops) {
    const Tile tile = tiles[xcd_tile(ntiles)];
    const int v = tile.begin + static_cast<int>(threadIdx.x);
    if (v >= tile.end) return;
    const ClothView& c = views[tile.cloth];
    if (!c.sdyn) return;
    float4* pos = parity ? c.pos1 : c.pos0;
    const float4 p4 = pos[v];
    if (p4.w == 0.0f) return;
    VertexState s{&c, v, f3(p4), f3(0.0f, 0.0f, 0.0f), f3(0.0f, 0.0f, 0.0f), false, false, false, false};
    signed_contacts(c, s, kInside, (ops & kOpCount) != 0u);
    if (s.p.x != p4.x || s.p.y != p4.y || s.p.z != p4.z) pos[v] = with_w(s.p, p4.w);
    if (s.q_dirty) c.prev[v] = with_w(s.q_, 0.0f);
    if (s.v_dirty) c.vel[v] = with_w(s.vel_, 0.0f);
}

// ---------------------------------------------------------- vertex pass
// Neighbour positions loaded together before any of their edges is
// evaluated: one round of memory latency per chunk instead of one per edge.
#ifndef SSBL_GATHER_CHUNK
#define SSBL_GATHER_CHUNK 4
#endif
constexpr int kGatherChunk = SSBL_GATHER_CHUNK;
static_assert(kEllMax % kGatherChunk == 0, "the gather chunk must divide the slot count");

// Element i of a wave-uniform array, addressed as base + 32-bit byte offset
// (a scalar base register and one vector register per load in flight).
template <typename T>
__device__ __forceinline__ T load_at(const T* __restrict__ base, unsigned i) {
    return gload(reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * static_cast<unsigned>(sizeof(T))));
}

// A41:1927-1980, one edge as seen from vertex s.v: the edge thread's
// correction, with the difference vector negated for the second endpoint
// (which flips the correction's sign bit-exactly).
__device__ __forceinline__ void stretch_edge(int2 a, float4 o4, F3 p0, float w, float alpha, F3& p, F3& q, bool& q_moved,
                                             unsigned& clamps) {
    const float restl = __int_as_float(a.y);
    const float wsum = w + o4.w;
    if (wsum <= 0.0f) return;
    const F3 delta = f3(o4) - p0;
    // len >= sqrt(kEps) > kEps, so max(len, eps) is len itself
    const float len = sqrt_eps(dot(delta, delta));
    const float cval = len - restl;
    const float dl = -cval / (wsum + alpha);
    F3 corr = mul(delta, 0.18f * dl / len);
    const float csq = fmaxf(dot(corr, corr), 0.0f);
    const float mc = fmaxf(0.0025f, fminf(fmaxf(restl, 0.0f) * 0.10f, 0.025f));
    if (sqrt_greater(csq, mc)) {
        corr = mul(corr, mc / fmaxf(sqrtf(csq), kEps));
        if (a.x < 0) ++clamps;
    }
    if (w > 0.0f) {
        const F3 d = mul(corr, -w);
        p = p + d;
        if (cval > 0.0f) { q = q + mul(d, 0.08f); q_moved = true; }
    }
}

// The first ell_width incidences of vertex v.  All kEllMax slots are loaded
// (those past the width load the last real slot again, from cache), so the
// loads issue back to back with no branch between them; the gather ignores
// slots at or past the width.  Requires ell_width > 0.
__device__ __forceinline__ void ell_load(const ClothView& c, int v, int2 (&a)[kEllMax]) {
    const int last = c.ell_width - 1;
    if (c.ell16) {
        // 6-byte slots, expanded to the 8-byte form
        short d[kEllMax];
        float r[kEllMax];
#pragma unroll
        for (int j = 0; j < kEllMax; ++j) {
            const unsigned slot = static_cast<unsigned>(min(j, last) * c.ell_pitch + v);
            d[j] = load_at(c.ell16, slot);
            r[j] = load_at(c.ell_rest, slot);
        }
        // an unused slot holds code 0 and rest length 0, so the rest length
        // is taken as loaded (no load waits on its slot's code)
#pragma unroll
        for (int j = 0; j < kEllMax; ++j) {
            const int code = d[j];
            const unsigned other = static_cast<unsigned>(v + (code >> 1)) | ((code & 1) ? 0x80000000u : 0u);
            a[j] = make_int2(code == 0 ? -1 : static_cast<int>(other), __float_as_int(r[j]));
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < kEllMax; ++j) a[j] = load_at(c.ell, static_cast<unsigned>(min(j, last) * c.ell_pitch + v));
}

// Neighbour positions of a gather: from the pass's source buffer in HBM, or
// from the workgroup's LDS copy of it (StagedSource, indices lo .. lo + n).
struct GlobalSource {
    const float4* src;
    __device__ __forceinline__ float4 operator()(unsigned i) const { return load_at(src, i); }
};
struct StagedSource {
    const float4* stage;
    int lo;
    __device__ __forceinline__ float4 operator()(unsigned i) const { return stage[static_cast<int>(i) - lo]; }
};

// Copies src[lo, lo + n) (n <= kStageMax) to stage[0, n).  Each wave
// instruction moves 64 consecutive positions from global memory straight
// into LDS, with no registers in between; lanes past n load position n - 1
// again into slots no gather reads.
constexpr int kStageMax = kTile + 2 * kStageHaloMax;
__device__ __forceinline__ void stage_positions(const float4* __restrict__ src, int lo, int n, float4* stage) {
    const int lane = static_cast<int>(threadIdx.x) & 63;
    for (int ch = static_cast<int>(threadIdx.x) >> 6; ch * 64 < n; ch += kTile / 64) {
        const int i = min(ch * 64 + lane, n - 1);
        __builtin_amdgcn_global_load_lds(src + lo + i, (__attribute__((address_space(3))) void*)(stage + ch * 64),
                                         16, 0, 0);
    }
}

// Direct stretch gathered per vertex, incident edges in ascending edge order:
// the slot-major first ell_width incidences (a, loaded by ell_load), then any
// overflow from adj (read from src in HBM).
template <typename Source>
__device__ inline void stretch_gather(const ClothView& c, Source fetch, const float4* __restrict__ src, VertexState& s,
                                      float w, const int2 (&a)[kEllMax], bool count) {
    const int v = s.v;
    const int width = c.ell_width;
    const float alpha = c.stretch_alpha;
    const F3 p0 = s.p;
    F3 p = p0;
    F3 q = w > 0.0f ? s.q() : p0;
    bool q_moved = false;
    unsigned clamps = 0;
#pragma unroll
    for (int j0 = 0; j0 < kEllMax; j0 += kGatherChunk) {
        if (j0 >= width) break;
        float4 o[kGatherChunk];
        bool live[kGatherChunk];
#pragma unroll
        for (int j = 0; j < kGatherChunk; ++j) {
            const int x = a[j0 + j].x;
            live[j] = j0 + j < width && x != -1;
            o[j] = fetch(static_cast<unsigned>(live[j] ? (x & 0x7fffffff) : v));
        }
#pragma unroll
        for (int j = 0; j < kGatherChunk; ++j)
            if (live[j]) stretch_edge(a[j0 + j], o[j], p0, w, alpha, p, q, q_moved, clamps);
    }
    if (c.ell_overflow) {
        const int e = c.adj_off[v + 1];
        for (int k = c.adj_off[v] + width; k < e; ++k) {
            const int2 ak = c.adj[k];
            stretch_edge(ak, src[ak.x & 0x7fffffff], p0, w, alpha, p, q, q_moved, clamps);
        }
    }
    s.p = p;
    if (q_moved) s.set_q(q);
    if (count && clamps) atomicAdd(&c.counts[kCountHardFallbacks], static_cast<unsigned long long>(clamps));
}

// A41:2710-2803 gathered the same way (tethers are one-sided: only a
// stretched tether acts).
__device__ inline void lra_gather(const ClothView& c, const float4* __restrict__ src, VertexState& s, float w, float dt,
                                  bool count) {
    const int b = c.lra_off[s.v], e = c.lra_off[s.v + 1];
    const F3 p0 = s.p;
    unsigned tacks = 0, guards = 0;
    for (int k = b; k < e; ++k) {
        const int4 a = c.lra_adj[k];
        const int other = a.x & 0x7fffffff;
        const bool first = a.x < 0;
        const float restl = __int_as_float(a.y);
        const float4 o4 = src[other];
        const float wsum = w + o4.w;
        if (!isfinite(wsum) || wsum <= 0.0f) { guards += first; continue; }
        const F3 po = f3(o4);
        if (!finite3(p0) || !finite3(po)) { guards += first; continue; }
        F3 dx = po - p0;
        const float lsq = __builtin_fmaf(dx.x, dx.x, __builtin_fmaf(dx.y, dx.y, dx.z * dx.z));
        if (!isfinite(lsq) || lsq <= 1.0e-10f || !isfinite(restl) || restl <= kEps) { guards += first; continue; }
        const float inv_len = rsqrt_exact(lsq);
        const float len = lsq * inv_len;
        const float stretch = len - restl;
        if (!isfinite(stretch) || stretch <= 0.0f) continue;
        dx = mul(dx, inv_len);
        const float bc = fmaxf(c.lra_compliance, 1.0e-12f);
        const float stiff = (1.0f / bc) * 500000.0f;
        const float ealpha = (bc / 500000.0f) / fmaxf(dt * dt, kEps);
        const float sw = stiff / fmaxf(stiff + 500000.0f, 1.0f);
        const float dl = -stretch / fmaxf(wsum + ealpha, kEps);
        F3 corr = mul(dx, 0.35f * sw * dl);
        if (!finite3(corr)) { guards += first; continue; }
        const float cl = len3(corr);
        const float mc = fmaxf(5.0e-4f, fminf(fmaxf(restl, c.thickness) * 0.20f, 0.035f));
        if (!isfinite(cl) || cl <= 0.0f) continue;
        if (cl > mc) { corr = mul(corr, mc / fmaxf(cl, kEps)); guards += first; }
        const bool th = c.thickness > 0.0f;
        const float cap = th ? 0.0025f : 0.012f;
        const float rl = fmaxf(restl, 0.0f) * (th ? 0.02f : 0.08f);
        const float tl = th ? c.thickness * 0.12f : cap;
        const float al = fmaxf(5.0e-4f, fminf(fminf(rl, tl), cap));
        if (w > 0.0f) {
            const F3 d = limit_delta(mul(corr, -w / fmaxf(wsum, kEps)), al);
            s.p = s.p + d;
            s.set_q(s.q() + mul(d, 0.32f));
        }
        tacks += first;
    }
    if (count && tacks) atomicAdd(&c.counts[kCountLraTacks], static_cast<unsigned long long>(tacks));
    if (count && guards) atomicAdd(&c.counts[kCountTackGuards], static_cast<unsigned long long>(guards));
}

// A41:2805-2884 (uncoloured bends) gathered like the stretch pass.
__device__ inline void bend_gather(const ClothView& c, const float4* __restrict__ src, VertexState& s, float w,
                                   bool count) {
    const int b = c.bend_off[s.v], e = c.bend_off[s.v + 1];
    const F3 p0 = s.p;
    unsigned wings = 0, guards = 0;
    for (int k = b; k < e; ++k) {
        const int4 a = c.bend_adj[k];
        const int other = a.x & 0x7fffffff;
        const bool first = a.x < 0;
        const float restl = __int_as_float(a.y);
        const float4 o4 = src[other];
        const float wsum = w + o4.w;
        if (!isfinite(wsum) || wsum <= 0.0f) continue;
        if (!isfinite(restl) || restl <= kEps) { guards += first; continue; }
        const F3 po = f3(o4);
        if (!finite3(p0) || !finite3(po)) { guards += first; continue; }
        const F3 delta = po - p0;
        const float lsq = __builtin_fmaf(delta.x, delta.x, __builtin_fmaf(delta.y, delta.y, delta.z * delta.z));
        if (!isfinite(lsq) || lsq <= 1.0e-10f) { guards += first; continue; }
        const float inv_len = rsqrt_exact(lsq);
        const float len = lsq * inv_len;
        const float cval = len - restl;
        if (!isfinite(cval) || fabsf(cval) <= 1.0e-6f) continue;
        const float dl = -cval / fmaxf(wsum + c.bend_alpha, kEps);
        F3 corr = mul(delta, 0.10f * dl * inv_len);
        if (!finite3(corr)) { guards += first; continue; }
        const float cl = len3(corr);
        const float mc = fmaxf(2.5e-4f, fminf(fmaxf(restl, c.thickness) * 0.08f, 0.020f));
        if (!isfinite(cl) || cl <= 0.0f) continue;
        if (cl > mc) { corr = mul(corr, mc / fmaxf(cl, kEps)); guards += first; }
        if (w > 0.0f) {
            const F3 d = mul(corr, -w);
            s.p = s.p + d;
            s.set_q(s.q() + mul(d, 0.06f));
        }
        wings += first;
    }
    if (count && wings) atomicAdd(&c.counts[kCountBendingWings], static_cast<unsigned long long>(wings));
    if (count && guards) atomicAdd(&c.counts[kCountBendingGuards], static_cast<unsigned long long>(guards));
}

enum GatherKind { kGatherNone = 0, kGatherStretch = 1, kGatherLra = 2, kGatherBend = 3 };

// The vertex-local steps of a pass after its constraint gather, then the
// vertex's writes.
__device__ __forceinline__ void vertex_tail(const ClothView& c, VertexState& s, float w, bool pinned, float dt, float4* dst,
                                            unsigned ops, bool count, const VertexPins& pins) {
    const bool had_pins = pins.b < pins.e;
    if (ops & kOpPin) pins.apply(c, s, pinned, dt);
    if ((ops & kOpAnalytic) && !pinned) analytic_contacts(c, s, dt, count);
    if ((ops & kOpSdf) && c.sdf_ready && !pinned) sdf_contact(c, s, dt, count);
    // A41:9360-9369: the second pin projection runs after any contact work,
    // and on every last iteration (dynamic collision is scheduled there).
    if ((ops & kOpPinPolish) && ((ops & kOpLastIter) || c.contacts_on))
        pins.apply_after_contacts(c, s, pinned, dt, had_pins, (ops & kOpPin) != 0u);
    if (ops & kOpVelocity) {
        if (pinned) s.set_vel(f3(0.0f, 0.0f, 0.0f));
        else velocity_update(c, s, dt);
    }
    gstore(dst + s.v, with_w(s.p, w));
    if (s.q_dirty) gstore(c.prev + s.v, with_w(s.q_, 0.0f));
    if (s.v_dirty) gstore(c.vel + s.v, with_w(s.vel_, 0.0f));
}

// The stretch pass of a cloth with stage_halo > 0.  The workgroup copies the
// source positions of its tile and of stage_halo indices either side into
// LDS while each vertex's incidence slots and previous position load; after
// one barrier the vertex's own position and every ELL neighbour come from
// LDS, so the pass makes one round trip to memory before its arithmetic.
// Same arithmetic, same order as the gather from HBM.
__device__ inline void vertex_pass_staged(const ClothView& c, const Tile& tile, const float4* __restrict__ src,
                                          float4* __restrict__ dst, unsigned ops, float4* stage) {
    const int v = tile.begin + static_cast<int>(threadIdx.x);
    const bool live = v < tile.end;
    const int lo = max(tile.begin - c.stage_halo, 0);
    const int hi = min(tile.end + c.stage_halo, c.nv);
    stage_positions(src, lo, hi - lo, stage);
    const int vl = live ? v : tile.begin;
    const F3 q = gload_xyz(c.prev + vl);
    VertexPins pins;
    pins.range(c, vl, ops);
    int2 ell[kEllMax];
    ell_load(c, vl, ell);
    __syncthreads();
    if (!live) return;
    const bool count = (ops & kOpCount) != 0u;
    const float dt = c.dt_sub;
    const float4 p4 = stage[v - lo];
    const float w = p4.w;
    const bool pinned = w == 0.0f;
    VertexState s{&c, v, f3(p4), q, f3(0.0f, 0.0f, 0.0f), true, false, false, false};
    stretch_gather(c, StagedSource{stage, lo}, src, s, w, ell, count);
    vertex_tail(c, s, w, pinned, dt, dst, ops, count, pins);
}

// One solver iteration for every vertex of the tile: the constraint gather
// selected at compile time, then the vertex-local steps selected by `ops`.
template <int kGather>
__global__ __launch_bounds__(kTile, 8) void k_vertex_pass(const ClothView* __restrict__ views, const Tile* __restrict__ tiles,
                                                       int ntiles, int src_parity, int dst_parity, unsigned ops) {
    const Tile tile = tiles[xcd_tile(ntiles)];
    const ClothView& c = views[tile.cloth];
    const float4* src = src_parity ? c.pos1 : c.pos0;
    float4* dst = dst_parity ? c.pos1 : c.pos0;
    if constexpr (kGather == kGatherStretch) {
        // uniform over the workgroup: every thread reaches the barrier
        if (c.direct_stretch != 0 && c.stage_halo > 0) {
            __shared__ float4 stage[kStageMax];
            vertex_pass_staged(c, tile, src, dst, ops, stage);
            return;
        }
    }
    const int v = tile.begin + static_cast<int>(threadIdx.x);
    if (v >= tile.end) return;
    const float dt = c.dt_sub;
    const bool count = (ops & kOpCount) != 0u;
    // Everything the stretch gather reads about the vertex itself (its
    // incidences, position and previous position) is requested before any of
    // it is used.
    bool gather = false;
    int2 ell[kEllMax];
    F3 q = f3(0.0f, 0.0f, 0.0f);
    const float4 p4 = gload(src + v);
    VertexPins pins;
    pins.range(c, v, ops);
    if constexpr (kGather == kGatherStretch) {
        gather = c.direct_stretch != 0;
        if (gather) {
            q = gload_xyz(c.prev + v);
            if (c.ell_width > 0) {
                ell_load(c, v, ell);
            } else {
#pragma unroll
                for (int j = 0; j < kEllMax; ++j) ell[j] = make_int2(-1, 0);   // never read: width 0
            }
        }
    }
    const float w = p4.w;
    // .w holds 0 exactly for the vertices the reference flags as pinned
    // (inverse mass <= 0, A41:8861-8865)
    const bool pinned = w == 0.0f;
    VertexState s{&c, v, f3(p4), f3(0.0f, 0.0f, 0.0f), f3(0.0f, 0.0f, 0.0f), false, false, false, false};
    if constexpr (kGather == kGatherStretch) {
        if (gather) {
            s.q_ = q;
            s.q_have = true;
            stretch_gather(c, GlobalSource{src}, src, s, w, ell, count);
        }
    } else if constexpr (kGather == kGatherLra) {
        lra_gather(c, src, s, w, dt, count);
    } else if constexpr (kGather == kGatherBend) {
        bend_gather(c, src, s, w, count);
    }
    vertex_tail(c, s, w, pinned, dt, dst, ops, count, pins);
}

// ------------------------------------------------------ coloured stretch
// Graph-coloured Gauss-Seidel form of the direct stretch projection
// (A41:1927-1980): one launch per edge colour over the colour's segments of
// every cloth ({cloth, begin, end}, 256 edges each), one thread per edge,
// positions and previous positions updated in place -- race free because no
// vertex appears twice in a colour, so the result equals a sequential sweep
// of the colour.  Each edge is evaluated once for both endpoints with the
// arithmetic of stretch_edge; the record {i, j, rest bits, 0} is one 16-byte
// load.  Consecutive segments (neighbouring edges of one cloth) go to the
// same XCD.
__global__ __launch_bounds__(kTile) void k_stretch_color(const ClothView* __restrict__ views, const int4* __restrict__ segs,
                                                         int nsegs, int parity, unsigned ops) {
    const int4 seg = segs[xcd_tile(nsegs)];
    const int k = seg.y + static_cast<int>(threadIdx.x);
    if (k >= seg.z) return;
    const ClothView& c = views[seg.x];
    float4* pos = parity ? c.pos1 : c.pos0;
    const int4 e = c.sedges[k];
    const int i = e.x, j = e.y;
    if (i >= c.nv || j >= c.nv) return;
    const float4 pi4 = pos[i], pj4 = pos[j];
    const float wi = pi4.w, wj = pj4.w;
    const float w = wi + wj;
    if (w <= 0.0f) return;
    const F3 pi = f3(pi4), pj = f3(pj4);
    const F3 delta = pj - pi;
    const float len = sqrt_eps(dot(delta, delta));
    const float restl = __int_as_float(e.z);
    const float cval = len - restl;
    const float dl = -cval / (w + c.stretch_alpha);
    F3 corr = mul(delta, 0.18f * dl / len);
    const float csq = fmaxf(dot(corr, corr), 0.0f);
    const float mc = fmaxf(0.0025f, fminf(fmaxf(restl, 0.0f) * 0.10f, 0.025f));
    const bool clamped = sqrt_greater(csq, mc);
    if (clamped) corr = mul(corr, mc / fmaxf(sqrtf(csq), kEps));
    if (wi > 0.0f) {
        const F3 d = mul(corr, -wi);
        pos[i] = with_w(pi + d, wi);
        if (cval > 0.0f) c.prev[i] = with_w(f3(c.prev[i]) + mul(d, 0.08f), 0.0f);
    }
    if (wj > 0.0f) {
        const F3 d = mul(corr, wj);
        pos[j] = with_w(pj + d, wj);
        if (cval > 0.0f) c.prev[j] = with_w(f3(c.prev[j]) + mul(d, 0.08f), 0.0f);
    }
    if (ops & kOpCount) wave_count(c.counts, kCountHardFallbacks, clamped);
}

// -------------------------------------------------------------- bending
// A41:2886-2970, one launch per colour; a segment is {cloth, begin, end}.
__global__ __launch_bounds__(kTile) void k_bend_color(const ClothView* __restrict__ views, const int4* __restrict__ segs,
                                                      int parity, unsigned ops) {
    const int4 seg = segs[blockIdx.x];
    const int k = seg.y + static_cast<int>(threadIdx.x);
    if (k >= seg.z) return;
    const ClothView& c = views[seg.x];
    float4* pos = parity ? c.pos1 : c.pos0;
    const bool count = (ops & kOpCount) != 0u;
    const int2 pr = c.bends[k];
    const int i = pr.x, j = pr.y;
    if (i < 0 || j < 0 || i >= c.nv || j >= c.nv || i == j) { if (count) atomicAdd(&c.counts[kCountBendingGuards], 1ull); return; }
    const float4 pi4 = pos[i], pj4 = pos[j];
    const float wi = pi4.w, wj = pj4.w;
    const float w = wi + wj;
    if (!isfinite(w) || w <= 0.0f) return;
    const float restl = c.bend_rest[k];
    if (!isfinite(restl) || restl <= kEps) { if (count) atomicAdd(&c.counts[kCountBendingGuards], 1ull); return; }
    const F3 pi = f3(pi4), pj = f3(pj4);
    if (!finite3(pi) || !finite3(pj)) { if (count) atomicAdd(&c.counts[kCountBendingGuards], 1ull); return; }
    const F3 delta = pj - pi;
    const float lsq = __builtin_fmaf(delta.x, delta.x, __builtin_fmaf(delta.y, delta.y, delta.z * delta.z));
    if (!isfinite(lsq) || lsq <= 1.0e-10f) { if (count) atomicAdd(&c.counts[kCountBendingGuards], 1ull); return; }
    const float inv_len = rsqrt_exact(lsq);
    const float len = lsq * inv_len;
    const float cval = len - restl;
    if (!isfinite(cval) || fabsf(cval) <= 1.0e-6f) return;
    const float dl = -cval / fmaxf(w + c.bend_alpha, kEps);
    F3 corr = mul(delta, 0.10f * dl * inv_len);
    if (!finite3(corr)) { if (count) atomicAdd(&c.counts[kCountBendingGuards], 1ull); return; }
    const float cl = len3(corr);
    const float mc = fmaxf(2.5e-4f, fminf(fmaxf(restl, c.thickness) * 0.08f, 0.020f));
    if (!isfinite(cl) || cl <= 0.0f) return;
    if (cl > mc) { corr = mul(corr, mc / fmaxf(cl, kEps)); if (count) atomicAdd(&c.counts[kCountBendingGuards], 1ull); }
    if (wi > 0.0f) {
        const F3 d = mul(corr, -wi);
        pos[i] = with_w(pi + d, wi);
        const float4 q = c.prev[i];
        c.prev[i] = with_w(f3(q) + mul(d, 0.06f), 0.0f);
    }
    if (wj > 0.0f) {
        const F3 d = mul(corr, wj);
        pos[j] = with_w(pj + d, wj);
        const float4 q = c.prev[j];
        c.prev[j] = with_w(f3(q) + mul(d, 0.06f), 0.0f);
    }
    if (count) wave_count(c.counts, kCountBendingWings, true);
}

// ---------------------------------------------------------- SDF build
// A41:1725-1747 +x ray against a triangle (Moller-Trumbore), for the sign.
__device__ inline bool ray_hits(F3 p, F3 a, F3 b, F3 c) {
    const F3 dir = f3(1.0f, 0.0f, 0.0f);
    const F3 e1 = b - a, e2 = c - a;
    const F3 h = cross(dir, e2);
    const float det = dot(e1, h);
    if (fabsf(det) <= 1.0e-8f) return false;
    const float inv = 1.0f / det;
    const F3 s = p - a;
    const float u = inv * dot(s, h);
    if (u < -1.0e-5f || u > 1.0f + 1.0e-5f) return false;
    const F3 q = cross(s, e1);
    const float vv = inv * dot(dir, q);
    if (vv < -1.0e-5f || u + vv > 1.0f + 1.0e-5f) return false;
    return inv * dot(e2, q) > 1.0e-6f;
}

// A41:1787-1829: brute-force distance of every voxel to the static triangles,
// triangles staged through LDS in chunks.
__global__ __launch_bounds__(kTile) void k_sdf_build(float* __restrict__ sdf, int dx, int dy, int dz, float ox, float oy,
                                                     float oz, float voxel, const float* __restrict__ tris, int ntris,
                                                     int signed_sdf) {
    __shared__ float chunk[kTile * 9];
    const long long vox = static_cast<long long>(blockIdx.x) * kTile + threadIdx.x;
    const long long total = static_cast<long long>(dx) * dy * dz;
    const bool active = vox < total;
    F3 p = f3(0.0f, 0.0f, 0.0f), rp = p;
    if (active) {
        const int x = static_cast<int>(vox % dx), y = static_cast<int>((vox / dx) % dy), z = static_cast<int>(vox / (1LL * dx * dy));
        p = f3(ox, oy, oz) + f3(static_cast<float>(x) * voxel, static_cast<float>(y) * voxel, static_cast<float>(z) * voxel);
        rp = p + f3(0.0f, voxel * 1.37e-4f, voxel * 2.11e-4f);
    }
    float best = 3.402823466e+38f;
    int parity = 0;
    for (int base = 0; base < ntris; base += kTile) {
        const int n = min(kTile, ntris - base);
        __syncthreads();
        for (int k = threadIdx.x; k < n * 9; k += kTile) chunk[k] = tris[static_cast<size_t>(base) * 9 + k];
        __syncthreads();
        if (active) {
            for (int t = 0; t < n; ++t) {
                const float* T = chunk + 9 * t;
                const F3 a = f3(T[0], T[1], T[2]), b = f3(T[3], T[4], T[5]), c = f3(T[6], T[7], T[8]);
                const F3 d = p - closest_point_on_triangle(p, a, b, c);
                best = fminf(best, dot(d, d));
                if (signed_sdf && ray_hits(rp, a, b, c)) parity ^= 1;
            }
        }
    }
    if (active) {
        float dist = sqrtf(fmaxf(best, 0.0f));
        if (signed_sdf && parity) dist = -dist;
        sdf[vox] = dist;
    }
}

__global__ void k_fill_u32(unsigned* p, unsigned value, size_t n) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// ------------------------------------------------------------ launches
void launch_integrate(const LaunchCtx& c, int parity, unsigned ops, bool full) {
    TimedLaunch t("integrate", c.stream);
    if (full)
        launch_k(k_integrate<true>, dim3(c.ntiles), dim3(kTile), 0, c.stream, c.views, c.tiles, c.ntiles, parity, ops);
    else
        launch_k(k_integrate<false>, dim3(c.ntiles), dim3(kTile), 0, c.stream, c.views, c.tiles, c.ntiles, parity, ops);
}

void launch_vertex_pass(const LaunchCtx& c, int src_parity, int dst_parity, unsigned ops) {
    const char* name = (ops & kOpStretch) ? "vertex_pass_stretch"
                       : (ops & kOpLra) ? "vertex_pass_lra"
                       : (ops & kOpBend) ? "vertex_pass_bend" : "vertex_pass_local";
    TimedLaunch t(name, c.stream);
    const dim3 grid(c.ntiles), block(kTile);
    if (ops & kOpStretch)
        launch_k(k_vertex_pass<kGatherStretch>, grid, block, 0, c.stream, c.views, c.tiles, c.ntiles, src_parity, dst_parity, ops);
    else if (ops & kOpLra)
        launch_k(k_vertex_pass<kGatherLra>, grid, block, 0, c.stream, c.views, c.tiles, c.ntiles, src_parity, dst_parity, ops);
    else if (ops & kOpBend)
        launch_k(k_vertex_pass<kGatherBend>, grid, block, 0, c.stream, c.views, c.tiles, c.ntiles, src_parity, dst_parity, ops);
    else
        launch_k(k_vertex_pass<kGatherNone>, grid, block, 0, c.stream, c.views, c.tiles, c.ntiles, src_parity, dst_parity, ops);
}

void launch_order_tiles(const LaunchCtx& c) {
    if (!c.tile_order || !c.tile_work || c.ntiles <= 0) return;
    TimedLaunch t("dynamic_order", c.stream);
    launch_k(k_order_tiles, dim3(1), dim3(1024), 0, c.stream, c.ntiles, c.tile_work, c.tile_order);
}

void launch_dynamic_pass(const LaunchCtx& c, int parity, unsigned ops, unsigned tail_ops, int pool_max, hipEvent_t begin,
                         hipEvent_t end) {
    const unsigned pops = ops | (static_cast<unsigned>(std::min(std::max(pool_max, 0), 64)) << kOpPoolShift);
    if (begin) hipEventRecord(begin, c.stream);
    {
        // the units (launch_order_tiles) fit 2 * ntiles workgroups
        TimedLaunch t("dynamic_triangles", c.stream);
        const int* units = c.tile_order + 2 * c.ntiles + 1;
        launch_k(k_dynamic_triangles, dim3(2 * c.ntiles), dim3(kTile), 0, c.stream, c.views, c.tiles, parity, ops, units,
                 static_cast<const int*>(c.tile_order + 2 * c.ntiles), c.tile_work + c.ntiles);
    }
    {
        TimedLaunch t("dynamic_contacts", c.stream);
        launch_k(k_dynamic_pass, dim3(c.ntiles), dim3(kTile), 0, c.stream, c.views, c.tiles, c.ntiles, parity, pops, tail_ops,
                 static_cast<const int*>(c.tile_order), c.tile_work);
    }
    if (end) hipEventRecord(end, c.stream);
}

void launch_signed_pass(const LaunchCtx& c, int parity, bool inside, unsigned ops) {
    TimedLaunch t("signed_contacts", c.stream);
    if (inside) {
        for (int k = 0; k < kSignedInsidePasses; ++k)
            launch_k(k_signed_pass<true>, dim3(c.ntiles), dim3(kTile), 0, c.stream, c.views, c.tiles, c.ntiles, parity, ops);
    } else {
        launch_k(k_signed_pass<false>, dim3(c.ntiles), dim3(kTile), 0, c.stream, c.views, c.tiles, c.ntiles, parity, ops);
    }
}

void launch_stretch_color(const LaunchCtx& c, const int4* segments, int nseg_blocks, int parity, unsigned ops) {
    if (nseg_blocks <= 0) return;
    TimedLaunch t("stretch_color", c.stream);
    launch_k(k_stretch_color, dim3(nseg_blocks), dim3(kTile), 0, c.stream, c.views, segments, nseg_blocks, parity, ops);
}

void launch_bend_color(const LaunchCtx& c, const int4* segments, int nseg_blocks, int parity, unsigned ops) {
    if (nseg_blocks <= 0) return;
    TimedLaunch t("bend_color", c.stream);
    launch_k(k_bend_color, dim3(nseg_blocks), dim3(kTile), 0, c.stream, c.views, segments, parity, ops);
}

void launch_sdf_build(hipStream_t s, float* sdf, int dx, int dy, int dz, float ox, float oy, float oz, float voxel,
                      const float* tris, int ntris, int signed_sdf) {
    const long long total = static_cast<long long>(dx) * dy * dz;
    const unsigned blocks = static_cast<unsigned>((total + kTile - 1) / kTile);
    TimedLaunch t("sdf_build", s);
    launch_k(k_sdf_build, dim3(blocks), dim3(kTile), 0, s, sdf, dx, dy, dz, ox, oy, oz, voxel, tris, ntris, signed_sdf);
}

void launch_fill_u32(hipStream_t s, unsigned* p, unsigned value, size_t n) {
    if (n == 0) return;
    TimedLaunch t("fill", s);
    launch_k(k_fill_u32, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, p, value, n);
}

}  // namespace ssbl
