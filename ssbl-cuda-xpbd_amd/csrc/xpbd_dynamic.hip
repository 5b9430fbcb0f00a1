// xpbd_dynamic.hip -- dynamic collider sets on gfx950: uploads, device
// gathers of a scene from cloth handles, bounds and the two hashes.
//
// Reference (native/src/ssbl_xpbd_cuda_abi41.cu, A41:<line>): bounds
// A41:3694-3760, triangle hash A41:3762-3932 with the >64-cell list, compact
// offsets A41:6354-6390 (a host scan in the reference; a device scan here),
// particle hash A41:4325-4352, cell sizes A41:6590-6601 and 8619-8627.
#include "xpbd_dynamic.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "xpbd_math.hpp"
#include "xpbd_timing.hpp"

namespace ssbl {

bool fail(const std::string& message);
bool hip_ok(hipError_t err, const char* what);
bool sync_stream();

namespace {

unsigned grid_for(long long n, int block) { return static_cast<unsigned>(std::max<long long>((n + block - 1) / block, 1)); }

int next_pow2(int v) {
    int r = 1;
    while (r < v && r < (1 << 30)) r <<= 1;
    return r;
}

__device__ __host__ inline float tri_cell_size(float max_extent, float margin, float thickness, bool fine) {  // A41:6590-6601
    const float coarse = fmaxf(fmaxf(max_extent + margin * 2.0f, thickness), 1.0e-3f);
    if (!fine) return coarse;
    const float fine_cell = fmaxf(fmaxf(thickness * 1.5f, margin * 4.0f), 0.02f);
    return fmaxf(fminf(coarse, fine_cell), 1.0e-3f);
}

__global__ void k_set_view(DynView* dst, DynView v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = v;
}

__global__ void k_fill_i32(int* p, int value, int n) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = value;
}

// A41:3694-3760: AABB, centre, bounding radius and (hashed sets) the cell
// span with the margin; a span of more than 64 cells goes to the large list.
__global__ __launch_bounds__(256) void k_bounds(const DynView* __restrict__ dv, float span_margin, TriBox* box, int* large_flag,
                                                unsigned long long* counts, int count_diag) {
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const int nt = dv->nt;
    bool large = false;
    if (t < nt) {
        const F3 a = f3(dv->tv[3 * t]), b = f3(dv->tv[3 * t + 1]), c = f3(dv->tv[3 * t + 2]);
        const F3 mn = f3(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)));
        const F3 mx = f3(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)));
        const F3 cen = mul((a + b) + c, 1.0f / 3.0f);
        const float r = sqrtf(fmaxf(fmaxf(dot(a - cen, a - cen), dot(b - cen, b - cen)), dot(c - cen, c - cen)));
        int4 c0 = make_int4(0, 0, 0, 0), c1 = make_int4(0, 0, __float_as_int(mn.x), __float_as_int(mn.y));
        if (dv->tri_hash) {
            const float cs = fmaxf(dv->tri_cell, 1.0e-3f), m = span_margin;
            int x0 = cell_coord(mn.x - m, cs), y0 = cell_coord(mn.y - m, cs), z0 = cell_coord(mn.z - m, cs);
            int x1 = cell_coord(mx.x + m, cs), y1 = cell_coord(mx.y + m, cs), z1 = cell_coord(mx.z + m, cs);
            const long long sx = 1LL * x1 - x0 + 1, sy = 1LL * y1 - y0 + 1, sz = 1LL * z1 - z0 + 1;
            long long cells = (sx > 0 && sy > 0 && sz > 0) ? sx * sy * sz : 0;
            if (cells <= 0) {
                x0 = x1 = cell_coord(cen.x, cs);
                y0 = y1 = cell_coord(cen.y, cs);
                z0 = z1 = cell_coord(cen.z, cs);
                cells = 1;
            }
            c0 = make_int4(x0, y0, z0, x1);
            c1.x = y1;
            c1.y = z1;
            large = cells > kDynMaxHashCells;
        }
        TriBox bx;
        bx.c0 = c0;
        bx.c1 = c1;
        bx.b0 = make_float4(mn.z, mx.x, mx.y, mx.z);
        bx.s = with_w(cen, r);
        box[t] = bx;
        large_flag[t] = large ? 1 : 0;
    }
    if (count_diag) wave_count(counts, kCountDynTriLarge, large);
}

__global__ void k_box_centres(const TriBox* __restrict__ box, float4* ctr, int n) {
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t < n) ctr[t] = box[t].s;
}

__global__ void k_large_scatter(int nt, const int* __restrict__ flag, const int* __restrict__ off, int* large) {
    const int t = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (t < nt && flag[t]) large[off[t]] = t;
}

// A triangle's cell span as extents and their reciprocals: cell idx of the
// span (z outer, then y, then x: the reference's loop order) decoded without
// integer division.  Spans hold at most 64 cells, so for idx < 64 and
// divisors up to 64 the quotient (idx + 0.5) / d lies at least 1/128 from an
// integer, far beyond the error of a hardware reciprocal, and truncation gives
// the exact integer quotient.
struct CellSpan {
    int4 lo;
    int sx, sxy, n;
    float rsx, rsxy;
    __device__ __forceinline__ void init(int4 a, int4 b) {
        lo = a;
        sx = b.x - a.x + 1;
        const int sy = b.y - a.y + 1, sz = b.z - a.z + 1;
        sxy = sx * sy;
        n = (sx > 0 && sy > 0 && sz > 0) ? sxy * sz : 0;
        rsx = __builtin_amdgcn_rcpf(static_cast<float>(max(sx, 1)));
        rsxy = __builtin_amdgcn_rcpf(static_cast<float>(max(sxy, 1)));
    }
    __device__ __forceinline__ void cell(int idx, int* x, int* y, int* z) const {
        const int q = static_cast<int>((static_cast<float>(idx) + 0.5f) * rsxy);
        const int r = idx - q * sxy;
        const int w = static_cast<int>((static_cast<float>(r) + 0.5f) * rsx);
        *x = lo.x + (r - w * sx);
        *y = lo.y + w;
        *z = lo.z + q;
    }
};

constexpr int kTriLanes = 4;      // lanes per triangle in the hash builds: lane l takes cells l, l + 4, ...
constexpr int kTriPerBlock = 256 / kTriLanes;
constexpr int kAggSlots = 1024;   // LDS table of the distinct buckets a block's entries fall in
constexpr int kAggProbes = 32;    // a bucket with no slot within this many probes counts globally, entry by entry

__device__ __forceinline__ int agg_slot(int* s_bucket, int* s_list, int* s_n, int b) {
    return lds_bucket_slot<kAggSlots, kAggProbes>(s_bucket, s_list, s_n, b);
}

// kTriLanes lanes per triangle (32 triangles per block).  Phase 0 counts the
// entries of every bucket; phase 1 counts them down again, each entry taking
// the place its decrement returns in the bucket's segment, in arrival order:
// ranked sets write raw[off[b] + slot] = {key, b} for k_tri_rank to sort;
// the others, whose queries take the deepest contact with the smallest key
// whatever the order, write ent[off[b] + slot] = {key, source, layer, 0}.
// Neighbouring triangles share most of their cells, so a block first gathers
// its entries per bucket in an LDS table and makes one global atomic per
// distinct bucket (device-scope atomics are the cost); phase 1 keeps each
// entry's table slot in LDS and then hands out the places of each bucket's
// block-wide range.
__global__ __launch_bounds__(256) void k_tri_hash(const DynView* __restrict__ dv, const TriBox* __restrict__ box,
                                                  const int* __restrict__ large_flag, int* count, const int* __restrict__ off,
                                                  int2* raw, int4* ent, const int* __restrict__ tsrc,
                                                  const int* __restrict__ tlayer, const int* __restrict__ order, int ranked,
                                                  int phase) {
    __shared__ int s_bucket[kAggSlots];
    __shared__ int s_count[kAggSlots];
    __shared__ int s_base[kAggSlots];
    __shared__ int s_list[kAggSlots];
    __shared__ short s_slot[kTriPerBlock * kDynMaxHashCells];
    __shared__ int s_n;
    for (int i = static_cast<int>(threadIdx.x); i < kAggSlots; i += static_cast<int>(blockDim.x)) {
        s_bucket[i] = -1;
        s_count[i] = 0;
    }
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int tl = static_cast<int>(threadIdx.x) / kTriLanes;
    const int nt = dv->nt;
    const int at_k = static_cast<int>(blockIdx.x) * kTriPerBlock + tl;   // the block's k-th triangle in visiting order
    const int t = (order && at_k < nt) ? order[at_k] : at_k;
    const int lane = static_cast<int>(threadIdx.x) % kTriLanes;
    const bool live = t < nt && !large_flag[t];   // a span of at most 64 cells
    CellSpan sp;
    sp.init(make_int4(0, 0, 0, 0), make_int4(-1, -1, -1, 0));
    if (live) {
        const TriBox bx = box[t];
        sp.init(box_cmin(bx), box_cmax(bx));
    }
    const int buckets = dv->tri_buckets;
    // the triangle's layer range (its layer's place among the scene's layers)
    const int classes = dv->tri_classes;
    int cls = 0;
    if (live && classes > 1 && tlayer) {
        const int layer = tlayer[t];
        for (int k = 1; k < classes; ++k) cls += layer >= dv->tri_class_layer[k] ? 1 : 0;
    }
    const int clo = dv->tri_class_base[cls], chi = dv->tri_class_base[cls + 1];
    short* my_slot = s_slot + tl * kDynMaxHashCells;
    for (int idx = lane; idx < sp.n; idx += kTriLanes) {
        int x, y, z;
        sp.cell(idx, &x, &y, &z);
        const int b = tri_bucket(classes, buckets, clo, chi, x, y, z);
        const int h = agg_slot(s_bucket, s_list, &s_n, b);
        if (h >= 0) atomicAdd(&s_count[h], 1);
        else if (phase == 0) atomicAdd(&count[b], 1);
        my_slot[idx] = static_cast<short>(h);
    }
    __syncthreads();
    const int n = s_n;
    for (int i = static_cast<int>(threadIdx.x); i < n; i += static_cast<int>(blockDim.x)) {
        const int h = s_list[i], b = s_bucket[h], k = s_count[h];
        if (phase == 0) {
            atomicAdd(&count[b], k);
        } else {
            s_base[h] = atomicSub(&count[b], k) - k;
            s_count[h] = 0;
        }
    }
    if (phase == 0) return;
    __syncthreads();
    int2 sl = make_int2(0, 0);
    if (live && !ranked && tsrc) sl = make_int2(tsrc[t], tlayer[t]);
    for (int idx = lane; idx < sp.n; idx += kTriLanes) {
        const int h = my_slot[idx];
        int b, slot;
        if (h >= 0) {
            b = s_bucket[h];
            slot = s_base[h] + atomicAdd(&s_count[h], 1);
        } else {
            int x, y, z;
            sp.cell(idx, &x, &y, &z);
            b = tri_bucket(classes, buckets, clo, chi, x, y, z);
            slot = atomicSub(&count[b], 1) - 1;
        }
        const int at = off[b] + slot, key = t * kDynMaxHashCells + idx;
        if (ranked) {
            raw[at] = make_int2(key, b);
        } else {
            int x, y, z;
            sp.cell(idx, &x, &y, &z);
            ent[at] = make_int4(key, sl.x, sl.y, cell_tag(x, y, z));
        }
    }
}

// Max over a 256-thread block (wave shuffles, then the four wave maxima).
__device__ __forceinline__ int block_max(int m) {
    __shared__ int wave_max[4];
    for (int d = 32; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d, 64));
    if ((threadIdx.x & 63u) == 0u) wave_max[threadIdx.x >> 6] = m;
    __syncthreads();
    return max(max(wave_max[0], wave_max[1]), max(wave_max[2], wave_max[3]));
}

// Largest bucket (diagnostics): one atomic per block.
__global__ __launch_bounds__(256) void k_bucket_max(const int* __restrict__ off, int buckets, int* out) {
    int m = 0;
    for (int b = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); b < buckets; b += static_cast<int>(gridDim.x * blockDim.x))
        m = max(m, off[b + 1] - off[b]);
    m = block_max(m);
    if (threadIdx.x == 0 && m > 0) atomicMax(out, m);
}

// Every scattered key lands at its rank among its bucket's keys, with its
// triangle's {source, layer} beside it (grid-stride over the entries).
__global__ __launch_bounds__(256) void k_tri_rank(const int* __restrict__ off, int buckets, const int2* __restrict__ raw, int* keys,
                                                  const int* __restrict__ tsrc, const int* __restrict__ tlayer, int2* ksl) {
    const int total = off[buckets];
    for (int j = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); j < total; j += static_cast<int>(gridDim.x * blockDim.x)) {
        const int2 e = raw[j];
        const int b = e.y, key = e.x;
        const int base = off[b], n = off[b + 1] - base;
        int rank = 0;
        for (int k = 0; k < n; ++k) rank += raw[base + k].x < key ? 1 : 0;
        keys[base + rank] = key;
        if (tsrc) {
            const int t = key / kDynMaxHashCells;
            ksl[base + rank] = make_int2(tsrc[t], tlayer[t]);
        }
    }
}

// One 16-byte bin per bucket for the queries: {offset, count, smallest layer,
// the source every one of the first `cap` entries shares (-1 when they
// differ, -2 when the bucket is empty)}.  A query whose filter rejects every
// entry of a bucket skips it (A41:3984-3997 filters each entry the same way,
// so results are equal).  Sets without sources get layer INT_MIN, source -1.
__device__ __forceinline__ int2 entry_source_layer(int2 e) { return e; }
__device__ __forceinline__ int2 entry_source_layer(const PartEntry& e) { return make_int2(e.sl.x, e.sl.y); }
__device__ __forceinline__ int2 entry_source_layer(int4 e) { return make_int2(e.y, e.z); }
__device__ __forceinline__ int entry_tag(const PartEntry&) { return 0; }
__device__ __forceinline__ int entry_tag(int4 e) { return e.w; }
__device__ __forceinline__ int entry_tag(int2) { return 0; }
// tagged: bins of a triangle hash with layer ranges, whose third field is
// the cell tag all entries share (0: several) instead of the minimum layer
// (one layer per range)
template <typename R>
__global__ __launch_bounds__(256) void k_bucket_bins(const int* __restrict__ off, int buckets, int cap, const R* __restrict__ rec,
                                                     int4* bins, int tagged) {
    for (int b = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); b < buckets; b += static_cast<int>(gridDim.x * blockDim.x)) {
        const int base = off[b], n = off[b + 1] - base;
        int min_layer = rec ? 0x7fffffff : static_cast<int>(0x80000000u), src = -2, tag = -1;
        if (rec) {
            // four entries in flight per round (the merge is order-free: a
            // minimum, and whether all sources and cells agree)
            const int m = min(n, cap);
            for (int k = 0; k < m; k += 4) {
                R r[4];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (k + u < m) r[u] = rec[base + k + u];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (k + u >= m) break;
                    const int2 e = entry_source_layer(r[u]);
                    const int g = entry_tag(r[u]);
                    min_layer = min(min_layer, e.y);
                    src = (src == -2 || src == e.x) ? e.x : -1;
                    tag = (tag == -1 || tag == g) ? g : 0;
                }
            }
        } else if (n > 0) {
            src = -1;
        }
        bins[b] = make_int4(base, n, tagged ? (tag == -1 ? 0 : tag) : min_layer, src);
    }
}

// A41:4325-4352 with rank placement: phase 0 counts, 1 scatters, 2 ranks
// (the first 16 particles of a bucket by index are visible to queries).
__global__ __launch_bounds__(256) void k_part_hash(const DynView* __restrict__ dv, int4* pcell, int* pcount, const int* __restrict__ poff,
                                                   int* pslot, int* praw, PartEntry* pent,
                                                   const int* __restrict__ psrc, const int* __restrict__ player,
                                                   unsigned long long* counts, int count_diag, int phase) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const bool active = i < dv->np;
    if (phase == 0) {
        if (!active) return;
        const float4 p = dv->pp[i];
        const float cs = dv->part_cell;
        const int cx = cell_coord(p.x, cs), cy = cell_coord(p.y, cs), cz = cell_coord(p.z, cs);
        const int b = static_cast<int>(hash_cell(cx, cy, cz, dv->part_buckets));
        pcell[i] = make_int4(cx, cy, cz, b);
        pslot[i] = atomicAdd(&pcount[b], 1);
        return;
    }
    if (phase == 1) {
        if (active) praw[poff[pcell[i].w] + pslot[i]] = i;
        return;
    }
    bool over = false;
    if (active) {
        const int b = pcell[i].w, base = poff[b], n = pcount[b];
        int rank = 0;
        for (int k = 0; k < n && rank < kDynParticleSlots; ++k) rank += praw[base + k] < i ? 1 : 0;
        if (rank < kDynParticleSlots) {
            const int4 cell = pcell[i];
            const int2 sl = psrc ? make_int2(psrc[i], player[i]) : make_int2(0, 0);
            PartEntry& e = pent[base + rank];
            e.ck = make_int4(cell.x, cell.y, cell.z, i);
            e.sl = make_int4(sl.x, sl.y, 0, 0);
            e.p = dv->pp[i];
        } else {
            over = true;
        }
    }
    if (count_diag) wave_count(counts, kCountDynParticleOverflow, over);
}

// Coarse particle summary (see DynView::coarse).  Each particle marks every
// coarse cell its reach box overlaps; the box is widened by a relative slack
// so that float rounding of the box bounds cannot leave out the coarse cell of
// a point the particle reaches.
__global__ __launch_bounds__(256) void k_coarse_fill(int2* t, int n, int* off) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) t[i] = make_int2(0x7fffffff, -2);
    if (i == 0) *off = 0;
}

// Merges a {smallest layer, one source | -1 mixed | -2 empty} summary into
// coarse entry e: order-free, so a block may merge its own summary first.
__device__ __forceinline__ void coarse_merge(int2* e, int layer, int src) {
    // plain coherent reads first: most entries already hold the outcome
    if (__hip_atomic_load(&e->x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > layer) atomicMin(&e->x, layer);
    int cur = __hip_atomic_load(&e->y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const int want = cur == -2 ? src : (cur == src ? cur : -1);
        if (want == cur) break;
        const int prev = atomicCAS(&e->y, cur, want);
        if (prev == cur) break;
        cur = prev;
    }
}

// Every particle's reach box marks the coarse cells it overlaps.  The
// particles of a block (neighbours on a cloth) share most of their cells, so
// the block merges its marks per bucket in LDS first and then merges each
// distinct bucket into the table once.
__global__ __launch_bounds__(256) void k_coarse_insert(const DynView* __restrict__ dv, int2* t, int* off) {
    __shared__ int s_bucket[kAggSlots];
    __shared__ int s_layer[kAggSlots];
    __shared__ int s_src[kAggSlots];
    __shared__ int s_list[kAggSlots];
    __shared__ int s_n;
    for (int k = static_cast<int>(threadIdx.x); k < kAggSlots; k += static_cast<int>(blockDim.x)) {
        s_bucket[k] = -1;
        s_layer[k] = 0x7fffffff;
        s_src[k] = -2;
    }
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const DynView& D = *dv;
    bool live = i < D.np;
    const float4 p = live ? D.pp[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (live && (!isfinite(p.x) || !isfinite(p.y) || !isfinite(p.z))) { atomicExch(off, 1); live = false; }
    const int src = D.psrc && live ? D.psrc[i] : -1;
    const int layer = D.psrc ? (D.player && live ? D.player[i] : 0) : static_cast<int>(0x80000000u);
    const float cc = D.coarse_cell, r = D.coarse_reach;
    int lo[3] = {0, 0, 0}, n[3] = {0, 0, 0};
    const float c3[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (!live) break;
        const float reach = r * 1.001f + fabsf(c3[a]) * 1.0e-5f + 1.0e-6f;
        lo[a] = cell_coord(c3[a] - reach, cc);
        const long long span = static_cast<long long>(cell_coord(c3[a] + reach, cc)) - lo[a] + 1;
        // the box is under a coarse cell wide, so this fails only where the
        // cell coordinates saturate: then the summary is not used at all
        if (span < 1 || span > 3) { atomicExch(off, 1); live = false; break; }
        n[a] = static_cast<int>(span);
    }
    if (live) {
        for (int kz = 0; kz < n[2]; ++kz)
            for (int ky = 0; ky < n[1]; ++ky)
                for (int kx = 0; kx < n[0]; ++kx) {
                    const int b = static_cast<int>(hash_cell(lo[0] + kx, lo[1] + ky, lo[2] + kz, D.coarse_buckets));
                    const int h = agg_slot(s_bucket, s_list, &s_n, b);
                    if (h < 0) { coarse_merge(t + b, layer, src); continue; }
                    atomicMin(&s_layer[h], layer);
                    int cur = s_src[h];
                    for (;;) {
                        const int want = cur == -2 ? src : (cur == src ? cur : -1);
                        if (want == cur) break;
                        const int prev = atomicCAS(&s_src[h], cur, want);
                        if (prev == cur) break;
                        cur = prev;
                    }
                }
    }
    __syncthreads();
    const int nb = s_n;
    for (int k = static_cast<int>(threadIdx.x); k < nb; k += static_cast<int>(blockDim.x)) {
        const int h = s_list[k];
        coarse_merge(t + s_bucket[h], s_layer[h], s_src[h]);
    }
}

// Scene gathers: a segment {source, begin, end} is a block's range.
__global__ __launch_bounds__(256) void k_gather_particles(const GatherSource* __restrict__ src, const int4* __restrict__ segs,
                                                          float4* pp, float4* pprev, int* psrc, int* player, int keep_prev,
                                                          int* extent_bits) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *extent_bits = 0;   // the triangle gather's running maximum
    const int4 sg = segs[blockIdx.x];
    const int v = sg.y + static_cast<int>(threadIdx.x);
    if (v >= sg.z) return;
    const GatherSource S = src[sg.x];
    const int i = S.pbase + v;
    const float4 p = S.pos[v];
    const float4 now = make_float4(p.x, p.y, p.z, S.radius);
    pprev[i] = keep_prev ? pp[i] : now;
    pp[i] = now;
    psrc[i] = S.source_id;
    player[i] = S.layer;
}

__global__ __launch_bounds__(256) void k_gather_triangles(const GatherSource* __restrict__ src, const int4* __restrict__ segs,
                                                          float4* tv, int* tsrc, int* tlayer, int* order, int* extent_bits) {
    const int4 sg = segs[blockIdx.x];
    const int t = sg.y + static_cast<int>(threadIdx.x);
    const bool active = t < sg.z;
    const GatherSource S = src[sg.x];
    const int i = S.tbase + (active ? t : 0);
    const int4 tr = active ? S.tris[t] : make_int4(-1, -1, -1, 0);
    const bool ok = tr.x >= 0 && tr.y >= 0 && tr.z >= 0 && tr.x < S.nv && tr.y < S.nv && tr.z < S.nv;
    const F3 a = ok ? f3(S.pos[tr.x]) : f3(0.0f, 0.0f, 0.0f);
    const F3 b = ok ? f3(S.pos[tr.y]) : f3(0.0f, 0.0f, 0.0f);
    const F3 c = ok ? f3(S.pos[tr.z]) : f3(0.0f, 0.0f, 0.0f);
    if (active) {
        tv[3 * i] = with_w(a, 0.0f);
        tv[3 * i + 1] = with_w(b, 0.0f);
        tv[3 * i + 2] = with_w(c, 0.0f);
        tsrc[i] = S.source_id;
        tlayer[i] = S.layer;
        order[i] = S.tbase + (S.order ? S.order[t] : t);
    }
    const float ext = fmaxf(fmaxf(a.x, fmaxf(b.x, c.x)) - fminf(a.x, fminf(b.x, c.x)),
                            fmaxf(fmaxf(a.y, fmaxf(b.y, c.y)) - fminf(a.y, fminf(b.y, c.y)),
                                  fmaxf(a.z, fmaxf(b.z, c.z)) - fminf(a.z, fminf(b.z, c.z))));
    // non-negative finite floats order as ints; one atomic per block
    const int e = block_max((active && isfinite(ext) && ext >= 0.0f) ? __float_as_int(ext) : 0);
    if (threadIdx.x == 0) atomicMax(extent_bits, e);
}

// The scene's triangle cell from the gathered extent (A41:6697-6702).
__global__ void k_scene_cell(DynView* dv, const int* extent_bits, float margin, float thickness) {
    if (blockIdx.x == 0 && threadIdx.x == 0) dv->tri_cell = tri_cell_size(__int_as_float(*extent_bits), margin, thickness, true);
}

}  // namespace

float dynamic_triangle_cell_size(float max_extent, float margin, float thickness, bool fine) {
    return tri_cell_size(max_extent, margin, thickness, fine);
}

// ------------------------------------------------------------------ host
DynamicColliders::~DynamicColliders() { release_all(); }

void* DynamicColliders::alloc(int group, size_t bytes) {
    bytes = std::max<size_t>(bytes, 256);
    void* p = device_alloc(bytes, "dynamic collider allocation");
    if (p) blocks_[group].emplace_back(p, bytes);
    return p;
}

void DynamicColliders::release(int group) {
    for (auto& b : blocks_[group]) device_free(b.first, b.second);
    blocks_[group].clear();
}

void DynamicColliders::release_all() {
    for (int g = 0; g < 5; ++g) release(g);
}

bool DynamicColliders::reserve_triangles(int n) {
    if (!d_view_) {
        d_view_ = static_cast<DynView*>(alloc(3, sizeof(DynView)));
        if (!d_view_) return false;
    }
    if (n <= tri_cap_) return true;
    release(0);
    tri_cap_ = 0;
    const size_t m = static_cast<size_t>(n);
    tv_ = static_cast<float4*>(alloc(0, sizeof(float4) * 3 * m));
    tsrc_ = static_cast<int*>(alloc(0, sizeof(int) * m));
    tlayer_ = static_cast<int*>(alloc(0, sizeof(int) * m));
    tri_order_ = static_cast<int*>(alloc(0, sizeof(int) * m));
    box_ = static_cast<TriBox*>(alloc(0, sizeof(TriBox) * m));
    pctr_ = static_cast<float4*>(alloc(0, sizeof(float4) * m));
    bounds_valid_ = false;
    large_flag_ = static_cast<int*>(alloc(0, sizeof(int) * m));
    large_off_ = static_cast<int*>(alloc(0, sizeof(int) * (m + 1)));
    large_ = static_cast<int*>(alloc(0, sizeof(int) * m));
    const int buckets = next_pow2(std::max(kDynMinBuckets, n * 4));
    tri_count_ = static_cast<int*>(alloc(0, sizeof(int) * buckets));
    tri_off_ = static_cast<int*>(alloc(0, sizeof(int) * (buckets + 1)));
    tri_raw_ = static_cast<int2*>(alloc(0, sizeof(int2) * m * kDynMaxHashCells));
    tri_keys_ = static_cast<int*>(alloc(0, sizeof(int) * m * kDynMaxHashCells));
    tri_ksl_ = static_cast<int2*>(alloc(0, sizeof(int2) * m * kDynMaxHashCells));
    tri_ent_ = static_cast<int4*>(alloc(0, sizeof(int4) * m * kDynMaxHashCells));
    tri_bin_ = static_cast<int4*>(alloc(0, sizeof(int4) * buckets));
    occupancy_ = static_cast<int*>(alloc(0, sizeof(int)));
    extent_bits_ = static_cast<int*>(alloc(0, sizeof(int)));
    const int need_scan = std::max(n, buckets);
    if (need_scan > scan_cap_) {
        release(2);
        scan_tmp_ = static_cast<int*>(alloc(2, sizeof(int) * scan_tmp_ints(need_scan)));
        scan_cap_ = scan_tmp_ ? need_scan : 0;
    }
    if (!tv_ || !tsrc_ || !tlayer_ || !tri_order_ || !box_ || !pctr_ || !large_flag_ || !large_off_ || !large_ || !tri_count_ || !tri_off_ ||
        !tri_raw_ || !tri_keys_ || !tri_ksl_ || !tri_ent_ || !tri_bin_ || !occupancy_ || !extent_bits_ || !scan_tmp_)
        return false;
    tri_cap_ = n;
    tri_buckets_ = buckets;
    return true;
}

bool DynamicColliders::reserve_particles(int n) {
    if (!d_view_) {
        d_view_ = static_cast<DynView*>(alloc(3, sizeof(DynView)));
        if (!d_view_) return false;
    }
    if (n <= part_cap_) return true;
    release(1);
    part_cap_ = 0;
    const size_t m = static_cast<size_t>(n);
    pp_ = static_cast<float4*>(alloc(1, sizeof(float4) * m));
    pprev_ = static_cast<float4*>(alloc(1, sizeof(float4) * m));
    psrc_ = static_cast<int*>(alloc(1, sizeof(int) * m));
    player_ = static_cast<int*>(alloc(1, sizeof(int) * m));
    pcell_ = static_cast<int4*>(alloc(1, sizeof(int4) * m));
    pslot_ = static_cast<int*>(alloc(1, sizeof(int) * m));
    praw_ = static_cast<int*>(alloc(1, sizeof(int) * m));
    pent_ = static_cast<PartEntry*>(alloc(1, sizeof(PartEntry) * m));
    const int buckets = next_pow2(std::max(kDynMinBuckets, n * 4));
    // a particle marks about 3.4 coarse cells and a sheet puts about four
    // particles in one, so 4n buckets keep unrelated cells mostly apart
    coarse_buckets_ = next_pow2(std::max(kDynMinBuckets, n * 4));
    coarse_ = static_cast<int2*>(alloc(1, sizeof(int2) * coarse_buckets_));
    coarse_off_ = static_cast<int*>(alloc(1, sizeof(int)));
    pbin_ = static_cast<int4*>(alloc(1, sizeof(int4) * buckets));
    pcount_ = static_cast<int*>(alloc(1, sizeof(int) * buckets));
    poff_ = static_cast<int*>(alloc(1, sizeof(int) * (buckets + 1)));
    if (buckets > scan_cap_) {
        release(2);
        scan_tmp_ = static_cast<int*>(alloc(2, sizeof(int) * scan_tmp_ints(buckets)));
        scan_cap_ = scan_tmp_ ? buckets : 0;
    }
    if (!pp_ || !pprev_ || !psrc_ || !player_ || !pcell_ || !pslot_ || !praw_ || !pbin_ || !pent_ || !pcount_ ||
        !poff_ || !scan_tmp_ || !coarse_ || !coarse_off_)
        return false;
    part_cap_ = n;
    part_buckets_ = buckets;
    return true;
}

bool DynamicColliders::upload_view(hipStream_t st) {
    DynView& v = view_;
    v = DynView{};
    v.nt = nt_;
    v.tri_hash = nt_ >= kDynTriangleHashMin ? 1 : 0;
    v.tri_buckets = next_pow2(std::max(kDynMinBuckets, nt_ * 4));
    v.tri_cell = tri_cell_;
    v.tv = tv_;
    v.tsrc = tri_filter_ ? tsrc_ : nullptr;
    v.tlayer = tri_filter_ ? tlayer_ : nullptr;
    v.box = box_;
    v.tri_keys = tri_keys_; v.tri_ksl = tri_filter_ ? tri_ksl_ : nullptr; v.large = large_;
    v.tri_bin = tri_bin_;
    v.tri_ent = signed_ ? nullptr : tri_ent_;
    v.nlarge = large_off_ ? large_off_ + nt_ : nullptr;
    v.np = np_;
    v.part_hash = np_ >= kDynParticleHashMin ? 1 : 0;
    v.part_buckets = next_pow2(std::max(kDynMinBuckets, np_ * 4));
    v.part_cell = part_cell_;
    v.pp = pp_;
    v.psrc = part_filter_ ? psrc_ : nullptr;
    v.player = part_filter_ ? player_ : nullptr;
    v.pcount = pcount_; v.poff = poff_; v.pcell = pcell_;
    v.pbin = pbin_;
    v.pent = pent_;
    if (v.part_hash) {
        v.coarse = coarse_;
        v.coarse_off = coarse_off_;
        v.coarse_buckets = coarse_buckets_;
        v.coarse_cell = 4.0f * part_cell_;
        v.coarse_reach = part_cell_;
        v.max_radius = max_radius_;
    }
    v.signed_set = signed_ ? 1 : 0;
    v.tri_motion = tri_motion_ ? 1 : 0;
    v.part_motion = part_motion_ ? 1 : 0;
    v.pctr = pctr_;
    v.pprev = pprev_;
    v.tri_classes = (v.tri_hash && tri_classes_ > 1 && !signed_ && tri_filter_) ? tri_classes_ : 1;
    for (int k = 0; k <= kMaxLayerClasses; ++k) v.tri_class_base[k] = v.tri_classes > 1 ? class_base_[k] : 0;
    for (int k = 0; k < kMaxLayerClasses; ++k) v.tri_class_layer[k] = v.tri_classes > 1 ? class_layer_[k] : 0;
    if (v.tri_hash && v.tri_buckets > tri_buckets_) return fail("dynamic triangle hash capacity");
    if (v.part_hash && v.part_buckets > part_buckets_) return fail("dynamic particle hash capacity");
    // the device copy already holds this view (a scene gathered again from
    // the same cloths); its one device-written field, the scene's triangle
    // cell, is recomputed after every gather anyway
    if (uploaded_valid_ && std::memcmp(&uploaded_, &v, sizeof(DynView)) == 0) return true;
    launch_k(k_set_view, dim3(1), dim3(1), 0, st, d_view_, v);
    uploaded_ = v;
    uploaded_valid_ = true;
    return hip_ok(hipGetLastError(), "dynamic collider view");
}

// A41:8338-8420 / 6686-6722: soup upload; the cell size from the largest extent.
bool DynamicColliders::set_triangles(hipStream_t st, const float* soup, int n, const int* src, const int* layer, float margin,
                                     float thickness, bool fine) {
    n = std::max(n, 0);
    if (!sync_stream()) return false;
    signed_ = false;
    tri_motion_ = false;
    tri_ordered_ = false;
    tri_classes_ = 1;
    if (n == 0) { nt_ = 0; bounds_valid_ = false; return d_view_ ? upload_view(st) : true; }
    if (!reserve_triangles(n)) return false;
    float max_extent = 0.0f;
    std::vector<float4> h(3 * static_cast<size_t>(n));
    for (int k = 0; k < n; ++k) {
        const float* T = soup + 9 * static_cast<size_t>(k);
        for (int j = 0; j < 3; ++j) h[3 * k + j] = make_float4(T[3 * j], T[3 * j + 1], T[3 * j + 2], 0.0f);
        const float ext = std::max(std::max(T[0], std::max(T[3], T[6])) - std::min(T[0], std::min(T[3], T[6])),
                                   std::max(std::max(T[1], std::max(T[4], T[7])) - std::min(T[1], std::min(T[4], T[7])),
                                            std::max(T[2], std::max(T[5], T[8])) - std::min(T[2], std::min(T[5], T[8]))));
        if (std::isfinite(ext) && ext >= 0.0f) max_extent = std::max(max_extent, ext);
    }
    margin = std::max(margin, 0.0f);
    tri_cell_ = tri_cell_size(max_extent, margin, std::max(thickness, 0.0f), fine);
    span_margin_ = margin;
    tri_filter_ = src != nullptr;
    bool ok = hip_ok(hipMemcpy(tv_, h.data(), sizeof(float4) * h.size(), hipMemcpyHostToDevice), "dynamic triangle upload");
    if (ok && src) ok = hip_ok(hipMemcpy(tsrc_, src, sizeof(int) * n, hipMemcpyHostToDevice), "dynamic triangle source upload");
    if (ok && src) {
        std::vector<int> zeros;
        const int* l = layer;
        if (!l) { zeros.assign(n, 0); l = zeros.data(); }
        ok = hip_ok(hipMemcpy(tlayer_, l, sizeof(int) * n, hipMemcpyHostToDevice), "dynamic triangle layer upload");
    }
    if (!ok) return false;
    nt_ = n;
    return upload_view(st) && compute_bounds(st);
}

// The bounds of the current corners, so that the centres always belong to the
// last upload (the signed path keeps them as the previous centres).
bool DynamicColliders::compute_bounds(hipStream_t st) {
    if (nt_ <= 0) return true;
    launch_k(k_bounds, dim3(grid_for(nt_, 256)), dim3(256), 0, st, static_cast<const DynView*>(d_view_), span_margin_, box_, large_flag_,
             static_cast<unsigned long long*>(nullptr), 0);
    bounds_valid_ = true;
    return hip_ok(hipGetLastError(), "dynamic triangle bounds");
}

bool DynamicColliders::set_indexed(hipStream_t st, const float* verts, int nverts, const int* idx, int ntri, bool topology_changed,
                                   float margin, float thickness) {
    nverts = std::max(nverts, 0);
    ntri = std::max(ntri, 0);
    if (!sync_stream()) return false;
    tri_ordered_ = false;
    tri_classes_ = 1;
    if (nverts == 0 || ntri == 0) {
        nt_ = 0;
        signed_ = tri_motion_ = bounds_valid_ = false;
        return d_view_ ? upload_view(st) : true;
    }
    float max_extent = 0.0f;
    std::vector<float4> h(3 * static_cast<size_t>(ntri), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (int t = 0; t < ntri; ++t) {
        const int ia = idx[3 * t], ib = idx[3 * t + 1], ic = idx[3 * t + 2];
        if (ia < 0 || ib < 0 || ic < 0 || ia >= nverts || ib >= nverts || ic >= nverts) continue;
        const float* A = verts + 3 * static_cast<size_t>(ia);
        const float* B = verts + 3 * static_cast<size_t>(ib);
        const float* C = verts + 3 * static_cast<size_t>(ic);
        h[3 * t] = make_float4(A[0], A[1], A[2], 0.0f);
        h[3 * t + 1] = make_float4(B[0], B[1], B[2], 0.0f);
        h[3 * t + 2] = make_float4(C[0], C[1], C[2], 0.0f);
        const float ext = std::max(std::max(std::max(A[0], std::max(B[0], C[0])) - std::min(A[0], std::min(B[0], C[0])),
                                            std::max(A[1], std::max(B[1], C[1])) - std::min(A[1], std::min(B[1], C[1]))),
                                   std::max(A[2], std::max(B[2], C[2])) - std::min(A[2], std::min(B[2], C[2])));
        if (std::isfinite(ext) && ext >= 0.0f) max_extent = std::max(max_extent, ext);
    }
    const bool had_previous = !topology_changed && nt_ == ntri && bounds_valid_ && tri_cap_ >= ntri;
    if (!reserve_triangles(ntri)) return false;
    if (had_previous) launch_k(k_box_centres, dim3(grid_for(ntri, 256)), dim3(256), 0, st, static_cast<const TriBox*>(box_), pctr_, ntri);
    if (!hip_ok(hipMemcpy(tv_, h.data(), sizeof(float4) * h.size(), hipMemcpyHostToDevice), "dynamic indexed triangle upload"))
        return false;
    margin = std::max(margin, 0.0f);
    tri_cell_ = tri_cell_size(max_extent, margin, std::max(thickness, 0.0f), true);
    span_margin_ = margin;
    tri_filter_ = false;
    nt_ = ntri;
    signed_ = true;
    tri_motion_ = had_previous;
    if (!upload_view(st) || !compute_bounds(st)) return false;
    if (!had_previous) launch_k(k_box_centres, dim3(grid_for(ntri, 256)), dim3(256), 0, st, static_cast<const TriBox*>(box_), pctr_, ntri);
    return hip_ok(hipGetLastError(), "dynamic previous centres");
}

bool DynamicColliders::clear_motion(hipStream_t st) {
    if (!tri_motion_ && !part_motion_) return true;
    tri_motion_ = part_motion_ = false;
    return upload_view(st);
}

// A41:8575-8697 / 6620-6668: the particle cell is recomputed when the radii change.
bool DynamicColliders::set_particles(hipStream_t st, const float* pos, const float* prev, const float* radii, const int* src,
                                     const int* layer, int n, float margin, float thickness) {
    n = std::max(n, 0);
    if (!sync_stream()) return false;
    if (n == 0) { np_ = 0; part_motion_ = false; return d_view_ ? upload_view(st) : true; }
    std::vector<float> clean(n);
    float max_radius = 0.0f;
    max_radius_ = 0.0f;   // the largest radius a query sees, max(r, 0) of the raw value (inf stays inf)
    for (int i = 0; i < n; ++i) {
        const float r = std::isfinite(radii[i]) ? radii[i] : 0.0f;
        clean[i] = std::max(r, 0.0f);
        max_radius = std::max(max_radius, clean[i]);
        if (radii[i] > max_radius_) max_radius_ = radii[i];
    }
    if (part_cell_ <= 0.0f || clean != radii_seen_ || np_ != n) {
        part_cell_ = std::max(std::max(max_radius + std::max(margin, 0.0f), std::max(thickness, 0.0f)), 1.0e-3f);
        radii_seen_ = clean;
    }
    const bool had_previous = np_ == n && part_cap_ >= n;
    part_motion_ = had_previous || prev != nullptr;   // A41:8625-8690
    if (!reserve_particles(n)) return false;
    std::vector<float4> h(n);
    for (int i = 0; i < n; ++i) h[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], radii[i]);
    bool ok = true;
    if (prev) {
        std::vector<float4> q(n);
        for (int i = 0; i < n; ++i) q[i] = make_float4(prev[3 * i], prev[3 * i + 1], prev[3 * i + 2], radii[i]);
        ok = hip_ok(hipMemcpy(pprev_, q.data(), sizeof(float4) * n, hipMemcpyHostToDevice), "dynamic particle upload");
    } else if (had_previous) {
        ok = hip_ok(hipMemcpy(pprev_, pp_, sizeof(float4) * n, hipMemcpyDeviceToDevice), "dynamic particle upload");
    } else {
        ok = hip_ok(hipMemcpy(pprev_, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice), "dynamic particle upload");
    }
    ok = ok && hip_ok(hipMemcpy(pp_, h.data(), sizeof(float4) * n, hipMemcpyHostToDevice), "dynamic particle upload");
    part_filter_ = src != nullptr;
    if (ok && src) ok = hip_ok(hipMemcpy(psrc_, src, sizeof(int) * n, hipMemcpyHostToDevice), "dynamic particle source upload");
    if (ok && src) {
        std::vector<int> zeros;
        const int* l = layer;
        if (!l) { zeros.assign(n, 0); l = zeros.data(); }
        ok = hip_ok(hipMemcpy(player_, l, sizeof(int) * n, hipMemcpyHostToDevice), "dynamic particle layer upload");
    }
    if (!ok) return false;
    np_ = n;
    return upload_view(st);
}

bool DynamicColliders::gather(hipStream_t st, const std::vector<GatherSource>& sources_in, float margin, float thickness) {
    std::vector<GatherSource> sources = sources_in;
    int np = 0, nt = 0;
    float max_radius = 0.0f, reach_radius = 0.0f;
    for (GatherSource& s : sources) {
        s.pbase = np;
        s.tbase = nt;
        np += std::max(s.nv, 0);
        nt += std::max(s.nt, 0);
        const float r = std::isfinite(s.radius) ? s.radius : 0.0f;
        max_radius = std::max(max_radius, std::max(r, 0.0f));
        if (s.radius > reach_radius) reach_radius = s.radius;
    }
    margin = std::max(margin, 0.0f);
    thickness = std::max(thickness, 0.0f);
    const bool keep_prev = np == np_ && np > 0;
    // one bucket range per distinct triangle layer, each sized by its share
    // of the triangles (one range when there are more layers than ranges)
    tri_classes_ = 1;
    {
        std::map<int, long long> per_layer;
        for (const GatherSource& s : sources)
            if (s.nt > 0) per_layer[s.layer] += s.nt;
        const long long buckets = next_pow2(std::max(kDynMinBuckets, nt * 4));
        if (nt >= kDynTriangleHashMin && per_layer.size() >= 2 && per_layer.size() <= static_cast<size_t>(kMaxLayerClasses)) {
            int k = 0;
            bool ok = true;
            class_base_[0] = 0;
            for (const auto& kv : per_layer) {
                const long long size = buckets * kv.second / nt;
                ok = ok && size >= 1;
                class_layer_[k] = kv.first;
                class_base_[k + 1] = class_base_[k] + static_cast<int>(size);
                ++k;
            }
            if (ok) tri_classes_ = k;
        }
    }
    // segment tables: {source, begin, end} per 256 items
    std::vector<int4> segs;
    for (size_t k = 0; k < sources.size(); ++k)
        for (int b = 0; b < sources[k].nv; b += 256) segs.push_back(make_int4(static_cast<int>(k), b, std::min(sources[k].nv, b + 256), 0));
    const int npseg = static_cast<int>(segs.size());
    for (size_t k = 0; k < sources.size(); ++k)
        for (int b = 0; b < sources[k].nt; b += 256) segs.push_back(make_int4(static_cast<int>(k), b, std::min(sources[k].nt, b + 256), 1));
    // The same cloths as last time (the common case: a scene refreshed every
    // frame): the device tables already hold these sources, so no upload and
    // no wait for the kernels still reading them.
    auto same_source = [](const GatherSource& x, const GatherSource& y) {
        return x.pos == y.pos && x.tris == y.tris && x.order == y.order && x.nv == y.nv && x.nt == y.nt && x.pbase == y.pbase && x.tbase == y.tbase &&
               std::memcmp(&x.radius, &y.radius, sizeof(float)) == 0 && x.source_id == y.source_id && x.layer == y.layer;
    };
    bool same = np == np_ && nt == nt_ && sources.size() == gathered_.size() && d_sources_ && d_segments_;
    for (size_t k = 0; same && k < sources.size(); ++k) same = same_source(sources[k], gathered_[k]);
    if (same) {
        np_ = np;
        nt_ = nt;
        return launch_gather(st, npseg, static_cast<int>(segs.size()) - npseg, keep_prev, max_radius, reach_radius, margin, thickness);
    }
    if (!sync_stream()) return false;
    gathered_.clear();
    if (!reserve_particles(std::max(np, 1)) || !reserve_triangles(std::max(nt, 1))) return false;
    if (static_cast<int>(sources.size()) > sources_cap_ || static_cast<int>(segs.size()) > segments_cap_) {
        release(4);
        const size_t ns = std::max<size_t>(sources.size(), static_cast<size_t>(sources_cap_));
        const size_t ng = std::max<size_t>(segs.size(), static_cast<size_t>(segments_cap_));
        d_sources_ = static_cast<GatherSource*>(alloc(4, sizeof(GatherSource) * ns));
        d_segments_ = static_cast<int4*>(alloc(4, sizeof(int4) * ng));
        sources_cap_ = d_sources_ ? static_cast<int>(ns) : 0;
        segments_cap_ = d_segments_ ? static_cast<int>(ng) : 0;
    }
    if ((!sources.empty() && !d_sources_) || (!segs.empty() && !d_segments_)) return false;
    if (!sources.empty() &&
        !hip_ok(hipMemcpy(d_sources_, sources.data(), sizeof(GatherSource) * sources.size(), hipMemcpyHostToDevice), "scene sources"))
        return false;
    if (!segs.empty() && !hip_ok(hipMemcpy(d_segments_, segs.data(), sizeof(int4) * segs.size(), hipMemcpyHostToDevice), "scene sources"))
        return false;
    gathered_ = sources;
    np_ = np;
    nt_ = nt;
    return launch_gather(st, npseg, static_cast<int>(segs.size()) - npseg, keep_prev, max_radius, reach_radius, margin, thickness);
}

bool DynamicColliders::launch_gather(hipStream_t st, int npseg, int ntseg, bool keep_prev, float max_radius, float reach_radius,
                                     float margin, float thickness) {
    bounds_valid_ = false;
    part_cell_ = std::max(std::max(max_radius + margin, thickness), 1.0e-3f);
    max_radius_ = reach_radius;
    radii_seen_.clear();
    span_margin_ = margin;
    tri_cell_ = 0.0f;   // computed on the device below
    tri_filter_ = part_filter_ = true;
    signed_ = tri_motion_ = false;   // a gathered scene is never signed (A41:6561)
    if (!upload_view(st)) return false;
    TimedLaunch t("scene_gather", st);
    if (npseg > 0)
        launch_k(k_gather_particles, dim3(npseg), dim3(256), 0, st, d_sources_, d_segments_, pp_, pprev_, psrc_, player_, keep_prev ? 1 : 0,
                 extent_bits_);
    else
        launch_k(k_fill_i32, dim3(1), dim3(1), 0, st, extent_bits_, 0, 1);
    if (ntseg > 0)
        launch_k(k_gather_triangles, dim3(ntseg), dim3(256), 0, st, d_sources_, d_segments_ + npseg, tv_, tsrc_, tlayer_, tri_order_,
                 extent_bits_);
    tri_ordered_ = ntseg > 0;
    launch_k(k_scene_cell, dim3(1), dim3(1), 0, st, d_view_, extent_bits_, margin, thickness);
    return hip_ok(hipGetLastError(), "scene gather");
}

bool DynamicColliders::build(hipStream_t st, unsigned long long* counts, bool count_diag, bool occupancy) {
    if (!d_view_) return true;
    const int cd = count_diag ? 1 : 0;
    TimedLaunch t("dynamic_hash", st);
    if (nt_ > 0) {
        const dim3 g(grid_for(nt_, 256)), blk(256);
        launch_k(k_bounds, g, blk, 0, st, static_cast<const DynView*>(d_view_), span_margin_, box_, large_flag_,
                 counts, cd);
        exclusive_scan(st, large_flag_, nt_, large_off_, scan_tmp_);
        launch_k(k_large_scatter, g, blk, 0, st, nt_, static_cast<const int*>(large_flag_), static_cast<const int*>(large_off_), large_);
        if (nt_ >= kDynTriangleHashMin) {
            const int buckets = next_pow2(std::max(kDynMinBuckets, nt_ * 4));
            const dim3 gw(grid_for(static_cast<long long>(nt_) * kTriLanes, 256));
            launch_k(k_fill_i32, dim3(grid_for(buckets, 256)), blk, 0, st, tri_count_, 0, buckets);
            // signed queries average their contacts in visiting order, so only
            // their buckets are sorted by key
            const int ranked = signed_ ? 1 : 0;
            const int* fsrc = tri_filter_ ? tsrc_ : nullptr;
            {
                TimedLaunch tc("tri_hash_count", st);
                launch_k(k_tri_hash, gw, blk, 0, st, static_cast<const DynView*>(d_view_), static_cast<const TriBox*>(box_),
                         static_cast<const int*>(large_flag_), tri_count_, static_cast<const int*>(tri_off_), tri_raw_, tri_ent_,
                         fsrc, static_cast<const int*>(tlayer_), static_cast<const int*>(tri_ordered_ ? tri_order_ : nullptr),
                         ranked, 0);
            }
            exclusive_scan(st, tri_count_, buckets, tri_off_, scan_tmp_);
            {
                TimedLaunch ts("tri_hash_scatter", st);
                launch_k(k_tri_hash, gw, blk, 0, st, static_cast<const DynView*>(d_view_), static_cast<const TriBox*>(box_),
                         static_cast<const int*>(large_flag_), tri_count_, static_cast<const int*>(tri_off_), tri_raw_, tri_ent_,
                         fsrc, static_cast<const int*>(tlayer_), static_cast<const int*>(tri_ordered_ ? tri_order_ : nullptr),
                         ranked, 1);
            }
            if (ranked) {
                const dim3 gr(std::min<unsigned>(grid_for(static_cast<long long>(nt_) * 16, 256), 16384u));
                launch_k(k_tri_rank, gr, blk, 0, st, static_cast<const int*>(tri_off_), buckets, static_cast<const int2*>(tri_raw_),
                         tri_keys_, fsrc, static_cast<const int*>(tlayer_), tri_ksl_);
            }
            const dim3 gb(std::min<unsigned>(grid_for(buckets, 256), 16384u));
            TimedLaunch tb("tri_bins", st);
            if (ranked)
                launch_k(k_bucket_bins<int2>, gb, blk, 0, st, static_cast<const int*>(tri_off_), buckets, 0x7fffffff,
                         static_cast<const int2*>(tri_filter_ ? tri_ksl_ : nullptr), tri_bin_, 0);
            else
                launch_k(k_bucket_bins<int4>, gb, blk, 0, st, static_cast<const int*>(tri_off_), buckets, 0x7fffffff,
                         static_cast<const int4*>(tri_filter_ ? tri_ent_ : nullptr), tri_bin_, view_.tri_classes > 1 ? 1 : 0);
            if (count_diag && occupancy) {
                launch_k(k_fill_i32, dim3(1), dim3(1), 0, st, occupancy_, 0, 1);
                launch_k(k_bucket_max, dim3(std::min<unsigned>(grid_for(buckets, 256), 4096u)), blk, 0, st,
                         static_cast<const int*>(tri_off_), buckets, occupancy_);
            }
        }
    }
    if (np_ >= kDynParticleHashMin) {
        TimedLaunch tp("part_hash", st);
        const int buckets = next_pow2(std::max(kDynMinBuckets, np_ * 4));
        const dim3 g(grid_for(np_, 256)), blk(256);
        launch_k(k_fill_i32, dim3(grid_for(buckets, 256)), blk, 0, st, pcount_, 0, buckets);
        for (int phase = 0; phase < 3; ++phase) {
            if (phase == 1) exclusive_scan(st, pcount_, buckets, poff_, scan_tmp_);
            launch_k(k_part_hash, g, blk, 0, st, static_cast<const DynView*>(d_view_), pcell_, pcount_, static_cast<const int*>(poff_),
                     pslot_, praw_, pent_, static_cast<const int*>(part_filter_ ? psrc_ : nullptr),
                     static_cast<const int*>(player_), counts, cd, phase);
        }
        launch_k(k_bucket_bins<PartEntry>, dim3(std::min<unsigned>(grid_for(buckets, 256), 16384u)), blk, 0, st,
                 static_cast<const int*>(poff_), buckets, kDynParticleSlots, static_cast<const PartEntry*>(part_filter_ ? pent_ : nullptr),
                 pbin_, 0);
        launch_k(k_coarse_fill, dim3(grid_for(coarse_buckets_, 256)), blk, 0, st, coarse_, coarse_buckets_, coarse_off_);
        launch_k(k_coarse_insert, g, blk, 0, st, static_cast<const DynView*>(d_view_), coarse_, coarse_off_);
    }
    return hip_ok(hipGetLastError(), "dynamic collider hash");
}

long long DynamicColliders::hash_entries() const {
    if (!tri_off_ || nt_ < kDynTriangleHashMin) return 0;
    const int buckets = next_pow2(std::max(kDynMinBuckets, nt_ * 4));
    int n = 0;
    if (!hip_ok(hipMemcpy(&n, tri_off_ + buckets, sizeof(int), hipMemcpyDeviceToHost), "dynamic hash entries")) return -1;
    return n;
}

int DynamicColliders::large_count() const {
    if (!large_off_ || nt_ <= 0) return 0;
    int n = 0;
    if (!hip_ok(hipMemcpy(&n, large_off_ + nt_, sizeof(int), hipMemcpyDeviceToHost), "dynamic large count")) return 0;
    return std::max(0, std::min(n, nt_));
}

int DynamicColliders::max_occupancy() const {
    if (!occupancy_ || nt_ < kDynTriangleHashMin) return 0;
    int n = 0;
    if (!hip_ok(hipMemcpy(&n, occupancy_, sizeof(int), hipMemcpyDeviceToHost), "dynamic bucket occupancy")) return 0;
    return std::max(n, 0);
}

}  // namespace ssbl
