// xpbd_self.hip -- self collision on gfx950: spatial hashes built by counting
// sort, the vertex-vertex candidate table, and the vertex-vertex,
// vertex-triangle and edge-edge responses (strict mode).
//
// Reference (native/src/ssbl_xpbd_cuda_abi41.cu, A41:<line>): cell and hash
// A41:654-664, hash builds A41:4550-4575 / 4632-4694 / 5532-5593, candidate
// table A41:4738-4814 with validity A41:4696-4717, vertex-vertex response
// A41:4816-4951, vertex-triangle A41:4990-5224, edge-edge A41:5595-5844,
// averaging A41:4953-4988, sizing A41:6784-6994.
//
// The reference's buckets are fixed 32-slot arrays filled by atomics, so their
// order (and which entries overflow) depends on scheduling.  Here a bucket is
// a counting-sorted segment whose first 32 entries are placed by rank in
// ascending key order -- the order a query visits -- so a build is
// deterministic and a query sees exactly the entries the CPU oracle sees.  Responses accumulate into
// per-vertex 64-bit fixed-point sums (2^-32 m), which makes the sum independent
// of the order the atomics land in.
#include "xpbd_self.hpp"

#include "xpbd_math.hpp"
#include "xpbd_timing.hpp"

#include <algorithm>
#include <cstdio>

namespace ssbl {

namespace {

constexpr int kScanBlock = 1024;          // threads of the scan kernels
constexpr int kScanItems = 4;             // items per thread per chunk
constexpr int kScanChunk = kScanBlock * kScanItems;
constexpr int kMaxPrimCells = 64;         // A41:34

int next_pow2(int v) {
    int r = 1;
    v = std::max(v, 1);
    while (r < v && r < (1 << 30)) r <<= 1;
    return r;
}

unsigned grid_for(long long n, int block) { return static_cast<unsigned>(std::max<long long>((n + block - 1) / block, 1)); }

// ----------------------------------------------------------- exclusive scan
// Inclusive wave64 scan through DPP-free lane shuffles.
__device__ __forceinline__ int wave_inclusive_scan(int v) {
    const int lane = static_cast<int>(threadIdx.x & 63u);
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total.
__device__ __forceinline__ int block_exclusive_scan(int v, int* total) {
    __shared__ int wave_sums[kScanBlock / 64];
    const int lane = static_cast<int>(threadIdx.x & 63u), wave = static_cast<int>(threadIdx.x >> 6);
    const int inc = wave_inclusive_scan(v);
    if (lane == 63) wave_sums[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        const int s = lane < kScanBlock / 64 ? wave_sums[lane] : 0;
        const int si = wave_inclusive_scan(s);
        if (lane < kScanBlock / 64) wave_sums[lane] = si - s;
        if (lane == kScanBlock / 64 - 1) *total = si;
    }
    __syncthreads();
    const int r = wave_sums[wave] + inc - v;
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const int* __restrict__ in, int n, int* __restrict__ block_sums) {
    const long long base = static_cast<long long>(blockIdx.x) * kScanChunk;
    int s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = base + static_cast<long long>(k) * kScanBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    __shared__ int total;
    block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_block_sums(int* __restrict__ sums, int nblocks) {
    __shared__ int total;
    int carry = 0;
    for (int base = 0; base < nblocks; base += kScanBlock) {
        const int i = base + static_cast<int>(threadIdx.x);
        const int v = i < nblocks ? sums[i] : 0;
        const int ex = block_exclusive_scan(v, &total);
        if (i < nblocks) sums[i] = carry + ex;
        carry += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_final(const int* __restrict__ in, int n, const int* __restrict__ block_sums,
                                                           int* __restrict__ out) {
    const long long base = static_cast<long long>(blockIdx.x) * kScanChunk;
    // each thread owns kScanItems consecutive items
    const long long first = base + static_cast<long long>(threadIdx.x) * kScanItems;
    int vals[kScanItems];
    int s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = first + k;
        vals[k] = i < n ? in[i] : 0;
        s += vals[k];
    }
    __shared__ int total;
    int run = block_exclusive_scan(s, &total) + block_sums[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k) {
        const long long i = first + k;
        if (i < n) out[i] = run;
        run += vals[k];
    }
    if (first <= n - 1 && first + kScanItems > n - 1) out[n] = run;   // grand total after the last item
}

}  // namespace

int scan_tmp_ints(int n) { return (std::max(n, 1) + kScanChunk - 1) / kScanChunk + 2; }

void exclusive_scan(hipStream_t s, const int* in, int n, int* out, int* tmp) {
    const int nb = (n + kScanChunk - 1) / kScanChunk;
    TimedLaunch t("hash_scan", s);
    launch_k(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, s, in, n, tmp);
    launch_k(k_scan_block_sums, dim3(1), dim3(kScanBlock), 0, s, tmp, nb);
    launch_k(k_scan_final, dim3(nb), dim3(kScanBlock), 0, s, in, n, tmp, out);
}

namespace {

// ------------------------------------------------------------ hash builds
__device__ __forceinline__ const float4* cur_pos(const ClothView& c, int parity) { return parity ? c.pos1 : c.pos0; }

__global__ void k_fill_i32(int* p, int value, int n) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = value;
}

// Rank of `key` among the keys raw[base .. base + n) of its bucket, counted up
// to kBucketSlots (a rank of kBucketSlots means "not among the first 32").
__device__ __forceinline__ int bucket_rank(const int* __restrict__ raw, int base, int n, int key) {
    int rank = 0;
    for (int k = 0; k < n && rank < kBucketSlots; ++k) rank += raw[base + k] < key ? 1 : 0;
    return rank;
}

// A41:4550-4575: vertex cell and bucket.  phase 0 counts, 1 scatters (the
// rank pass then keeps the first 32 of a bucket; A41:4570-4574 drops the
// others).
__device__ __forceinline__ void vertex_hash_item(const ClothView& c, const SelfView& sv, int parity, int at, int phase) {
    if (at >= sv.nv) return;
    const int i = sv.vorder ? sv.vorder[at] : at;
    if (phase == 0) {
        const float4 p = cur_pos(c, parity)[i];
        const int cx = cell_coord(p.x, sv.cell), cy = cell_coord(p.y, sv.cell), cz = cell_coord(p.z, sv.cell);
        const int b = static_cast<int>(hash_cell(cx, cy, cz, sv.vh.buckets));
        sv.vcell[i] = make_int4(cx, cy, cz, b);
        sv.vh.slot[i] = atomicAdd(&sv.vh.count[b], 1);
    } else {
        const int b = sv.vcell[i].w;
        const int j = sv.vh.off[b] + sv.vh.slot[i];
        sv.vh.raw[j] = i;
        sv.vh.rawb[j] = sv.vh.base + b;
    }
}

// Primitive box in cells (A41:4632-4694, 5532-5593) with the >64-cell guard
// (A41:871-900).  Returns the number of cells.
__device__ inline int prim_box(F3 lo, F3 hi, F3 fb, float cs, int* b0, int* b1, bool* guarded) {
    int x0 = cell_coord(lo.x, cs), y0 = cell_coord(lo.y, cs), z0 = cell_coord(lo.z, cs);
    int x1 = cell_coord(hi.x, cs), y1 = cell_coord(hi.y, cs), z1 = cell_coord(hi.z, cs);
    const long long sx = static_cast<long long>(x1) - x0 + 1, sy = static_cast<long long>(y1) - y0 + 1,
                    sz = static_cast<long long>(z1) - z0 + 1;
    *guarded = false;
    if (sx > 0 && sy > 0 && sz > 0 && sx * sy * sz > kMaxPrimCells) {
        x0 = x1 = cell_coord(fb.x, cs);
        y0 = y1 = cell_coord(fb.y, cs);
        z0 = z1 = cell_coord(fb.z, cs);
        *guarded = true;
    }
    b0[0] = x0; b0[1] = y0; b0[2] = z0;
    b1[0] = x1; b1[1] = y1; b1[2] = z1;
    if (x1 < x0 || y1 < y0 || z1 < z0) return 0;
    return (x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 1);
}

__device__ inline bool tri_valid(const SelfView& sv, int4 t) {
    return t.x < sv.nv && t.y < sv.nv && t.z < sv.nv && t.x != t.y && t.y != t.z && t.x != t.z;
}
__device__ inline bool edge_valid(const SelfView& sv, int2 e) { return e.x < sv.nv && e.y < sv.nv && e.x != e.y; }

__device__ inline bool tri_box(const ClothView& c, const SelfView& sv, int parity, int t, int* b0, int* b1, bool* g) {
    const int4 tr = sv.tris[t];
    if (!tri_valid(sv, tr)) return false;
    const float4* P = cur_pos(c, parity);
    const F3 a = f3(P[tr.x]), b = f3(P[tr.y]), cc = f3(P[tr.z]);
    const float m = sv.prim_margin;
    const F3 lo = f3(fminf(a.x, fminf(b.x, cc.x)) - m, fminf(a.y, fminf(b.y, cc.y)) - m, fminf(a.z, fminf(b.z, cc.z)) - m);
    const F3 hi = f3(fmaxf(a.x, fmaxf(b.x, cc.x)) + m, fmaxf(a.y, fmaxf(b.y, cc.y)) + m, fmaxf(a.z, fmaxf(b.z, cc.z)) + m);
    prim_box(lo, hi, mul((a + b) + cc, 1.0f / 3.0f), sv.cell, b0, b1, g);
    return true;
}

__device__ inline bool edge_box(const ClothView& c, const SelfView& sv, int parity, int e, int* b0, int* b1, bool* g) {
    const int2 ed = sv.edges[e];
    if (!edge_valid(sv, ed)) return false;
    const float4* P = cur_pos(c, parity);
    const F3 a = f3(P[ed.x]), b = f3(P[ed.y]);
    const float m = sv.prim_margin;
    const F3 lo = f3(fminf(a.x, b.x) - m, fminf(a.y, b.y) - m, fminf(a.z, b.z) - m);
    const F3 hi = f3(fmaxf(a.x, b.x) + m, fmaxf(a.y, b.y) + m, fmaxf(a.z, b.z) + m);
    prim_box(lo, hi, mul(a + b, 0.5f), sv.cell, b0, b1, g);
    return true;
}

// Half a wave per primitive, one lane per cell of its box and a second round
// for boxes over 32 cells (at most 64 cells: larger boxes are reduced to one
// cell by the guard; a box is usually 2-3 cells a side, as the margin is one
// cell).  kind 0 = triangles, 1 = edges; phase 0 counts, 1 scatters.
// Key = primitive * 64 + the cell's index in the box loop (z outer, then y,
// then x).  `blk` is the block's index within its hash's part of the launch;
// every thread of the block calls this (phase 0 has barriers).
//
// Neighbouring primitives share most of their cells, so phase 0 first counts
// the block's items per bucket in an LDS table and then makes one device
// atomic per distinct bucket: each item's slot is the block's base in its
// bucket plus its place among the block's items there (a bucket the table
// has no room for counts item by item).  Slots only need to be distinct
// within a bucket: the rank pass orders the bucket by key.
constexpr int kPrimLanes = 32;
constexpr int kSelfAggSlots = 1024;
constexpr int kSelfAggProbes = 32;
__device__ __forceinline__ void prim_hash_block(const ClothView& c, const SelfView& sv, int parity, int kind, int blk, int phase,
                                                int count_diag) {
    const int t = blk * 256 + static_cast<int>(threadIdx.x);
    const int lane = t % kPrimLanes;
    const int n = kind == 0 ? sv.nt : sv.ne;
    // the block's primitives in the visiting order
    const int* order = kind == 0 ? sv.torder : sv.eorder;
    const int at = t / kPrimLanes;
    const int k = (order && at < n) ? order[at] : at;
    int b0[3] = {0, 0, 0}, b1[3] = {-1, -1, -1};
    bool guarded = false;
    bool ok = false;
    if (k < n) ok = kind == 0 ? tri_box(c, sv, parity, k, b0, b1, &guarded) : edge_box(c, sv, parity, k, b0, b1, &guarded);
    if (ok && phase == 0 && guarded && count_diag && lane == 0) atomicAdd(&c.counts[kCountSelfOverflow], 1ull);
    const int sx = b1[0] - b0[0] + 1, sy = b1[1] - b0[1] + 1, sz = b1[2] - b0[2] + 1;
    const int cells = (ok && sx > 0 && sy > 0 && sz > 0) ? sx * sy * sz : 0;
    const HashView& h = kind == 0 ? sv.th : sv.eh;
    auto bucket_of = [&](int idx) {
        const int x = b0[0] + idx % sx, y = b0[1] + (idx / sx) % sy, z = b0[2] + idx / (sx * sy);
        return static_cast<int>(hash_cell(x, y, z, h.buckets));
    };
    if (phase == 1) {
        for (int idx = lane; idx < cells; idx += kPrimLanes) {
            const int b = bucket_of(idx);
            const int key = k * kMaxPrimCells + idx;
            const int j = h.off[b] + h.slot[key];
            h.raw[j] = key;
            h.rawb[j] = h.base + b;
        }
        return;
    }
    __shared__ int s_bucket[kSelfAggSlots];
    __shared__ int s_count[kSelfAggSlots];
    __shared__ int s_base[kSelfAggSlots];
    __shared__ int s_list[kSelfAggSlots];
    __shared__ int s_n;
    for (int i = static_cast<int>(threadIdx.x); i < kSelfAggSlots; i += static_cast<int>(blockDim.x)) {
        s_bucket[i] = -1;
        s_count[i] = 0;
    }
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    int hs[2] = {-1, -1}, place[2] = {0, 0};
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int idx = lane + r * kPrimLanes;
        if (idx < cells) {
            const int b = bucket_of(idx);
            const int slot = lds_bucket_slot<kSelfAggSlots, kSelfAggProbes>(s_bucket, s_list, &s_n, b);
            if (slot >= 0) {
                hs[r] = slot;
                place[r] = atomicAdd(&s_count[slot], 1);
            } else {
                h.slot[k * kMaxPrimCells + idx] = atomicAdd(&h.count[b], 1);
            }
        }
    }
    __syncthreads();
    const int distinct = s_n;
    for (int i = static_cast<int>(threadIdx.x); i < distinct; i += static_cast<int>(blockDim.x)) {
        const int slot = s_list[i];
        s_base[slot] = atomicAdd(&h.count[s_bucket[slot]], s_count[slot]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 2; ++r)
        if (hs[r] >= 0) h.slot[k * kMaxPrimCells + lane + r * kPrimLanes] = s_base[hs[r]] + place[r];
}

// One phase of all three hash builds in one launch: blocks [0, vh_blocks)
// serve the vertex hash, the next th_blocks the triangle hash, the rest the
// edge hash (each part uniform over its blocks).
__global__ __launch_bounds__(256) void k_self_hash(const ClothView* __restrict__ cv, SelfView sv, int parity, int phase,
                                                   int count_diag) {
    int blk = static_cast<int>(blockIdx.x);
    const int tid = static_cast<int>(threadIdx.x);
    if (blk < sv.vh_blocks) {
        vertex_hash_item(*cv, sv, parity, blk * 256 + tid, phase);
        return;
    }
    blk -= sv.vh_blocks;
    if (blk < sv.th_blocks) {
        prim_hash_block(*cv, sv, parity, 0, blk, phase, count_diag);
        return;
    }
    blk -= sv.th_blocks;
    if (sv.use_ehash) prim_hash_block(*cv, sv, parity, 1, blk, phase, count_diag);
}

constexpr int kBigBucket = 64;   // buckets with more items are ranked by one wave each

// Ascending bitonic sort of one value per lane across the wave.
__device__ __forceinline__ int wave_sort64(int v) {
    const int lane = static_cast<int>(threadIdx.x & 63u);
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const int o = __shfl_xor(v, j, 64);
            v = (((lane & j) == 0) == ((lane & k) == 0)) ? min(v, o) : max(v, o);
        }
    }
    return v;
}

// The 32 smallest keys of every bucket of more than kBigBucket items, one
// wave per bucket: each wave scans a group of 64 buckets and takes its big
// ones in turn.  Each 64-key chunk is sorted across the lanes and merged
// into the running 32 smallest (ascending in lanes 0-31, the chunk's 32
// smallest descending in lanes 32-63: a bitonic sequence, merged ascending).
// The same entries, in the same order, as the per-item ranks (keys are
// distinct); the other n - 32 items count as overflow.
__device__ inline void big_bucket_ranks(const SelfView& sv, const int* __restrict__ raw, int* __restrict__ entries, int count_diag,
                                        int wave, int nwaves) {
    const int lane = static_cast<int>(threadIdx.x & 63u);
    for (int g = wave * 64; g < sv.all_buckets; g += nwaves * 64) {
        const int mine = g + lane < sv.all_buckets ? sv.all_count[g + lane] : 0;
        unsigned long long big = __ballot(mine > kBigBucket);
        while (big != 0ull) {
            const int at = __ffsll(static_cast<long long>(big)) - 1;
            big &= big - 1ull;
            const int b = g + at;
            const int base = sv.all_off[b], n = __shfl(mine, at, 64);
            int top = 0x7fffffff;
            for (int c0 = 0; c0 < n; c0 += 64) {
                const int v = wave_sort64(c0 + lane < n ? raw[base + c0 + lane] : 0x7fffffff);
                // every lane takes part in the shuffle (a lane reading from
                // one that skips it would not get that lane's value)
                const int reversed = __shfl(v, 63 - lane, 64);
                int m = lane < 32 ? top : reversed;
#pragma unroll
                for (int j = 32; j > 0; j >>= 1) {
                    const int o = __shfl_xor(m, j, 64);
                    m = (lane & j) == 0 ? min(m, o) : max(m, o);
                }
                top = lane < 32 ? m : 0x7fffffff;
            }
            if (lane < kBucketSlots) entries[base + lane] = top;
            if (count_diag && lane == 0) atomicAdd(&sv.counts[kCountSelfOverflow], static_cast<unsigned long long>(n - kBucketSlots));
        }
    }
}

// Every scattered item of the shared table finds its rank in its bucket and
// lands there if it is among the first 32 (the others count as overflow).
// A grid-stride loop over the items actually scattered (their total is the
// scan's last offset) by the first item_blocks blocks; every lane runs the
// same number of rounds.  Buckets of more than kBigBucket items are left to
// the other blocks (big_bucket_ranks): the two parts touch different buckets.
__global__ __launch_bounds__(256) void k_self_rank(SelfView sv, const int* __restrict__ raw, const int* __restrict__ rawb,
                                                   int* __restrict__ entries, int count_diag, int item_blocks) {
    if (static_cast<int>(blockIdx.x) >= item_blocks) {
        const int waves = (static_cast<int>(gridDim.x) - item_blocks) * (static_cast<int>(blockDim.x) >> 6);
        big_bucket_ranks(sv, raw, entries, count_diag,
                         (static_cast<int>(blockIdx.x) - item_blocks) * (static_cast<int>(blockDim.x) >> 6) +
                             static_cast<int>(threadIdx.x >> 6),
                         waves);
        return;
    }
    const int total = sv.all_off[sv.all_buckets];
    const int stride = item_blocks * static_cast<int>(blockDim.x);
    const int rounds = (total + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {
        const int j = r * stride + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
        bool over = false;
        if (j < total) {
            const int b = rawb[j], key = raw[j];
            const int base = sv.all_off[b], n = sv.all_count[b];
            if (n <= kBigBucket) {
                const int rank = bucket_rank(raw, base, n, key);
                if (rank < kBucketSlots) entries[base + rank] = key;
                else over = true;
            }
        }
        if (count_diag) wave_count(sv.counts, kCountSelfOverflow, over);
    }
}

// ------------------------------------------------------------- responses
__device__ __forceinline__ void accumulate(const SelfView& sv, const float4* P, int v, F3 d) {
    if (v < 0 || v >= sv.nv || P[v].w == 0.0f) return;   // pinned vertices take no self correction
    atomicAdd(&sv.acc[3 * v + 0], static_cast<unsigned long long>(__double2ll_rn(static_cast<double>(d.x) * kFixedScale)));
    atomicAdd(&sv.acc[3 * v + 1], static_cast<unsigned long long>(__double2ll_rn(static_cast<double>(d.y) * kFixedScale)));
    atomicAdd(&sv.acc[3 * v + 2], static_cast<unsigned long long>(__double2ll_rn(static_cast<double>(d.z) * kFixedScale)));
    atomicAdd(&sv.accw[v], 1);
}

// A41:4594-4617
__device__ inline void triangle_barycentric(F3 p, F3 a, F3 b, F3 c, float* wa, float* wb, float* wc) {
    const F3 v0 = b - a, v1 = c - a, v2 = p - a;
    const float d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
    const float denom = d00 * d11 - d01 * d01;
    if (fabsf(denom) <= kEps) { *wa = *wb = *wc = 1.0f / 3.0f; return; }
    float v = (d11 * d20 - d01 * d21) / denom;
    float w = (d00 * d21 - d01 * d20) / denom;
    v = fminf(fmaxf(v, 0.0f), 1.0f);
    w = fminf(fmaxf(w, 0.0f), 1.0f - v);
    const float u = fmaxf(1.0f - v - w, 0.0f);
    const float sum = fmaxf(u + v + w, kEps);
    *wa = u / sum; *wb = v / sum; *wc = w / sum;
}

// A41:5458-5505
__device__ inline void closest_segment_parameters(F3 p1, F3 q1, F3 p2, F3 q2, float* s, float* t) {
    const F3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    if (a <= kEps && e <= kEps) { *s = 0.0f; *t = 0.0f; return; }
    if (a <= kEps) { *s = 0.0f; *t = clamp01f(f / fmaxf(e, kEps)); return; }
    const float c = dot(d1, r);
    if (e <= kEps) { *t = 0.0f; *s = clamp01f(-c / fmaxf(a, kEps)); return; }
    const float b = dot(d1, d2);
    const float denom = a * e - b * b;
    *s = fabsf(denom) > kEps ? clamp01f((b * f - c * e) / denom) : 0.0f;
    const float tnom = b * (*s) + f;
    if (tnom < 0.0f) { *t = 0.0f; *s = clamp01f(-c / fmaxf(a, kEps)); }
    else if (tnom > e) { *t = 1.0f; *s = clamp01f((b - c) / fmaxf(a, kEps)); }
    else { *t = tnom / e; }
}

__device__ inline bool vv_valid(const ClothView& c, const SelfView& sv, const float4* P, int s, int o) {  // A41:4696-4717
    if (s == o || sv.radius <= 0.0f || sv.onset <= sv.radius || P[s].w == 0.0f) return false;
    const F3 rd = f3(c.rest[s]) - f3(c.rest[o]);
    if (dot(rd, rd) < sv.rest_skip * sv.rest_skip) return false;
    const F3 cd = f3(P[s]) - f3(P[o]), pd = f3(c.prev[s]) - f3(c.prev[o]);
    return dot(cd, cd) < sv.onset * sv.onset || dot(pd, pd) < sv.onset * sv.onset;
}

// ------------------------------------------------------- pair evaluations
// One vertex-triangle (A41:4990-5107) or edge-edge (A41:5595-5722) candidate
// evaluated without side effects: rest = skipped by the rest-distance test,
// accepted = what the pair test returns (a contact within onset), point = the
// contact point that keys the fast mode's clusters, and when apply the
// response delta * w[k] goes to vertex id[k] for each w[k] != 0.
struct PairEval {
    bool rest = false, accepted = false, apply = false;
    int id[4] = {0, 0, 0, 0};
    float w[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    F3 delta, point;
};

__device__ __forceinline__ float contact_push(float r, float onset, float dist) {
    const float pen = r - dist;
    if (pen > 0.0f) return fminf(pen * 0.35f + r * 0.01f, r * 0.25f);
    const float x = 1.0f - fminf(dist / onset, 1.0f);
    return fminf(x * x * r * 0.06f, r * 0.04f);
}

__device__ inline PairEval vt_eval(const ClothView& c, const SelfView& sv, const float4* P, int v, int t) {
    PairEval out;
    const float r = sv.radius, onset = sv.onset;
    if (t < 0 || t >= sv.nt || r <= 0.0f || onset <= r) return out;
    const int4 tr = sv.tris[t];
    if (!tri_valid(sv, tr)) return out;
    const int ia = tr.x, ib = tr.y, ic = tr.z;
    bool rest_near = v == ia || v == ib || v == ic;   // A41:4615-4630
    if (!rest_near) {
        const F3 rv = f3(c.rest[v]);
        const F3 rq = closest_point_on_triangle(rv, f3(c.rest[ia]), f3(c.rest[ib]), f3(c.rest[ic]));
        const F3 rdd = rv - rq;
        rest_near = dot(rdd, rdd) < sv.rest_skip * sv.rest_skip;
    }
    if (rest_near) {
        out.rest = true;
        return out;
    }
    const F3 p = f3(P[v]), a = f3(P[ia]), b = f3(P[ib]), cc = f3(P[ic]);
    const F3 q = closest_point_on_triangle(p, a, b, cc);
    const F3 d = p - q;
    const float dsq = dot(d, d);
    float dist = sqrt_eps(dsq);
    if (dist >= onset) return out;
    F3 nn = f3(0.0f, 0.0f, 1.0f);
    if (dsq > kEps) nn = mul(d, 1.0f / dist);
    else {
        nn = cross(b - a, cc - a);
        const float nsq = dot(nn, nn);
        if (nsq > kEps) nn = mul(nn, rsqrt_exact(nsq));
        else {
            const F3 rv = f3(c.rest[v]);
            const F3 rdl = rv - closest_point_on_triangle(rv, f3(c.rest[ia]), f3(c.rest[ib]), f3(c.rest[ic]));
            const float rsq = dot(rdl, rdl);
            if (rsq > kEps) nn = mul(rdl, rsqrt_exact(rsq));
        }
        dist = 0.0f;
    }
    out.accepted = true;
    out.point = q;
    float wa = 1.0f / 3.0f, wb = wa, wc = wa;
    triangle_barycentric(q, a, b, cc, &wa, &wb, &wc);
    const float push = contact_push(r, onset, dist);
    const float wv = c.rest[v].w, wta = c.rest[ia].w * wa, wtb = c.rest[ib].w * wb, wtc = c.rest[ic].w * wc;
    const float total = wv + wta + wtb + wtc;
    if (push <= 0.0f || total <= 0.0f) return out;
    out.apply = true;
    out.delta = mul(nn, push / total);
    out.id[0] = v; out.id[1] = ia; out.id[2] = ib; out.id[3] = ic;
    out.w[0] = wv > 0.0f ? wv : 0.0f;
    out.w[1] = wta > 0.0f ? -wta : 0.0f;
    out.w[2] = wtb > 0.0f ? -wtb : 0.0f;
    out.w[3] = wtc > 0.0f ? -wtc : 0.0f;
    return out;
}

__device__ inline PairEval ee_eval(const ClothView& c, const SelfView& sv, const float4* P, int ea, int eb) {
    PairEval out;
    const float r = sv.radius, onset = sv.onset;
    if (eb <= ea || ea < 0 || eb >= sv.ne || r <= 0.0f || onset <= r) return out;
    const int2 A = sv.edges[ea], B = sv.edges[eb];
    if (!edge_valid(sv, A) || !edge_valid(sv, B)) return out;
    const int a0 = A.x, a1 = A.y, b0 = B.x, b1 = B.y;
    bool rest_near = a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1;   // A41:5507-5530
    if (!rest_near) {
        const F3 r0 = f3(c.rest[a0]), r1 = f3(c.rest[a1]), r2 = f3(c.rest[b0]), r3 = f3(c.rest[b1]);
        float s = 0.0f, t = 0.0f;
        closest_segment_parameters(r0, r1, r2, r3, &s, &t);
        const F3 dd = (r0 + mul(r1 - r0, s)) - (r2 + mul(r3 - r2, t));
        rest_near = dot(dd, dd) < sv.rest_skip * sv.rest_skip;
    }
    if (rest_near) {
        out.rest = true;
        return out;
    }
    const F3 p0 = f3(P[a0]), p1 = f3(P[a1]), p2 = f3(P[b0]), p3 = f3(P[b1]);
    float s = 0.0f, t = 0.0f;
    closest_segment_parameters(p0, p1, p2, p3, &s, &t);
    const F3 pa = p0 + mul(p1 - p0, s), pb = p2 + mul(p3 - p2, t);
    const F3 d = pa - pb;
    const float dsq = dot(d, d);
    float dist = sqrt_eps(dsq);
    if (dist >= onset) return out;
    F3 nn = f3(0.0f, 0.0f, 1.0f);
    if (dsq > kEps) nn = mul(d, 1.0f / dist);
    else {
        nn = cross(p1 - p0, p3 - p2);
        const float nsq = dot(nn, nn);
        if (nsq > kEps) nn = mul(nn, rsqrt_exact(nsq));
        else {
            const F3 r0 = f3(c.rest[a0]), r1 = f3(c.rest[a1]), r2 = f3(c.rest[b0]), r3 = f3(c.rest[b1]);
            float rs = 0.0f, rt = 0.0f;
            closest_segment_parameters(r0, r1, r2, r3, &rs, &rt);
            const F3 rd = (r0 + mul(r1 - r0, rs)) - (r2 + mul(r3 - r2, rt));
            const float rsq = dot(rd, rd);
            if (rsq > kEps) nn = mul(rd, rsqrt_exact(rsq));
        }
        dist = 0.0f;
    }
    out.accepted = true;
    out.point = mul(pa + pb, 0.5f);
    const float push = contact_push(r, onset, dist);
    const float w0 = 1.0f - s, w1 = s, w2 = 1.0f - t, w3 = t;
    const float ma0 = c.rest[a0].w, ma1 = c.rest[a1].w, mb0 = c.rest[b0].w, mb1 = c.rest[b1].w;
    const float wa0 = ma0 * w0, wa1 = ma1 * w1, wb0 = mb0 * w2, wb1 = mb1 * w3;
    const float total = ma0 * w0 * w0 + ma1 * w1 * w1 + mb0 * w2 * w2 + mb1 * w3 * w3;
    if (push <= 0.0f || total <= 0.0f) return out;
    out.apply = true;
    out.delta = mul(nn, push / total);
    out.id[0] = a0; out.id[1] = a1; out.id[2] = b0; out.id[3] = b1;
    out.w[0] = wa0 > 0.0f ? wa0 : 0.0f;
    out.w[1] = wa1 > 0.0f ? wa1 : 0.0f;
    out.w[2] = wb0 > 0.0f ? -wb0 : 0.0f;
    out.w[3] = wb1 > 0.0f ? -wb1 : 0.0f;
    return out;
}

__device__ __forceinline__ void apply_eval(const SelfView& sv, const float4* P, const PairEval& e) {
    for (int k = 0; k < 4; ++k)
        if (e.w[k] != 0.0f) accumulate(sv, P, e.id[k], mul(e.delta, e.w[k]));
}

// Vertex-vertex response of vertex s to table entry o (A41:4816-4951): the
// continuous time of impact of the two thick points over the substep, else
// the static overlap push, clamped to 0.35 thickness.  Returns whether a
// response is to be applied (in *resp).
__device__ inline bool vv_response(const ClothView& c, const float4* P, int s, int o, F3 pc, F3 pp, float thickness, F3* resp_out) {
    if (thickness <= 0.0f) return false;
    const F3 qc = f3(P[o]), qp = f3(c.prev[o]);
    F3 resp = f3(0.0f, 0.0f, 0.0f);
    bool hit = false;
    const F3 dref = pp - qp;
    const float cco = dot(dref, dref) - thickness * thickness;
    const F3 rel = (pc - pp) - (qc - qp);
    const float aco = dot(rel, rel), bco = dot(dref, rel);
    const bool other_pinned = P[o].w == 0.0f;
    const float scale = other_pinned ? 1.0f : 0.5f;
    if (cco >= 0.0f && bco < 0.0f && aco > 1.0e-5f) {
        const float disc = bco * bco - aco * cco;
        if (disc >= 0.0f) {
            const float t = (-bco - sqrtf(disc)) / aco;
            if (t > 0.0f && t <= 1.0f) {
                const float sv1 = 1.0f - t;
                const F3 ph = mul(pp, sv1) + mul(pc, t), qh = mul(qp, sv1) + mul(qc, t);
                F3 nn = ph - qh;
                const float nsq = dot(nn, nn);
                if (nsq > kEps) {
                    nn = mul(nn, rsqrt_exact(nsq));
                    const float vn = dot((pc - ph) - (qc - qh), nn);
                    if (vn > 0.0f) { resp = mul(nn, -vn * scale); hit = true; }
                }
            }
        }
    }
    if (!hit) {
        const F3 d = pc - qc;
        const float dsq = dot(d, d);
        if (dsq >= thickness * thickness) return false;
        F3 nn = f3(0.0f, 0.0f, 1.0f);
        float dist = sqrt_eps(dsq);
        if (dsq > kEps) nn = mul(d, 1.0f / dist);
        else {
            const F3 rd = f3(c.rest[s]) - f3(c.rest[o]);
            const float rsq = dot(rd, rd);
            if (rsq > kEps) nn = mul(rd, rsqrt_exact(rsq));
            dist = 0.0f;
        }
        resp = mul(nn, fminf((thickness - dist) * scale, thickness * 0.35f));
    }
    if (dot(resp, resp) <= kEps) return false;
    const float ml = fmaxf(thickness * 0.35f, 1.0e-4f);
    const float rl = len3(resp);
    if (rl > ml) resp = mul(resp, ml / rl);
    *resp_out = resp;
    return true;
}

// ----------------------------------------------------- fast mode, per thread
// Fast-mode contact filters (A41:708-849).  Every contact a thread accepts
// opens a cluster keyed by a coarse cell of its contact point, its normal
// direction and its pair type; at most 16 clusters per thread and one contact
// per cluster.  The cluster keys live in LDS, slot-major per thread.
constexpr int kClusterMaxLocal = 16;
constexpr int kIslandMaxContacts = 48;

struct FilterCounts {
    unsigned seen = 0, acc_vv = 0, acc_vt = 0, acc_ee = 0, rej_rest = 0, rej_dup = 0, rej_own = 0, clusters = 0;
};

__device__ __forceinline__ int filter_normal_bucket(float v) { return v > 0.45f ? 2 : (v < -0.45f ? 0 : 1); }

// A41:742-760
__device__ __forceinline__ unsigned filter_cluster_key(int type, F3 point, F3 normal, float cell) {
    const float cs = fmaxf(cell, 1.0e-3f);
    unsigned key = hash_cell(cell_coord(point.x, cs), cell_coord(point.y, cs), cell_coord(point.z, cs), 1 << 20);
    const unsigned nk = static_cast<unsigned>(filter_normal_bucket(normal.x)) | (static_cast<unsigned>(filter_normal_bucket(normal.y)) << 2) |
                        (static_cast<unsigned>(filter_normal_bucket(normal.z)) << 4);
    key ^= nk * 2654435761u;
    key ^= static_cast<unsigned>(type + 1) * 2246822519u;
    return key;
}

struct ClusterSet {
    unsigned* col;   // this thread's column of a [16][blockDim] LDS array
    int stride;
    int n;
    // A41:793-829 with one contact per cluster: a key seen before is refused.
    __device__ __forceinline__ bool try_accept(unsigned key, FilterCounts& fc) {
        for (int i = 0; i < n; ++i)
            if (col[i * stride] == key) { ++fc.rej_own; return false; }
        if (n >= kClusterMaxLocal) { ++fc.rej_own; return false; }
        col[n * stride] = key;
        ++n;
        ++fc.clusters;
        return true;
    }
};

__device__ __forceinline__ void flush_filter_counts(unsigned long long* counts, const FilterCounts& fc) {
    auto add = [&](int slot, unsigned v) { if (v) atomicAdd(&counts[slot], static_cast<unsigned long long>(v)); };
    add(kCountFilterSeen, fc.seen);
    add(kCountFilterAcceptedVv, fc.acc_vv);
    add(kCountFilterAcceptedVt, fc.acc_vt);
    add(kCountFilterAcceptedEe, fc.acc_ee);
    add(kCountFilterRejectedRest, fc.rej_rest);
    add(kCountFilterRejectedDuplicate, fc.rej_dup);
    add(kCountFilterRejectedOwnership, fc.rej_own);
    add(kCountClusterCount, fc.clusters);
    add(kCountClusterOwnedContacts, fc.clusters);   // one owned contact per cluster
}

// Vertex part: candidate table (A41:4738-4814) + vertex-vertex response
// (A41:4816-4951) + vertex-triangle repulsion (A41:5109-5224).  The table is
// built on the first pass of a substep only (`build`); later passes read it
// back.
__device__ inline void fast_vertex(const ClothView& c, const SelfView& sv, const float4* P, int s, bool count, bool build,
                                   ClusterSet& cl, FilterCounts& fc) {
    const float cs = sv.cell;
    const F3 ps = f3(P[s]);
    unsigned acc = 0;
    unsigned* out = sv.nb_idx + static_cast<size_t>(s) * kNeighborSlots;
    if (build) {
        if (sv.use_vhash) {
            const int cx = cell_coord(ps.x, cs), cy = cell_coord(ps.y, cs), cz = cell_coord(ps.z, cs);
            for (int dz = -1; dz <= 1 && acc < kNeighborSlots; ++dz)
                for (int dy = -1; dy <= 1 && acc < kNeighborSlots; ++dy)
                    for (int dx = -1; dx <= 1 && acc < kNeighborSlots; ++dx) {
                        const int qx = cx + dx, qy = cy + dy, qz = cz + dz;
                        const int b = static_cast<int>(hash_cell(qx, qy, qz, sv.vh.buckets));
                        const int base = sv.vh.off[b], lim = min(sv.vh.count[b], kBucketSlots);
                        for (int k = 0; k < lim && acc < kNeighborSlots; ++k) {
                            const int o = sv.vh.entries[base + k];
                            const int4 oc = sv.vcell[o];
                            if (oc.x != qx || oc.y != qy || oc.z != qz || !vv_valid(c, sv, P, s, o)) continue;
                            out[acc++] = static_cast<unsigned>(o);
                        }
                    }
        } else {
            for (int o = 0; o < sv.nv && acc < kNeighborSlots; ++o)
                if (vv_valid(c, sv, P, s, o)) out[acc++] = static_cast<unsigned>(o);
        }
        if (acc >= kNeighborSlots && count) atomicAdd(&c.counts[kCountSelfOverflow], 1ull);
        sv.nb_count[s] = acc;
    } else {
        acc = min(sv.nb_count[s], static_cast<unsigned>(kNeighborSlots));
    }

    const float r = sv.radius, onset = sv.onset;
    const float thickness = r * 0.5f + r * 0.5f;
    unsigned cand = 0, hits = 0, soft = 0;
    if (P[s].w != 0.0f && acc > 0) {
        const F3 pc = ps, pp = f3(c.prev[s]);
        for (unsigned k = 0; k < acc; ++k) {
            const int o = static_cast<int>(out[k]);
            if (o < 0 || o >= sv.nv || o == s) continue;
            ++fc.seen;
            F3 resp;
            if (!vv_response(c, P, s, o, pc, pp, thickness, &resp)) continue;
            const F3 nn = mul(resp, 1.0f / fmaxf(len3(resp), kEps));
            const F3 mid = mul(pc + f3(P[o]), 0.5f);
            if (!cl.try_accept(filter_cluster_key(0, mid, nn, fmaxf(thickness * 2.5f, 1.0e-3f)), fc)) continue;
            ++fc.acc_vv;
            ++cand;
            accumulate(sv, P, s, resp);
            ++hits;
        }
    }
    if (sv.nt > 0) {
        cl.n = 0;
        auto pair = [&](int t) -> bool {
            const PairEval e = vt_eval(c, sv, P, s, t);
            if (e.rest) ++fc.rej_rest;
            if (!e.accepted) return false;
            if (!cl.try_accept(filter_cluster_key(1, e.point, f3(0.0f, 0.0f, 0.0f), fmaxf(onset * 12.0f, 1.0e-3f)), fc)) return true;
            ++fc.acc_vt;
            ++cand;
            if (e.apply) {
                apply_eval(sv, P, e);
                ++soft;
            }
            return true;
        };
        int accepted = 0;
        if (sv.use_thash) {
            const int cx = cell_coord(ps.x, cs), cy = cell_coord(ps.y, cs), cz = cell_coord(ps.z, cs);
            const int b = static_cast<int>(hash_cell(cx, cy, cz, sv.th.buckets));
            const int base = sv.th.off[b], lim = min(sv.th.count[b], kBucketSlots);
            // A bucket lists its keys (triangle * 64 + cell) in ascending order,
            // so the copies of one triangle are adjacent and the seen-id list
            // of A41:762-791 reduces to a comparison with the previous entry
            // (a bucket holds at most 32 entries, below its 64 slots).
            int prev_t = -1;
            for (int k = 0; k < lim && accepted < sv.max_nb; ++k) {
                const int t = sv.th.entries[base + k] / kMaxPrimCells;
                ++fc.seen;
                if (t == prev_t) { ++fc.rej_dup; continue; }
                prev_t = t;
                if (pair(t)) ++accepted;
            }
        } else {
            for (int t = 0; t < sv.nt && accepted < sv.max_nb; ++t) {
                ++fc.seen;
                if (pair(t)) ++accepted;
            }
        }
    }
    if (count) {
        if (cand) atomicAdd(&c.counts[kCountSelfCandidates], static_cast<unsigned long long>(cand));
        if (hits) atomicAdd(&c.counts[kCountExactImpulse], static_cast<unsigned long long>(hits));
        if (soft) atomicAdd(&c.counts[kCountSoft], static_cast<unsigned long long>(soft));
    }
}

// Edge part: A41:5595-5844.
__device__ inline void fast_edge(const ClothView& c, const SelfView& sv, const float4* P, int e, bool count, ClusterSet& cl,
                                 FilterCounts& fc) {
    const float onset = sv.onset, cs = sv.cell;
    unsigned cand = 0, hits = 0;
    auto pair = [&](int o) -> bool {
        const PairEval ev = ee_eval(c, sv, P, e, o);
        if (ev.rest) ++fc.rej_rest;
        if (!ev.accepted) return false;
        if (!cl.try_accept(filter_cluster_key(2, ev.point, f3(0.0f, 0.0f, 0.0f), fmaxf(onset * 12.0f, 1.0e-3f)), fc)) return true;
        ++fc.acc_ee;
        ++cand;
        if (ev.apply) {
            apply_eval(sv, P, ev);
            ++hits;
        }
        return true;
    };
    int accepted = 0;
    if (sv.use_ehash) {
        const int2 E = sv.edges[e];
        if (!edge_valid(sv, E)) return;
        const F3 mid = mul(f3(P[E.x]) + f3(P[E.y]), 0.5f);
        const int cx = cell_coord(mid.x, cs), cy = cell_coord(mid.y, cs), cz = cell_coord(mid.z, cs);
        const int b = static_cast<int>(hash_cell(cx, cy, cz, sv.eh.buckets));
        const int base = sv.eh.off[b], lim = min(sv.eh.count[b], kBucketSlots);
        int prev_e = -1;   // ascending keys: copies of one edge are adjacent
        for (int k = 0; k < lim && accepted < sv.max_nb; ++k) {
            const int o = sv.eh.entries[base + k] / kMaxPrimCells;
            ++fc.seen;
            if (o == prev_e) { ++fc.rej_dup; continue; }
            prev_e = o;
            if (pair(o)) ++accepted;
        }
    } else {
        for (int o = e + 1; o < sv.ne && accepted < sv.max_nb; ++o) {
            ++fc.seen;
            if (pair(o)) ++accepted;
        }
    }
    if (count) {
        if (cand) atomicAdd(&c.counts[kCountSelfCandidates], static_cast<unsigned long long>(cand));
        if (hits) atomicAdd(&c.counts[kCountEdgeEdge], static_cast<unsigned long long>(hits));
    }
}

// Threads [0, V) do the vertex part, [V, V + E) the edge part; every read is
// of positions that stay fixed until the apply kernel.
__global__ __launch_bounds__(256) void k_self_fast(const ClothView* __restrict__ cv, SelfView sv, int parity, int count, int build) {
    __shared__ unsigned keys[kClusterMaxLocal * 256];
    const int k = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const ClothView& c = *cv;
    const float4* P = cur_pos(c, parity);
    FilterCounts fc;
    ClusterSet cl{keys + threadIdx.x, 256, 0};
    if (k < sv.nv) fast_vertex(c, sv, P, k, count != 0, build != 0, cl, fc);
    else if (k < sv.nv + sv.ne) fast_edge(c, sv, P, k - sv.nv, count != 0, cl, fc);
    if (count) flush_filter_counts(c.counts, fc);
}

// ------------------------------------------------ strict mode, lane per entry
// The strict mode's candidate work (A41:4738-4814 table, 4816-4951 VV,
// 4990-5224 VT, 5595-5844 EE) with 32 lanes per primitive: lane l takes entry
// l of every bucket its vertex or edge queries (a bucket shows at most 32
// entries), instead of one thread walking them all.  Results equal the
// sequential walk: the neighbour table is filled in the walk's order (a
// ballot and a prefix count place each cell's valid entries after the ones
// already taken, the first 32 kept), the responses are fixed-point sums
// (order-free), and the cap on accepted VT / EE candidates is applied by the
// same prefix count (entry l runs only when fewer than max_nb entries before
// it were accepted, which is when the walk would have reached it).

// This lane's half-wave bits of a ballot.
__device__ __forceinline__ unsigned half_ballot(bool pred) {
    return static_cast<unsigned>(__ballot(pred) >> (threadIdx.x & 32u));
}

// Threads 32p .. 32p + 31 serve primitive p: vertices [0, V), then edges.
constexpr int kStrictHalves = 256 / 32;
__global__ __launch_bounds__(256) void k_self_strict(const ClothView* __restrict__ cv, SelfView sv, int parity, int count) {
    __shared__ unsigned s_tab[kStrictHalves][kNeighborSlots];
    const int p = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) >> 5);
    const int l = static_cast<int>(threadIdx.x & 31u);
    unsigned* tab = s_tab[threadIdx.x >> 5];
    const unsigned below = (1u << l) - 1u;
    const ClothView& c = *cv;
    const float4* P = cur_pos(c, parity);
    unsigned cand = 0, hits = 0, soft = 0, ee_hits = 0, overflow = 0;
    // Takes entry l of a walk's chunk when accepted and fewer than max_nb
    // entries before it were.
    auto capped = [&](const PairEval& e, bool in_range, int taken, unsigned& applied) {
        const unsigned m = half_ballot(e.accepted);
        if (in_range && e.accepted && taken + __popc(m & below) < sv.max_nb) {
            ++cand;
            if (e.apply) {
                apply_eval(sv, P, e);
                ++applied;
            }
        }
        return __popc(m);
    };
    if (p < sv.nv) {
        const int s = p;
        const float cs = sv.cell;
        const F3 ps = f3(P[s]);
        unsigned acc = 0;
        unsigned* out = sv.nb_idx + static_cast<size_t>(s) * kNeighborSlots;
        // candidate table, A41:4738-4788 (hash) / 4790-4814 (brute force)
        auto take = [&](bool ok, int o) {
            const unsigned m = half_ballot(ok);
            const unsigned pos = acc + static_cast<unsigned>(__popc(m & below));
            if (ok && pos < kNeighborSlots) {
                tab[pos] = static_cast<unsigned>(o);
                out[pos] = static_cast<unsigned>(o);
            }
            acc = min(acc + static_cast<unsigned>(__popc(m)), static_cast<unsigned>(kNeighborSlots));
        };
        if (sv.use_vhash) {
            // The walk's 27 cells (dz outer, then dy, then dx) at once: lane l
            // < 27 takes cell l and its bucket's first min(count, 32) entries;
            // a prefix sum lays them out in walk order (cell, then entry), and
            // the half-wave takes that list 32 items at a time until the table
            // is full.  Item w belongs to the last lane whose start is <= w.
            const int cx = cell_coord(ps.x, cs), cy = cell_coord(ps.y, cs), cz = cell_coord(ps.z, cs);
            int qx = 0, qy = 0, qz = 0, first = 0, n = 0;
            if (l < 27) {
                qx = cx + l % 3 - 1;
                qy = cy + (l / 3) % 3 - 1;
                qz = cz + l / 9 - 1;
                const int b = static_cast<int>(hash_cell(qx, qy, qz, sv.vh.buckets));
                first = sv.vh.off[b];
                n = min(sv.vh.count[b], kBucketSlots);
            }
            int incl = n;
            for (int d = 1; d < 32; d <<= 1) {
                const int o = __shfl_up(incl, d, 32);
                if (l >= d) incl += o;
            }
            const int start = incl - n;
            const int total = __shfl(incl, 31, 32);
            for (int base = 0; base < total && acc < kNeighborSlots; base += 32) {
                const int w = base + l;
                int own = 0;
                for (int step = 16; step > 0; step >>= 1)
                    if (__shfl(start, own + step, 32) <= w) own += step;
                const int at = __shfl(first, own, 32) + (w - __shfl(start, own, 32));
                const int ox = __shfl(qx, own, 32), oy = __shfl(qy, own, 32), oz = __shfl(qz, own, 32);
                bool ok = false;
                int o = 0;
                if (w < total) {
                    o = sv.vh.entries[at];
                    const int4 oc = sv.vcell[o];
                    ok = oc.x == ox && oc.y == oy && oc.z == oz && vv_valid(c, sv, P, s, o);
                }
                take(ok, o);
            }
        } else {
            for (int o0 = 0; o0 < sv.nv && acc < kNeighborSlots; o0 += 32) {
                const int o = o0 + l;
                take(o < sv.nv && vv_valid(c, sv, P, s, o), o);
            }
        }
        if (l == 0) {
            sv.nb_count[s] = acc;
            overflow = acc >= kNeighborSlots ? 1u : 0u;
        }
        wave_lds_sync();
        // vertex-vertex: lane l takes table entry l
        const float r = sv.radius;
        const float thickness = r * 0.5f + r * 0.5f;
        if (P[s].w != 0.0f && static_cast<unsigned>(l) < acc) {
            const int o = static_cast<int>(tab[l]);
            if (o >= 0 && o < sv.nv && o != s) {
                ++cand;
                F3 resp;
                if (vv_response(c, P, s, o, ps, f3(c.prev[s]), thickness, &resp)) {
                    accumulate(sv, P, s, resp);
                    ++hits;
                }
            }
        }
        // vertex-triangle: the own triangle bucket (A41:5109-5182) or every triangle
        if (sv.nt > 0) {
            if (sv.use_thash) {
                const int cx = cell_coord(ps.x, cs), cy = cell_coord(ps.y, cs), cz = cell_coord(ps.z, cs);
                const int b = static_cast<int>(hash_cell(cx, cy, cz, sv.th.buckets));
                const int base = sv.th.off[b], lim = min(sv.th.count[b], kBucketSlots);
                PairEval e;
                if (l < lim) e = vt_eval(c, sv, P, s, sv.th.entries[base + l] / kMaxPrimCells);
                capped(e, l < lim, 0, soft);
            } else {
                for (int t0 = 0, taken = 0; t0 < sv.nt && taken < sv.max_nb; t0 += 32) {
                    PairEval e;
                    if (t0 + l < sv.nt) e = vt_eval(c, sv, P, s, t0 + l);
                    taken += capped(e, t0 + l < sv.nt, taken, soft);
                }
            }
        }
    } else if (p < sv.nv + sv.ne) {
        // edge-edge: the own edge bucket (A41:5724-5802) or every later edge
        const int e0 = p - sv.nv;
        if (sv.use_ehash) {
            const int2 E = sv.edges[e0];
            if (edge_valid(sv, E)) {
                const float cs = sv.cell;
                const F3 mid = mul(f3(P[E.x]) + f3(P[E.y]), 0.5f);
                const int cx = cell_coord(mid.x, cs), cy = cell_coord(mid.y, cs), cz = cell_coord(mid.z, cs);
                const int b = static_cast<int>(hash_cell(cx, cy, cz, sv.eh.buckets));
                const int base = sv.eh.off[b], lim = min(sv.eh.count[b], kBucketSlots);
                PairEval e;
                if (l < lim) e = ee_eval(c, sv, P, e0, sv.eh.entries[base + l] / kMaxPrimCells);
                capped(e, l < lim, 0, ee_hits);
            }
        } else {
            for (int o0 = e0 + 1, taken = 0; o0 < sv.ne && taken < sv.max_nb; o0 += 32) {
                PairEval e;
                if (o0 + l < sv.ne) e = ee_eval(c, sv, P, e0, o0 + l);
                taken += capped(e, o0 + l < sv.ne, taken, ee_hits);
            }
        }
    }
    if (count) {
        wave_add(c.counts, kCountSelfCandidates, cand);
        wave_add(c.counts, kCountExactImpulse, hits);
        wave_add(c.counts, kCountSoft, soft);
        wave_add(c.counts, kCountEdgeEdge, ee_hits);
        wave_add(c.counts, kCountSelfOverflow, overflow);
    }
}

// A41:5226-5439: a vertex that sits inside three or more nearby triangles
// (searched in the 27 triangle-hash cells around it) is pushed out along
// their penetration-weighted mean normal.  Runs on the last pass of a fast
// substep, before the averaging.
constexpr int kIslandBlock = 128;
__global__ __launch_bounds__(kIslandBlock) void k_self_island(const ClothView* __restrict__ cv, SelfView sv, int parity, int count) {
    __shared__ int accepted_tris[kIslandMaxContacts * kIslandBlock];
    const int v = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= sv.nv) return;
    const ClothView& c = *cv;
    const float4* P = cur_pos(c, parity);
    if (P[v].w == 0.0f) return;
    const float r = sv.radius;
    if (r <= 0.0f) return;
    int* seen = accepted_tris + threadIdx.x;
    const F3 p = f3(P[v]), pp = f3(c.prev[v]);
    unsigned guarded = 0;
    if (!finite3(p) || !finite3(pp)) {
        if (count) atomicAdd(&c.counts[kCountIslandGuarded], 1ull);
        return;
    }
    const float cs = sv.cell;
    const int max_contacts = sv.island_contacts;
    const int scan_limit = max_contacts * 10;
    const int cx = cell_coord(p.x, cs), cy = cell_coord(p.y, cs), cz = cell_coord(p.z, cs);
    const F3 rv = f3(c.rest[v]);
    int scanned = 0, contacts = 0;
    float pen_sum = 0.0f, sample_sum = 0.0f, max_pen = 0.0f;
    F3 axis = f3(0.0f, 0.0f, 0.0f), best = f3(0.0f, 0.0f, 1.0f);
    for (int dz = -1; dz <= 1 && contacts < max_contacts && scanned < scan_limit; ++dz)
        for (int dy = -1; dy <= 1 && contacts < max_contacts && scanned < scan_limit; ++dy)
            for (int dx = -1; dx <= 1 && contacts < max_contacts && scanned < scan_limit; ++dx) {
                const int b = static_cast<int>(hash_cell(cx + dx, cy + dy, cz + dz, sv.th.buckets));
                const int base = sv.th.off[b], lim = min(sv.th.count[b], kBucketSlots);
                for (int k = 0; k < lim && contacts < max_contacts && scanned < scan_limit; ++k) {
                    ++scanned;
                    const int t = sv.th.entries[base + k] / kMaxPrimCells;
                    bool dup = false;
                    for (int i = 0; i < contacts; ++i)
                        if (seen[i * kIslandBlock] == t) { dup = true; break; }
                    if (dup) continue;
                    const int4 tr = sv.tris[t];
                    if (!tri_valid(sv, tr)) continue;
                    const int ia = tr.x, ib = tr.y, ic = tr.z;
                    if (v == ia || v == ib || v == ic) continue;
                    {
                        const F3 rd = rv - closest_point_on_triangle(rv, f3(c.rest[ia]), f3(c.rest[ib]), f3(c.rest[ic]));
                        if (dot(rd, rd) < sv.rest_skip * sv.rest_skip) continue;
                    }
                    const F3 a = f3(P[ia]), bb = f3(P[ib]), cc = f3(P[ic]);
                    const F3 ap = f3(c.prev[ia]), bp = f3(c.prev[ib]), cp = f3(c.prev[ic]);
                    if (!finite3(a) || !finite3(bb) || !finite3(cc) || !finite3(ap) || !finite3(bp) || !finite3(cp)) {
                        ++guarded;
                        continue;
                    }
                    const F3 q = closest_point_on_triangle(p, a, bb, cc);
                    F3 d = p - q;
                    const float dsq = dot(d, d);
                    float dist = sqrt_eps(dsq);
                    if (dist >= r) continue;
                    F3 nn = f3(0.0f, 0.0f, 1.0f);
                    if (dsq > kEps) nn = mul(d, 1.0f / dist);
                    else {
                        nn = cross(bb - a, cc - a);
                        const float nsq = dot(nn, nn);
                        if (nsq > kEps) nn = mul(nn, rsqrt_exact(nsq));
                        else {
                            const F3 rd = rv - closest_point_on_triangle(rv, f3(c.rest[ia]), f3(c.rest[ib]), f3(c.rest[ic]));
                            const float rsq = dot(rd, rd);
                            if (rsq > kEps) nn = mul(rd, rsqrt_exact(rsq));
                        }
                        d = mul(nn, dist);
                        dist = 0.0f;
                    }
                    if (!finite3(nn)) { ++guarded; continue; }
                    const F3 pd = pp - closest_point_on_triangle(pp, ap, bp, cp);
                    if (dot(pd, nn) > r * 0.25f && dot(d, nn) > -r * 0.10f) continue;
                    float wa = 1.0f / 3.0f, wb = wa, wc = wa;
                    triangle_barycentric(q, a, bb, cc, &wa, &wb, &wc);
                    const float pen = r - dist;
                    if (pen <= 0.0f) continue;
                    seen[contacts * kIslandBlock] = t;
                    ++contacts;
                    pen_sum += pen;
                    sample_sum += c.rest[ia].w * wa + c.rest[ib].w * wb + c.rest[ic].w * wc;
                    axis = axis + mul(nn, pen);
                    if (pen > max_pen) { max_pen = pen; best = nn; }
                }
            }
    if (contacts > 0) {
        bool applied = false;
        float dl = 0.0f;
        F3 agg = best;
        if (contacts < 3) ++guarded;
        else {
            const float avg = pen_sum / static_cast<float>(contacts);
            if (avg < r * 0.04f) ++guarded;
            else {
                const float axis_len = len3(axis);
                agg = axis_len > kEps ? mul(axis, 1.0f / axis_len) : best;
                const float source_w = c.rest[v].w;
                const float sample_w = sample_sum / static_cast<float>(contacts);
                const float share = source_w / fmaxf(source_w + sample_w, kEps);
                const float max_delta = fmaxf(r * 0.30f, 1.0e-5f);
                dl = fminf(avg * 0.50f * share, max_delta);
                if (!finite3(agg) || !isfinite(dl) || dl <= kEps) ++guarded;
                else applied = true;
            }
        }
        if (applied) accumulate(sv, P, v, mul(agg, dl));
        if (count) {
            atomicAdd(&c.counts[kCountIslandCandidates], static_cast<unsigned long long>(contacts));
            if (applied) {
                atomicAdd(&c.counts[kCountIslandClusters], 1ull);
                atomicAdd(&c.counts[kCountIslandVertexRefs], static_cast<unsigned long long>(contacts));
                atomicAdd(&c.counts[kCountIslandApplied], 1ull);
                // A41:852-863
                atomicMax(&c.counts[kCountIslandMaxDeltaMicro], static_cast<unsigned long long>(fminf(dl, 1024.0f) * 1000000.0f));
            }
        }
    }
    if (count && guarded) atomicAdd(&c.counts[kCountIslandGuarded], static_cast<unsigned long long>(guarded));
}

// Cache counters of a fast substep (A41:831-850): the table built once serves
// every pass.
__global__ void k_self_cache_counts(unsigned long long* counts, unsigned long long hits, unsigned long long misses) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    if (hits) atomicAdd(&counts[kCountFilterCacheHits], hits);
    if (misses) atomicAdd(&counts[kCountFilterCacheMisses], misses);
}

// A41:4953-4988: average, clamp, apply; the accumulators are cleared for the
// next pass.
__global__ __launch_bounds__(256) void k_self_apply(const ClothView* __restrict__ cv, SelfView sv, int parity) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    const ClothView& c = *cv;
    float alen = 0.0f;
    const int w = i < sv.nv ? sv.accw[i] : 0;
    if (w > 0) {
        float4* P = parity ? c.pos1 : c.pos0;
        const double inv = 1.0 / kFixedScale;
        const F3 sum = f3(static_cast<float>(static_cast<double>(static_cast<long long>(sv.acc[3 * i + 0])) * inv),
                          static_cast<float>(static_cast<double>(static_cast<long long>(sv.acc[3 * i + 1])) * inv),
                          static_cast<float>(static_cast<double>(static_cast<long long>(sv.acc[3 * i + 2])) * inv));
        sv.acc[3 * i + 0] = 0ull;
        sv.acc[3 * i + 1] = 0ull;
        sv.acc[3 * i + 2] = 0ull;
        sv.accw[i] = 0;
        F3 avg = mul(sum, 1.0f / static_cast<float>(w));
        const float max_delta = fmaxf(sv.radius * 0.35f, 1.0e-4f);
        const float dl = len3(avg);
        if (dl > max_delta) avg = mul(avg, max_delta / dl);
        alen = len3(avg);
        const float4 p = P[i];
        P[i] = with_w(f3(p) + avg, p.w);
        const float4 q = c.prev[i];
        c.prev[i] = with_w(f3(q) + mul(avg, 0.55f), 0.0f);
    }
    // the largest averaged delta (order-free): one atomic per wave
    wave_max_nonneg(c.max_smoothed, alen);
}

}  // namespace

// ------------------------------------------------------------------ host
SelfCollision::~SelfCollision() {
    for (auto& a : allocations_) device_free(a.first, a.second);
}

void* SelfCollision::alloc_bytes(size_t bytes) {
    const size_t n = std::max<size_t>(bytes, 256);
    void* p = device_alloc(n, "self collision allocation");
    if (p) allocations_.emplace_back(p, n);
    return p;
}

// A41:6784-6994 sizing: cell = max(1.8 r, 1 mm); buckets = next_pow2(max(1024, 4n)).
bool SelfCollision::prepare(const SsblXpbdConfig& cfg, const std::vector<int>& edges, const std::vector<int>& tris, int nv,
                            unsigned long long* counts, const std::vector<int>& vorder, const std::vector<int>& torder,
                            const std::vector<int>& eorder) {
    ready_ = false;
    const float req = std::max(cfg.self_collision_distance, 0.0f);
    const float r = req > 0.0f ? req : std::max(std::max(cfg.cloth_thickness, cfg.collision_margin), 0.0f);
    if (!cfg.self_collision || nv <= 0 || r <= 0.0f) return true;
    SelfView& s = sv_;
    s = SelfView{};
    s.nv = nv;
    s.nt = cfg.triangle_count;
    s.ne = cfg.edge_count;
    s.radius = r;
    s.onset = std::max(r * 1.8f, r + 1.0e-7f);
    s.rest_skip = std::max(s.onset * 1.25f, r * 2.5f);
    s.cell = std::max(r * 1.8f, 1.0e-3f);
    s.prim_margin = r * 1.8f;
    s.max_nb = std::max(1, std::min(cfg.max_self_collision_neighbors, kBucketSlots));
    s.fast = cfg.self_collision_mode == 1 ? 1 : 0;   // A41:140
    if (s.fast && s.max_nb > 16) s.max_nb = 16;       // A41:5133-5135, 5749-5751
    {
        const long long mn = std::max(cfg.max_self_collision_neighbors, 1);
        s.island_contacts = static_cast<int>(std::max<long long>(std::min<long long>(2 * mn, kIslandMaxContacts), 3));
    }
    s.counts = counts;
    s.use_vhash = nv >= 128 ? 1 : 0;
    s.use_thash = s.nt >= 64 ? 1 : 0;
    s.use_ehash = s.ne >= 64 ? 1 : 0;
    passes_ = std::max(1, std::min(cfg.fast_self_collision_passes, 8));
    interval_ = std::max(cfg.self_collision_interval, 1);
    // the shared bucket table: vertex, triangle, edge buckets in that order,
    // item slots likewise
    long long buckets = 0, items = 0;
    std::vector<std::pair<HashView*, long long>> used;   // hash, first item slot
    auto plan = [&](HashView& h, int n, long long capacity) {
        h.buckets = next_pow2(std::max(1024, n * 4));
        h.capacity = static_cast<int>(capacity);
        h.base = static_cast<int>(buckets);
        used.emplace_back(&h, items);
        buckets += h.buckets;
        items += capacity;
    };
    if (s.use_vhash) plan(s.vh, nv, nv);
    if (s.use_thash) plan(s.th, s.nt, static_cast<long long>(s.nt) * kMaxPrimCells);
    if (s.use_ehash) plan(s.eh, s.ne, static_cast<long long>(s.ne) * kMaxPrimCells);
    if (buckets >= (1ll << 30) || items >= (1ll << 31)) return fail("self-collision hashes too large");
    s.all_buckets = static_cast<int>(buckets);
    s.hash_items = static_cast<int>(items);
    if (!used.empty()) {
        const size_t cap = static_cast<size_t>(std::max<long long>(items, 1));
        s.all_count = static_cast<int*>(alloc_bytes(sizeof(int) * buckets));
        s.all_off = static_cast<int*>(alloc_bytes(sizeof(int) * (buckets + 1)));
        int* slot = static_cast<int*>(alloc_bytes(sizeof(int) * cap));
        int* raw = static_cast<int*>(alloc_bytes(sizeof(int) * cap));
        int* rawb = static_cast<int*>(alloc_bytes(sizeof(int) * cap));
        int* entries = static_cast<int*>(alloc_bytes(sizeof(int) * cap));
        if (!s.all_count || !s.all_off || !slot || !raw || !rawb || !entries) return false;
        for (auto& [h, first_item] : used) {
            h->count = s.all_count + h->base;
            h->off = s.all_off + h->base;
            h->slot = slot + first_item;
            h->raw = raw;
            h->rawb = rawb;
            h->entries = entries;
        }
    }
    s.vh_blocks = s.use_vhash ? static_cast<int>(grid_for(nv, 256)) : 0;
    s.th_blocks = s.use_thash ? static_cast<int>(grid_for(static_cast<long long>(s.nt) * kPrimLanes, 256)) : 0;
    s.scan_tmp = static_cast<int*>(alloc_bytes(sizeof(int) * scan_tmp_ints(std::max(s.all_buckets, 1))));
    s.vcell = static_cast<int4*>(alloc_bytes(sizeof(int4) * nv));
    s.nb_count = static_cast<unsigned*>(alloc_bytes(sizeof(unsigned) * nv));
    s.nb_idx = static_cast<unsigned*>(alloc_bytes(sizeof(unsigned) * static_cast<size_t>(nv) * kNeighborSlots));
    s.acc = static_cast<unsigned long long*>(alloc_bytes(sizeof(unsigned long long) * 3 * static_cast<size_t>(nv)));
    s.accw = static_cast<int*>(alloc_bytes(sizeof(int) * nv));
    int4* d_tris = static_cast<int4*>(alloc_bytes(sizeof(int4) * std::max(s.nt, 1)));
    int2* d_edges = static_cast<int2*>(alloc_bytes(sizeof(int2) * std::max(s.ne, 1)));
    if (!s.scan_tmp || !s.vcell || !s.nb_count || !s.nb_idx || !s.acc || !s.accw || !d_tris || !d_edges) return false;
    std::vector<int4> ht(s.nt);
    for (int t = 0; t < s.nt; ++t) ht[t] = make_int4(tris[3 * t], tris[3 * t + 1], tris[3 * t + 2], 0);
    std::vector<int2> he(s.ne);
    for (int e = 0; e < s.ne; ++e) he[e] = make_int2(edges[2 * e], edges[2 * e + 1]);
    if (s.nt > 0 && !hip_ok(hipMemcpy(d_tris, ht.data(), sizeof(int4) * s.nt, hipMemcpyHostToDevice), "self triangle upload")) return false;
    if (s.ne > 0 && !hip_ok(hipMemcpy(d_edges, he.data(), sizeof(int2) * s.ne, hipMemcpyHostToDevice), "self edge upload")) return false;
    s.tris = d_tris;
    s.edges = d_edges;
    // visiting orders (a missing or mis-sized one: mesh order)
    auto upload_order = [&](const std::vector<int>& order, int n, const int** out) {
        *out = nullptr;
        if (n <= 0 || order.size() != static_cast<size_t>(n)) return true;
        int* d = static_cast<int*>(alloc_bytes(sizeof(int) * static_cast<size_t>(n)));
        if (!d || !hip_ok(hipMemcpy(d, order.data(), sizeof(int) * static_cast<size_t>(n), hipMemcpyHostToDevice), "self order upload"))
            return false;
        *out = d;
        return true;
    };
    if (!upload_order(vorder, nv, &s.vorder) || !upload_order(torder, s.nt, &s.torder) || !upload_order(eorder, s.ne, &s.eorder))
        return false;
    if (!hip_ok(hipMemset(s.acc, 0, sizeof(unsigned long long) * 3 * static_cast<size_t>(nv)), "self reset") ||
        !hip_ok(hipMemset(s.accw, 0, sizeof(int) * nv), "self reset") ||
        !hip_ok(hipMemset(s.nb_count, 0, sizeof(unsigned) * nv), "self reset") ||
        !hip_ok(hipMemset(s.vcell, 0, sizeof(int4) * nv), "self reset"))
        return false;
    ready_ = true;
    return true;
}

SelfBatch* SelfCollision::single_batch(const ClothView&, const ClothView* dview) {
    dview_ = dview;
    return ready_ ? this : nullptr;
}

bool SelfCollision::run(const LaunchCtx& ctx, int parity, int substep, int substeps, bool count) {
    if (!(((substep + 1) % interval_) == 0 || substep == substeps - 1)) return true;
    return run_for(ctx.stream, dview_, parity, count);
}

bool SelfCollision::run_for(hipStream_t st, const ClothView* cv, int parity, bool count) {
    if (!ready_) return true;
    const SelfView& s = sv_;
    if (s.all_buckets > 0) {
        // the three hashes in one shared table: fill, count, scan, scatter, rank
        TimedLaunch t("self_hash", st);
        const dim3 block(256);
        const unsigned eh_blocks = s.use_ehash ? grid_for(static_cast<long long>(s.ne) * kPrimLanes, 256) : 0u;
        const dim3 grid(static_cast<unsigned>(s.vh_blocks + s.th_blocks) + eh_blocks);
        const HashView& any = s.use_vhash ? s.vh : (s.use_thash ? s.th : s.eh);
        launch_k(k_fill_i32, dim3(grid_for(s.all_buckets, 256)), block, 0, st, s.all_count, 0, s.all_buckets);
        launch_k(k_self_hash, grid, block, 0, st, cv, s, parity, 0, count ? 1 : 0);
        exclusive_scan(st, s.all_count, s.all_buckets, s.all_off, s.scan_tmp);
        launch_k(k_self_hash, grid, block, 0, st, cv, s, parity, 1, 0);
        // about one item per vertex and 27 per primitive (3 cells a side): one thread each
        const long long rank_threads = std::min<long long>(
            s.hash_items, (s.use_vhash ? s.nv : 0) + 32ll * ((s.use_thash ? s.nt : 0) + (s.use_ehash ? s.ne : 0)));
        // and a wave per group of 64 buckets for the big ones, four waves per block
        const unsigned item_blocks = grid_for(std::max<long long>(rank_threads, 1), 256);
        const unsigned group_blocks = grid_for((static_cast<long long>(s.all_buckets) + 63) / 64, 4);
        launch_k(k_self_rank, dim3(item_blocks + group_blocks), block, 0, st, s, any.raw, any.rawb, any.entries, count ? 1 : 0,
                 static_cast<int>(item_blocks));
    }
    const dim3 contacts_grid(grid_for(static_cast<long long>(s.nv) + s.ne, 256));
    if (s.fast && count) {
        // A41:9291-9301: the fast mode builds the table once per substep
        launch_k(k_self_cache_counts, dim3(1), dim3(1), 0, st, s.counts,
                 static_cast<unsigned long long>(s.nv) * static_cast<unsigned long long>(passes_ - 1),
                 static_cast<unsigned long long>(s.nv));
    }
    for (int pass = 0; pass < passes_; ++pass) {
        {
            TimedLaunch t("self_contacts", st);
            if (s.fast)
                launch_k(k_self_fast, contacts_grid, dim3(256), 0, st, cv, s, parity, count ? 1 : 0, pass == 0 ? 1 : 0);
            else
                launch_k(k_self_strict, dim3(grid_for((static_cast<long long>(s.nv) + s.ne) * 32, 256)), dim3(256), 0, st, cv, s,
                         parity, count ? 1 : 0);
        }
        if (s.fast && pass == passes_ - 1 && s.nt > 0 && s.use_thash) {
            TimedLaunch t("self_island", st);
            launch_k(k_self_island, dim3(grid_for(s.nv, kIslandBlock)), dim3(kIslandBlock), 0, st, cv, s, parity, count ? 1 : 0);
        }
        {
            TimedLaunch t("self_apply", st);
            launch_k(k_self_apply, dim3(grid_for(s.nv, 256)), dim3(256), 0, st, cv, s, parity);
        }
    }
    return hip_ok(hipGetLastError(), "launch self collision");
}

bool SelfCollision::download_cells(int* cells, int nv) {
    if (!ready_ || !sv_.use_vhash) return fail("no self-collision vertex hash");
    std::vector<int4> h(nv);
    if (!hip_ok(hipMemcpy(h.data(), sv_.vcell, sizeof(int4) * nv, hipMemcpyDeviceToHost), "download hash cells")) return false;
    for (int i = 0; i < nv; ++i) { cells[3 * i] = h[i].x; cells[3 * i + 1] = h[i].y; cells[3 * i + 2] = h[i].z; }
    return true;
}

bool SelfCollision::download_neighbors(unsigned* counts, unsigned* indices, int nv) {
    if (!ready_) return fail("no self-collision table");
    return hip_ok(hipMemcpy(counts, sv_.nb_count, sizeof(unsigned) * nv, hipMemcpyDeviceToHost), "download neighbours") &&
           hip_ok(hipMemcpy(indices, sv_.nb_idx, sizeof(unsigned) * static_cast<size_t>(nv) * kNeighborSlots, hipMemcpyDeviceToHost),
                  "download neighbours");
}

}  // namespace ssbl

namespace ssbl {

SelfBatch* SelfBatchHolder::get(const std::vector<Solver*>& members, const std::vector<ClothView>&, const ClothView* dviews) {
    runner_.entries.clear();
    for (size_t m = 0; m < members.size(); ++m) {
        SelfCollision* sc = solver_self(members[m]);
        if (sc && sc->enabled()) runner_.entries.push_back(Entry{sc, dviews + m, sc->interval()});
    }
    return runner_.entries.empty() ? nullptr : &runner_;
}

bool SelfBatchHolder::Runner::run(const LaunchCtx& ctx, int parity, int substep, int substeps, bool count) {
    for (const Entry& e : entries) {
        if (!(((substep + 1) % e.substep_interval) == 0 || substep == substeps - 1)) continue;
        if (!e.self->run_for(ctx.stream, e.dview, parity, count)) return false;
    }
    return true;
}

}  // namespace ssbl
