// xpbd_dynamic.hpp -- dynamic colliders a cloth collides with: its own
// triangle soup and particles, or a global dynamic scene holding every cloth
// of a session, with their spatial hashes built on the device.
//
// Reference: uploads A41:8319-8420 and 8575-8697, bounds and hashes
// A41:3694-3932, 4325-4352 and 6220-6449, global scene A41:6451-6782 and
// 8733-8789, contacts A41:3934-4511 (unsigned).
//
// The reference fills its buckets with atomics, so the order of a bucket (and
// which particles a full 16-slot bucket keeps) depends on scheduling.  Here
// every bucket is a counting-sorted segment whose entries are placed by rank:
// triangle buckets hold (triangle * 64 + cell index) keys in ascending order,
// particle buckets their first 16 particles in ascending order, and the list
// of triangles too large to hash is in ascending order (a prefix sum of
// flags).  A query therefore visits the same primitives in the same order as
// the CPU oracle.
#pragma once

#include <vector>

#include "xpbd_device.hpp"
#include "xpbd_math.hpp"

namespace ssbl {

constexpr int kDynParticleSlots = 16;       // A41:26
constexpr int kDynMaxHashCells = 64;        // A41:79
constexpr int kDynParticleHashMin = 128;    // A41:25
constexpr int kDynTriangleHashMin = 64;     // A41:39
constexpr int kDynMinBuckets = 1024;        // A41:27, 40
constexpr int kMaxLayerClasses = 8;         // layer ranges of a gathered scene's triangle hash

enum DynCountSlot {
    kCountDynParticleCandidates = 4, kCountDynParticleContacts = 5, kCountDynParticleOverflow = 6, kCountTrianglePairs = 7,
    kCountDynTriCandidates = 34, kCountDynTriBucketOverflow = 35, kCountDynTriLarge = 36, kCountDynTriAabbRejects = 37
};

// One triangle's bounds in one 64-byte record, so that a query's canonical
// cell test, box test and sphere test read one memory segment: the hashed
// cell span (cmin, cmax), the box (lo, hi), the centre and the bounding radius.
struct __align__(64) TriBox {
    int4 c0;     // cmin.x, cmin.y, cmin.z, cmax.x
    int4 c1;     // cmax.y, cmax.z, lo.x bits, lo.y bits
    float4 b0;   // lo.z, hi.x, hi.y, hi.z
    float4 s;    // centre xyz, bounding radius
};

__device__ __forceinline__ TriBox gload(const TriBox* p) {
    TriBox b;
    b.c0 = gload(&p->c0);
    b.c1 = gload(&p->c1);
    b.b0 = gload(&p->b0);
    b.s = gload(&p->s);
    return b;
}

__device__ __forceinline__ int4 box_cmin(const TriBox& b) { return make_int4(b.c0.x, b.c0.y, b.c0.z, 0); }
__device__ __forceinline__ int4 box_cmax(const TriBox& b) { return make_int4(b.c0.w, b.c1.x, b.c1.y, 0); }
__device__ __forceinline__ float4 box_lo(const TriBox& b) { return make_float4(__int_as_float(b.c1.z), __int_as_float(b.c1.w), b.b0.x, 0.0f); }
__device__ __forceinline__ float4 box_hi(const TriBox& b) { return make_float4(b.b0.y, b.b0.z, b.b0.w, 0.0f); }

// A visible particle bucket entry (the first 16 particles of a bucket by
// index) in one 48-byte record: its particle cell and index, {source, layer},
// and the particle's position and radius (fixed while the set's hash is
// current), so a query reads an entry in one round trip.
struct PartEntry {
    int4 ck;     // cell x, y, z, particle index
    int4 sl;     // source id, layer, 0, 0
    float4 p;    // position, radius
};

__device__ __forceinline__ PartEntry gload(const PartEntry* e) {
    PartEntry r;
    r.ck = gload(&e->ck);
    r.sl = gload(&e->sl);
    r.p = gload(&e->p);
    return r;
}

// What the contact pass reads; lives in device memory (one per collider set).
struct DynView {
    int nt;
    int tri_hash, tri_buckets;
    float tri_cell;
    const float4* tv;        // 3 per triangle (.w unused)
    const int* tsrc;         // null: no source filter
    const int* tlayer;
    const TriBox* box;       // bounds of every triangle (cell span only in hashed sets)
    const int* tri_keys;     // triangle * 64 + cell index, ascending per bucket
    const int2* tri_ksl;     // {source id, layer} of each key's triangle (filtered sets)
    const int4* tri_bin;     // per bucket: {offset, count, min layer (with layer ranges: the entries' one cell tag | 0), the one source of all entries | -1 mixed | -2 empty}
    const int4* tri_ent;     // unranked sets: {key, source, layer, cell tag} per entry, in no particular order
    const int* large;        // triangles spanning more than 64 cells, ascending
    const int* nlarge;       // their count (device)
    int np;
    int part_hash, part_buckets;
    float part_cell;
    const float4* pp;        // position, .w = radius as uploaded
    const int* psrc;
    const int* player;
    const int* pcount;       // buckets
    const int* poff;         // buckets + 1
    const int4* pbin;        // per bucket, over its visible entries: {offset, count, min layer, one source | -1 | -2}
    const PartEntry* pent;   // first min(count, 16) particles of a bucket by index: cell, index, source, layer, position
    const int4* pcell;
    // Coarse particle summary: per bucket of coarse cells (4 particle cells
    // wide), {smallest layer, the one source | -1 mixed | -2 empty} over every
    // particle whose reach box (position +- coarse_reach) overlaps the cell.
    // A query whose own coarse cell has no particle its filter accepts cannot
    // find a contact and skips the particle search.
    const int2* coarse;      // null: no summary
    const int* coarse_off;   // device flag: a particle position was not finite, do not use the summary
    int coarse_buckets;
    float coarse_cell, coarse_reach, max_radius;
    // signed (indexed) triangles, A41:8433-8590: their contacts take the
    // place of the unsigned ones; centres and particle positions of the
    // previous upload give the colliders' velocities for one step
    int signed_set, tri_motion, part_motion, pad5;
    const float4* pctr;      // previous triangle centres
    const float4* pprev;     // previous particle positions (.w = radius)
    // A gathered scene whose triangles carry 2 .. kMaxLayerClasses distinct
    // layers keeps one bucket range per layer, ascending: a triangle of
    // layer tri_class_layer[k] is entered in range k, buckets
    // [tri_class_base[k], tri_class_base[k + 1]) (sized by the layer's share
    // of the triangles), so a query that accepts only lower layers probes
    // only the ranges below its own and walks no entry its filter rejects by
    // layer.  tri_classes 1: one range, hash_cell over tri_buckets.
    int tri_classes;
    int tri_class_base[kMaxLayerClasses + 1];
    int tri_class_layer[kMaxLayerClasses];
};

// Nonzero tag of a cell.  The entries of a triangle hash with layer ranges
// carry their cell's tag, and a bucket bin there holds the tag its entries
// share (0: several cells): a query from a cell of another tag cannot pass
// the canonical-cell test on any of them (a triangle whose span holds the
// query's cell is listed under that cell, in the query's own bucket), so it
// skips the bucket with the bin load alone.
__device__ __forceinline__ int cell_tag(int x, int y, int z) {
    const unsigned h = static_cast<unsigned>(x) * 0x8da6b343u ^ static_cast<unsigned>(y) * 0xd8163841u ^ static_cast<unsigned>(z) * 0xcb1ab31fu;
    return static_cast<int>(h | 1u);
}

// Bucket of cell (x, y, z) in the layer range [lo, hi) of a triangle hash
// with `classes` ranges (multiply-shift into the range for two or more).
__device__ __forceinline__ int tri_bucket(int classes, int buckets, int lo, int hi, int x, int y, int z) {
    if (classes <= 1) return static_cast<int>(hash_cell(x, y, z, buckets));
    const unsigned h = (static_cast<unsigned>(x) * 73856093u ^ static_cast<unsigned>(y) * 19349663u ^ static_cast<unsigned>(z) * 83492791u) *
                       2654435761u;
    return lo + static_cast<int>((static_cast<unsigned long long>(h) * static_cast<unsigned>(hi - lo)) >> 32);
}

// One contributing cloth of a device-gathered scene.
struct GatherSource {
    const float4* pos;       // current positions
    const int4* tris;        // its triangles
    const int* order;        // the order the hash build visits them (a permutation of 0 .. nt - 1), or null
    int nv, nt;
    int pbase, tbase;        // first particle / triangle in the scene
    float radius;
    int source_id, layer;
};

class DynamicColliders {
public:
    DynamicColliders() = default;
    DynamicColliders(const DynamicColliders&) = delete;
    DynamicColliders& operator=(const DynamicColliders&) = delete;
    ~DynamicColliders();

    // Host arrays (copied).  `fine` selects the scene's finer triangle cells
    // (A41:6590-6601).  Both sync the library stream before replacing data.
    bool set_triangles(hipStream_t st, const float* soup, int n, const int* src, const int* layer, float margin, float thickness,
                       bool fine);
    // Signed triangles from indexed vertices (A41:8433-8590); an index out of
    // range collapses its triangle to the origin.
    bool set_indexed(hipStream_t st, const float* verts, int nverts, const int* idx, int ntri, bool topology_changed, float margin,
                     float thickness);
    // The colliders' velocities hold for the step after an upload only (A41:9389-9390).
    bool clear_motion(hipStream_t st);
    bool is_signed() const { return signed_; }
    bool set_particles(hipStream_t st, const float* pos, const float* prev, const float* radii, const int* src, const int* layer,
                       int n, float margin, float thickness);
    // Device gather of every source's current positions (particles) and
    // triangles, ordered by source; asynchronous on `st`.
    bool gather(hipStream_t st, const std::vector<GatherSource>& sources, float margin, float thickness);
    // Bounds and both hashes; build-time counters go to `counts` when
    // count_diag, and the largest triangle bucket is measured when occupancy.
    bool build(hipStream_t st, unsigned long long* counts, bool count_diag, bool occupancy);

    const DynView* device_view() const { return d_view_; }
    int triangles() const { return nt_; }
    int particles() const { return np_; }
    bool triangles_allocated() const { return tri_cap_ > 0; }
    // Synchronous reads for diagnostics (the caller has synchronised).
    // (triangle, cell) entries of the triangle hash after the last build
    // (waits for the stream's work first); -1 on a copy error
    long long hash_entries() const;
    int large_count() const;
    int max_occupancy() const;

private:
    bool reserve_triangles(int n);
    bool reserve_particles(int n);
    bool upload_view(hipStream_t st);
    // the gather kernels over the uploaded source and segment tables
    bool launch_gather(hipStream_t st, int npseg, int ntseg, bool keep_prev, float max_radius, float reach_radius, float margin,
                       float thickness);
    bool compute_bounds(hipStream_t st);
    // allocation groups: 0 triangles, 1 particles, 2 scan scratch, 3 the view, 4 gather tables
    void* alloc(int group, size_t bytes);
    void release(int group);
    void release_all();

    int nt_ = 0, np_ = 0, tri_cap_ = 0, part_cap_ = 0, tri_buckets_ = 0, part_buckets_ = 0, scan_cap_ = 0;
    float tri_cell_ = 0.0f, span_margin_ = 0.0f, part_cell_ = 0.0f;
    std::vector<float> radii_seen_;
    std::vector<std::pair<void*, size_t>> blocks_[5];
    // triangles
    float4* tv_ = nullptr;
    int* tsrc_ = nullptr;
    int* tlayer_ = nullptr;
    int* tri_order_ = nullptr;   // the hash build's triangle order (gathered scenes; others in index order)
    bool tri_ordered_ = false;
    bool tri_filter_ = false;
    TriBox* box_ = nullptr;
    int *large_flag_ = nullptr, *large_off_ = nullptr, *large_ = nullptr;
    int *tri_count_ = nullptr, *tri_off_ = nullptr, *tri_keys_ = nullptr;
    int2* tri_raw_ = nullptr;   // {key, bucket} in arrival order
    int2* tri_ksl_ = nullptr;
    int4* tri_ent_ = nullptr;   // {key, source, layer, 0} in arrival order (sets not ranked)
    int4* tri_bin_ = nullptr;
    int* occupancy_ = nullptr;
    int* extent_bits_ = nullptr;
    // particles
    float4 *pp_ = nullptr, *pprev_ = nullptr;
    int *psrc_ = nullptr, *player_ = nullptr;
    bool part_filter_ = false;
    int4* pcell_ = nullptr;
    int *pslot_ = nullptr, *pcount_ = nullptr, *poff_ = nullptr, *praw_ = nullptr;
    int4* pbin_ = nullptr;
    PartEntry* pent_ = nullptr;
    float4* pctr_ = nullptr;
    bool signed_ = false, tri_motion_ = false, part_motion_ = false, bounds_valid_ = false;
    int2* coarse_ = nullptr;
    int* coarse_off_ = nullptr;
    int coarse_buckets_ = 0;
    float max_radius_ = 0.0f;
    int* scan_tmp_ = nullptr;
    // layer ranges of the triangle hash (DynView::tri_classes)
    int tri_classes_ = 1;
    int class_base_[kMaxLayerClasses + 1] = {};
    int class_layer_[kMaxLayerClasses] = {};
    std::vector<GatherSource> gathered_;   // the sources the device tables hold (empty: none valid)
    GatherSource* d_sources_ = nullptr;
    int sources_cap_ = 0;
    int4* d_segments_ = nullptr;
    int segments_cap_ = 0;
    DynView* d_view_ = nullptr;
    DynView view_{};
    DynView uploaded_{};            // the last view copied to d_view_
    bool uploaded_valid_ = false;
};

// Triangle hash cell of a soup whose largest triangle extent is max_extent.
float dynamic_triangle_cell_size(float max_extent, float margin, float thickness, bool fine);

}  // namespace ssbl
