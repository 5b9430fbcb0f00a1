// xpbd_self.hpp -- self-collision state of one cloth and its per-step driver.
//
// Reference: A41:4542-5844 (hash builds, candidate table, vertex-vertex,
// vertex-triangle and edge-edge responses, averaging), host set-up
// A41:6784-6994, schedule A41:9255-9345; fast-mode filters A41:708-849,
// overlap islands A41:5226-5439.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "xpbd_device.hpp"

namespace ssbl {

bool fail(const std::string& message);
bool hip_ok(hipError_t err, const char* what);

struct Solver;

// One spatial hash as a counting-sorted bucket table.  Items are counted per
// bucket (each item keeps the arrival slot its atomic returned), scattered to
// raw[off[b] + slot], and then every item computes its rank among the keys of
// its bucket: entries[off[b] + rank] holds the min(count, 32) smallest keys of
// bucket b in ascending order (key = vertex, or primitive * 64 + cell index in
// the primitive's box loop), which is the order queries visit.  The ranks do
// not depend on the order the atomics landed in.
//
// A cloth's three hashes (vertex, triangle, edge) share one table: their
// buckets are consecutive ranges of one count / offset array, scanned
// together, so an offset is an index into one raw / rawb / entries array, and
// each build phase is one launch for all three.
struct HashView {
    int buckets;
    int capacity;
    int base;        // first bucket of this hash in the shared table
    int* count;      // buckets (a range of the shared counts)
    int* off;        // buckets + 1 (a range of the shared offsets)
    int* slot;       // capacity: arrival slot of item (vertex, or prim * 64 + cell)
    int* raw;        // shared: keys in arrival order
    int* rawb;       // shared: shared-table bucket of raw[j]
    int* entries;    // shared: first min(count, 32) keys of a bucket, ascending
};

struct SelfView {
    int nv, nt, ne;
    int use_vhash, use_thash, use_ehash;
    float cell, radius, onset, rest_skip, prim_margin;
    int max_nb;                  // vertex-triangle / edge-edge candidates per query (16 at most in fast mode)
    int fast;                    // fast mode: contact filters, table reuse, overlap islands
    int island_contacts;         // overlap-island contact cap, clamp(2 * max neighbours, 3, 48)
    int4* vcell;                 // V: cell x, y, z, bucket (vertex hash)
    HashView vh, th, eh;
    int* all_count;              // the shared table's counts (all_buckets)
    int* all_off;                // and offsets (all_buckets + 1)
    int all_buckets;
    int hash_items;              // largest number of items the three hashes hold together
    int vh_blocks, th_blocks;    // blocks of the vertex and triangle parts of a build launch
    unsigned* nb_count;          // V
    unsigned* nb_idx;            // V * 32
    unsigned long long* acc;     // 3V fixed-point 2^-32 m
    int* accw;                   // V
    const int4* tris;
    const int2* edges;           // mesh order, endpoints clamped >= 0
    // The order in which a hash build visits the vertices, triangles and
    // edges (permutations, by Morton code of their rest positions, so that a
    // block's items share buckets); the items and their keys are unchanged.
    const int* vorder;
    const int* torder;
    const int* eorder;
    int* scan_tmp;               // block sums for the bucket scans
    unsigned long long* counts;  // the cloth's diagnostic counters
};

// Runs the self-collision passes of one step for a set of cloths.
class SelfBatch {
public:
    virtual ~SelfBatch() = default;
    virtual bool run(const LaunchCtx& ctx, int parity, int substep, int substeps, bool count) = 0;
};

class SelfCollision : public SelfBatch {
public:
    SelfCollision() = default;
    SelfCollision(const SelfCollision&) = delete;
    SelfCollision& operator=(const SelfCollision&) = delete;
    ~SelfCollision() override;
    bool prepare(const SsblXpbdConfig& cfg, const std::vector<int>& edges, const std::vector<int>& tris, int nv,
                 unsigned long long* counts, const std::vector<int>& vorder, const std::vector<int>& torder,
                 const std::vector<int>& eorder);
    // The driver for a lone cloth: `view` is the cloth's host view mirror and
    // `dview` its device copy.
    SelfBatch* single_batch(const ClothView& view, const ClothView* dview);
    bool run(const LaunchCtx& ctx, int parity, int substep, int substeps, bool count) override;
    bool run_for(hipStream_t stream, const ClothView* dview, int parity, bool count);
    bool download_cells(int* cells, int nv);
    bool download_neighbors(unsigned* counts, unsigned* indices, int nv);
    bool enabled() const { return ready_; }
    int interval() const { return interval_; }

private:
    bool ready_ = false;
    int passes_ = 1;
    int interval_ = 1;
    SelfView sv_{};
    std::vector<std::pair<void*, size_t>> allocations_;
    const ClothView* dview_ = nullptr;
    void* alloc_bytes(size_t bytes);
};

// Self collision of every member of a batch (members run one after another).
// The self-collision state of a solver (defined with the solver).
SelfCollision* solver_self(Solver* s);

class SelfBatchHolder {
public:
    SelfBatch* get(const std::vector<Solver*>& members, const std::vector<ClothView>& views, const ClothView* dviews);

private:
    struct Entry { SelfCollision* self; const ClothView* dview; int substep_interval; };
    class Runner : public SelfBatch {
    public:
        std::vector<Entry> entries;
        bool run(const LaunchCtx& ctx, int parity, int substep, int substeps, bool count) override;
    };
    Runner runner_;
};

}  // namespace ssbl
