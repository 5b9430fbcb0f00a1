// xpbd_device.hpp -- device-side data views and kernel launch interface of the
// gfx950 XPBD cloth library.
//
// HBM layout of one cloth (all arrays 256-byte aligned, vertex arrays padded to
// a multiple of the 256-vertex tile):
//   pos[2]  float4 x V  ping-pong positions; .w = effective inverse mass
//                       (0 for pinned/kinematic vertices)
//   prev    float4 x V  previous positions
//   vel     float4 x V  velocities
//   rest    float4 x V  rest positions; .w = inverse mass as uploaded
//   adj     int2 x 2E   vertex -> incident stretch edges, ascending edge index:
//                       {other vertex | first-endpoint bit 31, rest length bits}
//   ell     int2 x W*Vp the first W <= 8 of those per vertex, slot-major
// Per-vertex passes (integrate, the fused constraint/contact pass, velocity)
// read one float4 per array per vertex -- 16-byte coalesced loads -- and gather
// neighbour positions through L2.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "ssbl_xpbd_hip.h"

namespace ssbl {

constexpr int kTile = 256;              // vertices per workgroup tile
constexpr int kCountSlots = 38;         // diagnostic counters per cloth
constexpr int kNeighborSlots = 32;      // vertex-vertex candidates per vertex
constexpr int kBucketSlots = 32;        // entries a query sees per hash bucket
constexpr int kMaxForceFields = 64;
constexpr int kEllMax = 8;              // incidence slots held slot-major per vertex
constexpr int kStageHaloMax = 384;      // widest neighbour distance a staged stretch gather covers
constexpr double kFixedScale = 4294967296.0;  // 2^32: self-collision accumulators

enum CountSlot {
    kCountSoft = 0, kCountExactImpulse = 1, kCountEdgeEdge = 2, kCountHardFallbacks = 3,
    kCountSelfCandidates = 9, kCountSelfOverflow = 10, kCountStaticSdf = 11, kCountLraTacks = 12,
    kCountBendingWings = 13, kCountTackGuards = 14, kCountBendingGuards = 15,
    // fast self-collision mode (A41:56-72)
    kCountIslandCandidates = 16, kCountIslandClusters = 17, kCountIslandVertexRefs = 18, kCountIslandApplied = 19,
    kCountIslandGuarded = 20, kCountIslandMaxDeltaMicro = 21, kCountFilterSeen = 22, kCountFilterAcceptedVv = 23,
    kCountFilterAcceptedVt = 24, kCountFilterAcceptedEe = 25, kCountFilterRejectedRest = 26, kCountFilterRejectedDuplicate = 27,
    kCountFilterRejectedOwnership = 28, kCountFilterCacheHits = 29, kCountFilterCacheMisses = 30, kCountClusterCount = 31,
    kCountClusterOwnedContacts = 32,
    kCountFriction = 33
};

// Work of the fused per-vertex pass, as a bit mask.
enum PassOps : unsigned {
    kOpStretch = 1u << 0,     // Jacobi stretch gather (src -> dst)
    kOpPin = 1u << 1,         // pin projection
    kOpAnalytic = 1u << 2,    // ground / sphere / wall contacts
    kOpSdf = 1u << 3,         // static SDF contact
    kOpPinPolish = 1u << 4,   // second pin projection after contacts
    kOpVelocity = 1u << 5,    // velocity update (end of substep)
    kOpLra = 1u << 6,         // LRA tether gather (src -> dst), instead of stretch
    kOpCount = 1u << 7,       // count diagnostics
    kOpLastIter = 1u << 8,    // last iteration of the substep (pin polish always runs)
    kOpBend = 1u << 9,        // uncoloured bend gather (src -> dst), instead of stretch
    kOpIntegrate = 1u << 10,  // the next substep's gravity-only integration (parity -> parity ^ 1)
    kOpIntegratePin = 1u << 11,   // ... followed by its pin projection
    // bits 16-22 of a dynamic contact launch's ops: the most probing
    // vertices a wave pools into one item list (more search per lane)
    kOpPoolShift = 16,
};

struct DynView;

// Scalars of one cloth's hard-stretch PCG solve (device resident).
struct PcgScalars {
    double rz, dad, rznext, initial;
    float beta;
    int status, dir_ok, runs;
};

// One cloth's hard-stretch PCG arrays (null / zero when the cloth has none).
struct PcgView {
    int nv, ne, hard, ntiles;
    float strength_cfg;   // configured strength (polish / cap thresholds)
    float strength;       // clamp01(strength * 2.5), the system stiffness
    float thick, wb_limit;
    float4 *rhs, *sol, *res, *z, *dir, *adir;
    float *diag, *pinv, *off;   // 6 floats per block
    // 1: off holds the block of incidence j of vertex v at 6 * (j * ell_pitch
    // + v), slot-major like ClothView::ell (every vertex's incidences fit its
    // slots); 0: at 6 * k for CSR incidence k
    int ell_off;
    // previous positions as they were before an uncoloured cap gather (the
    // gather reads its neighbours' from here while it rewrites its own);
    // null unless the cloth polishes without an edge colouring
    float4* qsnap;
    unsigned long long* partials;   // two 64-bit words per tile of the cloth
    PcgScalars* sc;
    unsigned long long* guard;
    float* max_delta;
    float* tile_max;   // per tile of the cloth: the largest step of the last update
    const int2* edges;
    const float* edge_rest;
    unsigned long long* counts;
};

// Everything a kernel needs to know about one cloth.  Lives in device memory
// (a table of these for a batch); host keeps a mirror.
struct ClothView {
    int nv, ne, nt, nbend, nlra, npin, nfield;
    int direct_stretch;        // 0 when the cloth uses the hard-stretch PCG path
    float4* pos0;
    float4* pos1;
    float4* prev;
    float4* vel;
    float4* rest;
    const int* adj_off;        // V+1
    const int2* adj;           // stretch incidences
    const int* lra_off;        // V+1
    const int4* lra_adj;       // {other | first bit, rest bits, lra index, 0}
    const int* tri_off;        // V+1 incident triangles (surface normal)
    const int* tri_inc;
    const int4* tris;          // {a, b, c, 0}
    const int2* bends;         // bend pairs
    const float* bend_rest;
    const int* bend_off;       // V+1, uncoloured bends only
    const int4* bend_adj;      // {other | first bit, rest bits, bend index, 0}
    const int* pin_off;        // V+1 -> pins of that vertex (ascending pin index)
    const int* pin_ids;        // pin index of each vertex-ordered record
    const float4* pins;        // vertex-ordered {target xyz, relaxation}; relaxation 2 = hard, <=0 ignored
    const float* area_share;
    const float* rest_scale;
    const SsblXpbdForceField* fields;
    const float* sdf;
    const float* static_tris;     // 9 floats per static triangle
    int nstatic;
    int pad2;
    unsigned long long* counts;   // kCountSlots
    float* max_smoothed;          // self-collision max averaged delta
    // parameters
    float dt_sub;
    float damping;
    float gravity[3];
    float stretch_alpha;          // compliance / max(dt^2, eps), per substep
    float bend_alpha;
    float lra_compliance;
    float collision_margin;
    float thickness;
    float friction;
    float tangent_damping;
    int use_pressure;
    float pressure_strength;
    float pressure_accel_scale;
    float pressure_impulse_rest_scale;
    float pressure_velocity_rest_scale;
    float pressure_velocity_length_scale;
    SsblXpbdRuntimeColliders colliders;
    int sdf_ready;
    int sdf_dim[3];
    float sdf_origin[3];
    float sdf_voxel;
    // 1 / voxel and 1 / max(voxel, eps), divided once on the host (IEEE
    // single precision both sides, so bit-equal to the per-sample division),
    // and the largest grid coordinate of each axis, dim - 1, as a float.
    float sdf_inv;
    float sdf_inv_clamped;
    float sdf_top[3];
    int need_normal;
    int contacts_on;              // analytic colliders or a static SDF present
    // The first ell_width incidences of every vertex again, slot-major
    // (slot j of vertex v at ell[j * ell_pitch + v], {-1, 0} when unused), so
    // a wave loads each slot with one coalesced access and without first
    // reading adj_off.  ell_overflow: some vertex has more incidences, the
    // rest are read from adj.
    const int2* ell;
    int ell_width;
    int ell_pitch;
    int ell_overflow;
    // Stretch gathers of this cloth read their neighbours from LDS: a
    // workgroup first copies the source positions of its tile and of the
    // stage_halo indices either side (every ELL neighbour of the cloth lies
    // within that distance, kStageHaloMax at most).  0: gathered from HBM.
    int stage_halo;
    // The same slots in 6 bytes for the stretch gather, when every
    // neighbour of the cloth lies within 16383 indices (else null, and ell
    // is used): ell16 = (other - v) * 2 + first-endpoint bit (0: unused),
    // ell_rest = the rest length.
    const short* ell16;
    const float* ell_rest;
    // Dynamic colliders (null when none): the cloth's own, or an attached
    // global scene filtered by source and layer (A41:3984-3997).
    const DynView* dyn;
    int dyn_mode;                 // 0 no filter, 1 every other source, 2 lower layers only
    int dyn_target_src;
    int dyn_target_layer;
    // The cloth's own signed (indexed) dynamic triangles and its particles
    // (null when none), contacts of A41:4520-4540; collider_dt turns their
    // centre motion into a velocity (A41:4513-4518).
    const DynView* sdyn;
    float collider_dt;
    int pad5;
    // Per vertex, the deepest triangle contact key of the pooled searches of
    // the substep's dynamic contacts (0 between launches: the contact pass
    // clears what it reads).
    unsigned long long* dyn_best;
    // Stretch edges {i, j, rest length bits, 0} in input (colour-sorted)
    // order, for the coloured Gauss-Seidel stretch mode (null otherwise).
    const int4* sedges;
    // The hard-stretch solve's arrays: the cloth's own, or the static
    // scratch of the allocation-free step.
    PcgView pcg;
};

struct Tile { int cloth, begin, end, pad; };

// Kernel launch wrappers (xpbd_kernels.hip).  All launches go to `stream`.
struct LaunchCtx {
    hipStream_t stream;
    const ClothView* views;     // device table
    const Tile* tiles;          // device tile table
    int ntiles;
    // The dynamic contacts' launch plan (required when a launch has dynamic
    // contacts): their launches add each tile's work to tile_work
    // (dynamic_work_ints), and launch_order_tiles turns the work of the step
    // before into tile_order (dynamic_order_ints): the contact pass's tile
    // order and the triangle searches' units (k_order_tiles).
    int* tile_work;
    int* tile_order;
};

// Triangle search units: a tile's pooled lists split over up to
// kDynMaxSlices workgroups, one more per kDynSliceItems items of its longest
// wave list in the step before.
constexpr int kDynSliceItems = 512;
constexpr int kDynMaxSlices = 8;   // at most 16 (4 bits of a unit)
constexpr int kUnitSliceShift = 23, kUnitSlicesShift = 27, kUnitTileMask = (1 << kUnitSliceShift) - 1;
inline int dynamic_work_ints(int ntiles) { return 2 * ntiles; }
inline int dynamic_order_ints(int ntiles) { return 4 * ntiles + 1; }

void launch_integrate(const LaunchCtx& c, int parity, unsigned ops, bool full);
void launch_vertex_pass(const LaunchCtx& c, int src_parity, int dst_parity, unsigned ops);
// tail_ops: the closing pin polish and velocity update, fused into the
// particle launch when nothing else runs between (0: none)
// The dynamic contacts: the triangle search units, then the contact pass
// (triangles from the searches' keys, then particles), which pools the
// particle searches of a wave with at most `pool_max` probing vertices
// (0..64); the events, when given, bracket both launches.  The context's
// launch plan must have been built by launch_order_tiles in the same step.
constexpr int kDefaultPoolMax = 16;
void launch_dynamic_pass(const LaunchCtx& c, int parity, unsigned ops, unsigned tail_ops, int pool_max = kDefaultPoolMax,
                         hipEvent_t begin = nullptr, hipEvent_t end = nullptr);
// Signed contacts: the pass after self collision (inside = false) and the
// inside-only passes of a step's last iteration (A41:9322-9359).
void launch_order_tiles(const LaunchCtx& c);
void launch_signed_pass(const LaunchCtx& c, int parity, bool inside, unsigned ops);
void launch_bend_color(const LaunchCtx& c, const int4* segments, int nseg_blocks, int parity, unsigned ops);
// One colour of the coloured stretch mode; segments {cloth, begin, end}.
void launch_stretch_color(const LaunchCtx& c, const int4* segments, int nseg_blocks, int parity, unsigned ops);
void launch_sdf_build(hipStream_t s, float* sdf, int dx, int dy, int dz, float ox, float oy, float oz, float voxel,
                      const float* tris, int ntris, int signed_sdf);
void launch_fill_u32(hipStream_t s, unsigned* p, unsigned value, size_t n);
// Exclusive prefix sum of n ints into out[0..n] (out[n] = total), on the
// stream; tmp holds scan_tmp_ints(n) ints.
void exclusive_scan(hipStream_t s, const int* in, int n, int* out, int* tmp);
int scan_tmp_ints(int n);

// Device allocations of the library go through these, so the bytes a solver
// holds can be reported (ssbl_ext_device_bytes).
void* device_alloc(size_t bytes, const char* what);
void device_free(void* p, size_t bytes);
long long device_bytes_in_use();

}  // namespace ssbl
