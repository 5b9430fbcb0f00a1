// xpbd_solver.hip -- host side of libssbl_xpbd_hip_abi41.so: the ABI41 entry
// points (include/ssbl_xpbd_hip.h), the ssbl_ext_* extensions
// (include/ssbl_xpbd_hip_ext.h), device memory set-up and the substep
// schedule.
//
// Reference: native/src/ssbl_xpbd_cuda_abi41.cu (A41:<line>); create
// A41:8791-8967, step A41:9087-9398, uploads A41:5886-7109, 8282-8317.
//
// Ordering model: the library owns one HIP stream; every entry point that
// changes device data first waits for that stream (the reference's cudaMemcpy
// on the legacy default stream waits for all earlier work the same way), so the
// caller may free its arrays on return.  ssbl_step_solver_ex only enqueues
// work unless force_sync / fetch_diagnostics is set.
#include "ssbl_xpbd_hip.h"
#include "ssbl_xpbd_hip_ext.h"
#include "xpbd_device.hpp"
#include "xpbd_dynamic.hpp"
#include "xpbd_pcg.hpp"
#include "xpbd_self.hpp"
#include "xpbd_timing.hpp"

#include <algorithm>
#include <atomic>
#include <memory>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

namespace ssbl {

thread_local std::string g_error;

bool fail(const std::string& message) {
    g_error = message;
    return false;
}

bool hip_ok(hipError_t err, const char* what) {
    if (err == hipSuccess) return true;
    g_error = std::string(what) + ": " + hipGetErrorString(err);
    return false;
}

hipStream_t lib_stream() {
    static hipStream_t stream = [] {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
        return s;
    }();
    return stream;
}

// The stream a batch's second part runs on (forked from and joined back to
// the library stream within each step; default priority like the first).
static hipStream_t part_stream() {
    static hipStream_t stream = [] {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
        return s;
    }();
    return stream;
}

// ---------------------------------------------------------------- timing
namespace {
std::atomic<bool> g_timing{false};
std::mutex g_timing_mutex;
struct Pending { const char* name; hipEvent_t begin, end; };
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_event_pool;
struct KStat { long long launches = 0; double total_ms = 0.0; };
std::map<std::string, KStat> g_kstats;

hipEvent_t take_event() {
    if (!g_event_pool.empty()) { hipEvent_t e = g_event_pool.back(); g_event_pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

void resolve_pending() {
    for (const Pending& p : g_pending) {
        float ms = 0.0f;
        hipEventSynchronize(p.end);
        hipEventElapsedTime(&ms, p.begin, p.end);
        KStat& k = g_kstats[p.name];
        k.launches += 1;
        k.total_ms += ms;
        g_event_pool.push_back(p.begin);
        g_event_pool.push_back(p.end);
    }
    g_pending.clear();
}
}  // namespace

bool timing_enabled() { return g_timing.load(std::memory_order_relaxed); }
const char*& timing_scope_slot() {
    thread_local const char* name = nullptr;
    return name;
}
void timing_take(hipEvent_t* begin, hipEvent_t* end) {
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    *begin = take_event();
    *end = take_event();
}
void timing_push(const char* name, hipEvent_t begin, hipEvent_t end) {
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    g_pending.push_back({name, begin, end});
    if (g_pending.size() > 4096) resolve_pending();
}

// ------------------------------------------------------- device buffers
// Bytes allocated minus bytes freed by the calling thread: a handle's
// creation and updates are measured on the thread that makes them, so
// handles created on one thread while another steps are accounted exactly.
namespace {
thread_local long long t_device_bytes = 0;
}  // namespace

void* device_alloc(size_t bytes, const char* what) {
    void* p = nullptr;
    if (!hip_ok(hipMalloc(&p, bytes), what)) return nullptr;
    t_device_bytes += static_cast<long long>(bytes);
    return p;
}
void device_free(void* p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    t_device_bytes -= static_cast<long long>(bytes);
}
long long device_bytes_in_use() { return t_device_bytes; }

template <class T> struct Dev {
    T* p = nullptr;
    size_t n = 0;
    Dev() = default;
    Dev(const Dev&) = delete;
    Dev& operator=(const Dev&) = delete;
    ~Dev() { release(); }
    void release() { device_free(p, sizeof(T) * n); p = nullptr; n = 0; }
    bool alloc(size_t count, const char* what) {
        release();
        if (count == 0) return true;
        p = static_cast<T*>(device_alloc(sizeof(T) * count, what));
        if (!p) return false;
        n = count;
        return true;
    }
    bool upload(const std::vector<T>& h, const char* what) {
        if (h.size() != n && !alloc(h.size(), what)) return false;
        if (h.empty()) return true;
        return hip_ok(hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice), what);
    }
    // Grow-only forms: reallocate only when more elements are needed.
    bool reserve(size_t count, const char* what) { return count <= n || alloc(count, what); }
    bool upload_grow(const std::vector<T>& h, const char* what) {
        if (!reserve(h.size(), what)) return false;
        if (h.empty()) return true;
        return hip_ok(hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice), what);
    }
};

// Pinned staging ring for per-step parameter uploads that are ordered on the
// stream (the previous step may still be reading the device copy).
struct Staging {
    static constexpr int kRing = 4;
    void* host[kRing] = {};
    hipEvent_t done[kRing] = {};
    size_t bytes = 0;
    int next = 0;
    ~Staging() { release(); }
    void release() {
        for (int i = 0; i < kRing; ++i) {
            if (host[i]) hipHostFree(host[i]);
            if (done[i]) hipEventDestroy(done[i]);
            host[i] = nullptr;
            done[i] = nullptr;
        }
        bytes = 0;
    }
    // Sizes the ring for uploads of up to n bytes (done at set-up, so that a
    // step allocates nothing).
    bool reserve(size_t n) {
        if (n <= bytes) return true;
        release();
        for (int i = 0; i < kRing; ++i) {
            if (!hip_ok(hipHostMalloc(&host[i], n), "staging allocation")) return false;
            if (!hip_ok(hipEventCreateWithFlags(&done[i], hipEventDisableTiming), "staging event")) return false;
        }
        bytes = n;
        return true;
    }
    bool upload(void* dst, const void* src, size_t n, hipStream_t s) {
        if (!reserve(n)) return false;
        const int slot = next;
        next = (next + 1) % kRing;
        if (!hip_ok(hipEventSynchronize(done[slot]), "staging wait")) return false;
        std::memcpy(host[slot], src, n);
        if (!hip_ok(hipMemcpyAsync(dst, host[slot], n, hipMemcpyHostToDevice, s), "parameter upload")) return false;
        return hip_ok(hipEventRecord(done[slot], s), "staging record");
    }
};

int next_pow2(int v) {
    int r = 1;
    v = std::max(v, 1);
    while (r < v && r < (1 << 30)) r <<= 1;
    return r;
}

// ----------------------------------------------------------------- solver
// ---------------------------------------------------------------- graphs
std::atomic<int> g_graph_mode{1};

// The launches of a step's substep loop, captured once into a HIP graph and
// replayed whenever the launch sequence is one already seen: the key holds
// everything the sequence depends on (substeps, iterations, which passes run,
// the tuned launch limit), and the launches themselves take only buffers the
// owner keeps for its lifetime, so a replay issues exactly the launches a
// fresh capture would.  A capture the
// runtime refuses turns the holder off; it then launches directly.
struct StepGraph {
    static constexpr size_t kSlots = 4;   // launch sequences kept (least recently used goes)
    struct Slot { std::vector<long long> key; hipGraphExec_t exec = nullptr; long long used = 0; };
    std::vector<Slot> slots;
    bool refused = false;
    long long captures = 0, replays = 0, clock = 0;
    StepGraph() = default;
    StepGraph(const StepGraph&) = delete;
    StepGraph& operator=(const StepGraph&) = delete;
    // the owner waits for the stream before it goes
    ~StepGraph() {
        for (Slot& sl : slots)
            if (sl.exec) hipGraphExecDestroy(sl.exec);
    }
    // Drops every kept graph (the owner's launch tables changed); waits for
    // the stream first, since a dropped graph may still run.
    bool clear(hipStream_t st) {
        if (slots.empty()) return true;
        if (!hip_ok(hipStreamSynchronize(st), "step graph reset")) return false;
        for (Slot& sl : slots)
            if (sl.exec) hipGraphExecDestroy(sl.exec);
        slots.clear();
        return true;
    }
    template <class F>
    bool run(hipStream_t st, const std::vector<long long>& k, F&& record) {
        if (refused) return record();
        ++clock;
        for (Slot& sl : slots) {
            if (sl.key != k) continue;
            sl.used = clock;
            ++replays;
            return hip_ok(hipGraphLaunch(sl.exec, st), "step graph launch");
        }
        if (slots.size() == kSlots) {
            // an evicted graph may still run: wait for the stream first
            size_t old = 0;
            for (size_t i = 1; i < slots.size(); ++i)
                if (slots[i].used < slots[old].used) old = i;
            if (!hip_ok(hipStreamSynchronize(st), "step graph eviction")) return false;
            hipGraphExecDestroy(slots[old].exec);
            slots.erase(slots.begin() + static_cast<long>(old));
        }
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            refused = true;
            return record();
        }
        const std::string error_before = g_error;
        const bool ok = record();
        hipGraph_t g = nullptr;
        const hipError_t end = hipStreamEndCapture(st, &g);
        hipGraphExec_t e = nullptr;
        const hipError_t inst = (ok && end == hipSuccess) ? hipGraphInstantiate(&e, g, nullptr, nullptr, 0) : end;
        if (g) hipGraphDestroy(g);
        if (!ok || inst != hipSuccess) {
            // nothing captured ran: launch the same sequence directly
            if (e) hipGraphExecDestroy(e);
            (void)hipGetLastError();
            g_error = error_before;
            refused = true;
            return record();
        }
        slots.push_back(Slot{k, e, clock});
        ++captures;
        return hip_ok(hipGraphLaunch(e, st), "step graph launch");
    }
};

// ---------------------------------------------------------------- tuning
std::atomic<int> g_autotune{0};

// Picks the dynamic contact pass's particle pooling limit (kOpPoolShift) for one
// handle or batch by timing the candidates on its own launches.  Every
// kEvery-th step with dynamic contacts (the first included) is a timed step:
// substep k runs candidate (k + round) mod kCount, its contact launch
// bracketed by events; a later step reads the events once they have
// completed (no wait) and keeps the candidate with the least mean time, unless
// the one in use is within 3 % of it.  Contacts change over a run (cloths meet
// and pile up), and with them the best limit, hence the repeats.  Every
// candidate gives the same contacts, so tuning changes timing only.  Timed
// steps are never captured into a graph.
struct PoolTuner {
    static constexpr int kCount = 5;
    static constexpr int kCandidates[kCount] = {0, 8, 16, 32, 64};
    static constexpr int kEvery = 16;
    int chosen = -1;
    long long step = 0;    // steps with dynamic contacts so far
    int round = 0;         // timed steps so far
    bool pending = false;  // a timed step whose events are not read yet
    double total_ms[kCount] = {};
    int samples[kCount] = {};
    std::vector<hipEvent_t> events;   // begin/end per substep of the timed step
    std::vector<int> which;           // candidate index per substep, -1 when not launched
    PoolTuner() = default;
    PoolTuner(const PoolTuner&) = delete;
    PoolTuner& operator=(const PoolTuner&) = delete;
    ~PoolTuner() {
        for (hipEvent_t e : events)
            if (e) hipEventDestroy(e);
    }
    bool tuning() const { return g_autotune.load(std::memory_order_relaxed) != 0 && step % kEvery == 0; }
    int value() const { return chosen >= 0 ? chosen : kDefaultPoolMax; }
    bool begin_step(int substeps) {
        for (int k = 0; k < kCount; ++k) {
            total_ms[k] = 0.0;
            samples[k] = 0;
        }
        while (events.size() < 2 * static_cast<size_t>(substeps)) {
            hipEvent_t e = nullptr;
            if (!hip_ok(hipEventCreate(&e), "tuning event")) return false;
            events.push_back(e);
        }
        which.assign(static_cast<size_t>(substeps), -1);
        return true;
    }
    // The limit substep `sub` of the timed step runs, and its events.
    int candidate(int sub, hipEvent_t* begin, hipEvent_t* end) {
        const int k = (sub + round) % kCount;
        which[static_cast<size_t>(sub)] = k;
        *begin = events[2 * static_cast<size_t>(sub)];
        *end = events[2 * static_cast<size_t>(sub) + 1];
        return kCandidates[k];
    }
    void end_step() { pending = true; }
    // Reads the last timed step's events when they have completed, or waits
    // for them when `wait`, and updates the choice.
    bool poll(bool wait) {
        if (!pending) return true;
        int last = -1;
        for (size_t sub = 0; sub < which.size(); ++sub)
            if (which[sub] >= 0) last = static_cast<int>(sub);
        if (last >= 0 && !wait) {
            const hipError_t q = hipEventQuery(events[2 * static_cast<size_t>(last) + 1]);
            if (q == hipErrorNotReady) {
                (void)hipGetLastError();
                return true;
            }
            if (!hip_ok(q, "tuning event")) return false;
        }
        pending = false;
        for (size_t sub = 0; sub < which.size(); ++sub) {
            const int k = which[sub];
            if (k < 0) continue;
            float ms = 0.0f;
            if (!hip_ok(hipEventSynchronize(events[2 * sub + 1]), "tuning event") ||
                !hip_ok(hipEventElapsedTime(&ms, events[2 * sub], events[2 * sub + 1]), "tuning event"))
                return false;
            total_ms[k] += ms;
            samples[k] += 1;
        }
        ++round;
        int best = -1, cur = -1;
        for (int k = 0; k < kCount; ++k) {
            if (samples[k] == 0) continue;
            if (kCandidates[k] == chosen) cur = k;
            if (best < 0 || total_ms[k] / samples[k] < total_ms[best] / samples[best]) best = k;
        }
        if (best >= 0 && (cur < 0 || total_ms[best] / samples[best] < 0.97 * (total_ms[cur] / samples[cur]))) chosen = kCandidates[best];
        return true;
    }
};
constexpr int PoolTuner::kCandidates[PoolTuner::kCount];

struct HostPin { int index; float target[3]; float weight; };

// The triangles of a mesh in Morton order of their centroids (10 bits per
// axis over the mesh's bounds), ties by index.  A scene's hash build takes a
// gathered cloth's triangles in this order, so the triangles of one
// workgroup lie side by side and share most of their hash cells: fewer
// distinct buckets per workgroup, so fewer device atomics, and each bucket's
// entries from one workgroup are written together.  The order changes no
// bucket's contents.
static std::vector<int> spatial_order(const std::vector<float4>& pos, const std::vector<int>& tri, int count) {
    std::vector<int> order(std::max(count, 0));
    for (int t = 0; t < count; ++t) order[t] = t;
    const int n = static_cast<int>(pos.size());
    if (count < 2 || n == 0) return order;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    std::vector<float> ctr(3 * static_cast<size_t>(count), 0.0f);
    for (int t = 0; t < count; ++t) {
        const int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
        if (a >= n || b >= n || c >= n) continue;
        const float x[3] = {(pos[a].x + pos[b].x + pos[c].x) / 3.0f, (pos[a].y + pos[b].y + pos[c].y) / 3.0f,
                            (pos[a].z + pos[b].z + pos[c].z) / 3.0f};
        for (int k = 0; k < 3; ++k) {
            if (!std::isfinite(x[k])) continue;
            ctr[3 * t + k] = x[k];
            lo[k] = std::min(lo[k], x[k]);
            hi[k] = std::max(hi[k], x[k]);
        }
    }
    auto spread = [](unsigned v) {   // 10 bits to every third bit
        v &= 0x3ffu;
        v = (v | (v << 16)) & 0x030000ffu;
        v = (v | (v << 8)) & 0x0300f00fu;
        v = (v | (v << 4)) & 0x030c30c3u;
        v = (v | (v << 2)) & 0x09249249u;
        return v;
    };
    std::vector<std::pair<unsigned, int>> keys(count);
    for (int t = 0; t < count; ++t) {
        unsigned q[3];
        for (int k = 0; k < 3; ++k) {
            const float w = hi[k] - lo[k];
            const float u = (std::isfinite(w) && w > 0.0f) ? (ctr[3 * t + k] - lo[k]) / w : 0.0f;
            q[k] = static_cast<unsigned>(std::min(std::max(u, 0.0f), 1.0f) * 1023.0f);
        }
        keys[t] = {spread(q[0]) | (spread(q[1]) << 1) | (spread(q[2]) << 2), t};
    }
    std::sort(keys.begin(), keys.end());
    for (int t = 0; t < count; ++t) order[t] = keys[t].second;
    return order;
}

// A global dynamic scene (A41:309-340): the colliders every attached cloth
// sees, built by an update from host arrays or gathered on the device.
struct GlobalScene {
    DynamicColliders set;
    float margin = 0.0f, thickness = 0.0f;
    Dev<unsigned long long> counts;   // build-time counters
    long long hash_overflow = 0;
    bool overflow_pending = false;    // counters of an asynchronous build not read yet
};

struct Solver {
    long long dev_bytes = 0;   // device memory held (ssbl_ext_device_bytes)
    SsblXpbdConfig cfg{};
    SsblXpbdRuntimeColliders colliders{};
    int stretch_mode = SSBL_EXT_STRETCH_JACOBI;
    int pcg_cadence = 1;   // SSBL_ABI41_PCG_SUBSTEP_CADENCE when set above 0 (A41:284-292)
    int nv = 0;
    // host mirrors needed after create
    std::vector<float4> h_rest;          // w = inverse mass as uploaded
    std::vector<float> h_weff;           // effective inverse mass
    std::vector<int> edge_colors, bend_colors;
    bool bend_coloring_valid = false;
    bool bend_gather = false;
    int lra_invalid = 0;
    std::vector<HostPin> pins;
    int pins_iterations = -1;
    std::vector<SsblXpbdForceField> fields;
    int unsupported_fields = 0;
    float pressure_accel_scale = 1792.0f, pressure_impulse_rest_scale = 0.85f;
    float pressure_velocity_rest_scale = 2.25f, pressure_velocity_length_scale = 1.0f;
    // static SDF
    int sdf_ready = 0, sdf_dim[3] = {0, 0, 0}, sdf_signed = 0, nstatic = 0;
    float sdf_origin[3] = {0, 0, 0}, sdf_voxel = 0.0f;
    long long sdf_rebuilds = 0, sdf_unsigned = 0;
    float sdf_build_ms = 0.0f;
    // device state
    Dev<float4> pos0, pos1, prev, vel, rest;
    Dev<int> adj_off, lra_off, tri_off, tri_inc, pin_off, pin_ids, bend_off;
    Dev<int> tri_order;   // the order a gathered scene's hash build visits this cloth's triangles
    Dev<unsigned long long> dyn_best;   // per vertex: the triangle searches' deepest contact key (zero between launches)
    Dev<int2> adj, bends, ell;
    Dev<short> ell16;
    Dev<float> ell_rest;
    int ell_width = 0, ell_pitch = 0, ell_overflow = 0, stage_halo = 0;
    Dev<int4> lra_adj, tris, bend_segs, bend_adj;
    Dev<float> bend_rest, area_share, rest_scale, sdf, static_tris;
    Dev<float4> d_pins;
    Dev<SsblXpbdForceField> d_fields;
    Dev<unsigned long long> counts;
    Dev<float> max_smoothed;
    Dev<ClothView> d_view;
    Dev<Tile> d_tiles;
    Dev<int> tile_work, tile_order;   // the dynamic contact pass's tile order (LaunchCtx)
    std::vector<int> bend_seg_first, bend_seg_count;   // per colour, into bend_segs
    std::vector<int4> h_bend_segs;
    // coloured stretch mode: the edges as uploaded, whether their colouring
    // is valid (no vertex twice in a colour), and, once the mode is chosen,
    // the device edge records and per-colour segments
    std::vector<int4> h_sedges;
    bool edge_coloring_valid = false;
    Dev<int4> sedges, stretch_segs;
    std::vector<int4> h_stretch_segs;
    std::vector<int> stretch_seg_first, stretch_seg_count;
    int ntiles = 0;
    Staging staging;
    SelfCollision self;
    PcgPath pcg;
    PcgGroup pcg_group;   // the solve's launches over this cloth alone
    DynamicColliders dyn;              // the cloth's own dynamic colliders
    GlobalScene* scene = nullptr;      // attached global scene
    int target_src = 0, target_layer = 0, cross_mode = 0;
    bool use_scene() const { return scene != nullptr && cross_mode != 0; }
    const DynamicColliders& dyn_set() const { return use_scene() ? scene->set : dyn; }
    bool dyn_on() const { return dyn_set().triangles() > 0 || dyn_set().particles() > 0; }
    long long wing_count = 0;
    ClothView view{};
    SsblXpbdDiagnostics diag{};
    bool counts_pending = false;
    StepGraph graph;
    PoolTuner tuner;

    bool create(const SsblXpbdConfig* config, const SsblXpbdMesh* mesh);
    bool build_sdf(const float* triangles, int count);
    bool set_sdf_grid(const int* dims, const float* origin_and_voxel, const float* values, int count);
    bool upload_pins(const int* idx, const float* p, const float* w, int count);
    bool refresh_pins(int iterations);
    bool upload_fields(const SsblXpbdForceField* f, int count, int unsupported);
    bool set_stretch_mode(int mode);
    bool colored_stretch() const {
        return stretch_mode == SSBL_EXT_STRETCH_COLORED && cfg.edge_count > 0 && !cfg.stretch_optimization_enabled;
    }
    void fill_view(int substeps);
    float self_radius() const {
        const float req = std::max(cfg.self_collision_distance, 0.0f);
        if (req > 0.0f) return req;
        return std::max(std::max(cfg.cloth_thickness, cfg.collision_margin), 0.0f);
    }
    bool analytic_on() const { return colliders.use_ground || colliders.use_wall || colliders.use_sphere; }
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
};

SelfCollision* solver_self(Solver* s) { return &s->self; }

bool sync_stream() { return hip_ok(hipStreamSynchronize(lib_stream()), "stream synchronize"); }

// A41:7413-7584 per-vertex pressure area shares and rest scales (host set-up).
static void pressure_setup(Solver& s, const float* rest_f, const std::vector<int>& tri, const std::vector<int>& edges,
                           const std::vector<float>& edge_rest, std::vector<float>& share, std::vector<float>& scale) {
    const int n = s.nv;
    auto P = [&](int i, int k) { return rest_f[3 * i + k]; };
    auto cross_dot = [&](int a, int b, int c) {
        const float ux = P(b, 0) - P(a, 0), uy = P(b, 1) - P(a, 1), uz = P(b, 2) - P(a, 2);
        const float vx = P(c, 0) - P(a, 0), vy = P(c, 1) - P(a, 1), vz = P(c, 2) - P(a, 2);
        const float cx = std::fma(uy, vz, -(uz * vy)), cy = std::fma(uz, vx, -(ux * vz)), cz = std::fma(ux, vy, -(uy * vx));
        return std::fma(cz, cz, std::fma(cy, cy, std::fma(cx, cx, 0.0f)));
    };
    share.assign(n, 0.0f);
    double total_area = 0.0;
    int valid = 0;
    const int nt = static_cast<int>(tri.size() / 3);
    for (int t = 0; t < nt; ++t) {
        const int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
        if (a >= n || b >= n || c >= n) continue;
        const float sq = cross_dot(a, b, c);
        if (!std::isfinite(sq) || sq <= 0.0f) continue;
        const float area = 0.5f * std::sqrt(sq);
        if (!std::isfinite(area) || area <= 0.0f) continue;
        const float third = area / 3.0f;
        share[a] += third; share[b] += third; share[c] += third;
        total_area += static_cast<double>(area);
        ++valid;
    }
    if (valid <= 0 || total_area <= 0.0 || !std::isfinite(total_area)) {
        std::fill(share.begin(), share.end(), 0.0f);
    } else {
        const float mean = static_cast<float>(total_area / static_cast<double>(std::max(n, 1)));
        const float cap = std::max(mean * 16.0f, 1.0e-12f);
        for (float& x : share) x = (!std::isfinite(x) || x <= 0.0f) ? 0.0f : std::min(x, cap);
    }
    scale.assign(n, 0.0f);
    std::vector<double> sums(n, 0.0);
    std::vector<int> cnt(n, 0);
    const int ne = static_cast<int>(edge_rest.size());
    for (int e = 0; e < ne; ++e) {
        const int i = edges[2 * e], j = edges[2 * e + 1];
        const float r = edge_rest[e];
        if (i >= n || j >= n || !std::isfinite(r) || r <= 0.0f) continue;
        sums[i] += r; sums[j] += r; cnt[i]++; cnt[j]++;
    }
    double total = 0.0;
    int vc = 0;
    for (int i = 0; i < n; ++i) {
        float x = 0.0f;
        if (cnt[i] > 0) x = static_cast<float>(sums[i] / static_cast<double>(cnt[i]));
        else if (std::isfinite(share[i]) && share[i] > 0.0f) x = std::sqrt(share[i]);
        if (std::isfinite(x) && x > 0.0f) { scale[i] = x; total += x; ++vc; }
    }
    const float mean_scale = vc > 0 ? static_cast<float>(total / static_cast<double>(vc)) : 1.0f;
    const float lo = std::max(mean_scale * 0.02f, 1.0e-7f), hi = std::max(mean_scale * 16.0f, lo);
    for (float& x : scale) {
        if (!std::isfinite(x) || x <= 0.0f) x = mean_scale;
        x = std::max(lo, std::min(x, hi));
    }
    float obj = 0.08f;
    if (const char* raw = std::getenv("SSBL_ABI41_PRESSURE_VELOCITY_OBJECT_SCALE")) {
        char* end = nullptr;
        const float v = std::strtof(raw, &end);
        if (raw[0] && end != raw && std::isfinite(v)) obj = v;
    }
    obj = std::max(0.005f, std::min(obj, 1.0f));
    float ms = 0.0f;
    int mc = 0;
    for (float x : scale) if (std::isfinite(x) && x > 0.0f) { ms += x; ++mc; }
    ms = mc > 0 ? ms / static_cast<float>(mc) : 1.0f;
    s.pressure_velocity_length_scale = std::max(ms, 1.0e-6f);
    bool any = false;
    float lo3[3] = {0, 0, 0}, hi3[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(P(i, 0)) || !std::isfinite(P(i, 1)) || !std::isfinite(P(i, 2))) continue;
        for (int k = 0; k < 3; ++k) {
            if (!any) { lo3[k] = hi3[k] = P(i, k); }
            else { lo3[k] = std::min(lo3[k], P(i, k)); hi3[k] = std::max(hi3[k], P(i, k)); }
        }
        any = true;
    }
    if (any) {
        const float sx = hi3[0] - lo3[0], sy = hi3[1] - lo3[1], sz = hi3[2] - lo3[2];
        const float d = std::sqrt(std::max(std::fma(sz, sz, std::fma(sy, sy, std::fma(sx, sx, 0.0f))), 0.0f));
        if (std::isfinite(d) && d > 0.0f) s.pressure_velocity_length_scale = std::max(d * obj, ms * 4.0f);
    }
}

static float env_float(const char* name, float def) {
    const char* raw = std::getenv(name);
    if (!raw || !raw[0]) return def;
    char* end = nullptr;
    const float v = std::strtof(raw, &end);
    return (end == raw || !std::isfinite(v)) ? def : v;
}

bool Solver::create(const SsblXpbdConfig* config, const SsblXpbdMesh* mesh) {
    cfg = *config;  // A41:8797-8848
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
    pressure_accel_scale = std::max(1.0f, std::min(env_float("SSBL_ABI41_PRESSURE_UI_ACCEL_SCALE", 1792.0f), 8192.0f));
    pressure_impulse_rest_scale = std::max(0.05f, std::min(env_float("SSBL_ABI41_PRESSURE_IMPULSE_REST_SCALE", 0.85f), 8.0f));
    pressure_velocity_rest_scale = std::max(0.10f, std::min(env_float("SSBL_ABI41_PRESSURE_VELOCITY_REST_SCALE", 2.25f), 16.0f));
    // A41:266-292: an integer after strtol, used when above 0, default cadence 1
    pcg_cadence = 1;
    if (const char* raw = std::getenv("SSBL_ABI41_PCG_SUBSTEP_CADENCE")) {
        char* end = nullptr;
        const long v = std::strtol(raw, &end, 10);
        if (end != raw && v > 0 && v <= std::numeric_limits<int>::max()) pcg_cadence = static_cast<int>(v);
    }
    colliders.use_ground = cfg.use_ground;
    colliders.ground_height = cfg.ground_height;
    colliders.use_wall = cfg.use_wall;
    std::memcpy(colliders.wall_origin, cfg.wall_origin, sizeof(float) * 3);
    std::memcpy(colliders.wall_normal, cfg.wall_normal, sizeof(float) * 3);
    colliders.use_sphere = cfg.use_sphere;
    std::memcpy(colliders.sphere_center, cfg.sphere_center, sizeof(float) * 3);
    colliders.sphere_radius = cfg.sphere_radius;
    if (!lib_stream()) return fail("cannot create the HIP stream");

    const int n = nv = cfg.vertex_count;
    const float* P = mesh->positions;
    h_rest.resize(n);
    h_weff.resize(n);
    std::vector<float4> h_pos(n), h_zero(n, make_float4(0.0f, 0.0f, 0.0f, 0.0f)), h_prev(n);
    for (int i = 0; i < n; ++i) {
        const float im = mesh->inv_mass[i];
        h_weff[i] = im <= 0.0f ? 0.0f : im;   // A41:8861-8865 pinned flag
        h_rest[i] = make_float4(P[3 * i], P[3 * i + 1], P[3 * i + 2], im);
        h_pos[i] = make_float4(P[3 * i], P[3 * i + 1], P[3 * i + 2], h_weff[i]);
        h_prev[i] = make_float4(P[3 * i], P[3 * i + 1], P[3 * i + 2], 0.0f);
    }
    // stretch edges, A41:8866-8873; incidence lists in ascending edge order
    std::vector<int> edges(2 * static_cast<size_t>(cfg.edge_count));
    std::vector<float> edge_rest(cfg.edge_count);
    for (int e = 0; e < cfg.edge_count; ++e) {
        edges[2 * e] = std::max(mesh->edges ? mesh->edges[2 * e] : 0, 0);
        edges[2 * e + 1] = std::max(mesh->edges ? mesh->edges[2 * e + 1] : 0, 0);
        edge_rest[e] = mesh->edge_rest_lengths ? mesh->edge_rest_lengths[e] : 0.0f;
    }
    std::vector<int> h_adj_off(n + 1, 0);
    for (int e = 0; e < cfg.edge_count; ++e) {
        const int i = edges[2 * e], j = edges[2 * e + 1];
        if (i >= n || j >= n || i == j) continue;
        h_adj_off[i + 1]++; h_adj_off[j + 1]++;
    }
    for (int i = 0; i < n; ++i) h_adj_off[i + 1] += h_adj_off[i];
    std::vector<int2> h_adj(h_adj_off[n]);
    {
        std::vector<int> cur(h_adj_off.begin(), h_adj_off.end() - 1);
        for (int e = 0; e < cfg.edge_count; ++e) {
            const int i = edges[2 * e], j = edges[2 * e + 1];
            if (i >= n || j >= n || i == j) continue;
            int rb;
            std::memcpy(&rb, &edge_rest[e], 4);
            h_adj[cur[i]++] = make_int2(static_cast<int>(static_cast<unsigned>(j) | 0x80000000u), rb);
            h_adj[cur[j]++] = make_int2(i, rb);
        }
    }
    // slot-major copy of the first incidences (ClothView::ell)
    int max_deg = 0;
    for (int i = 0; i < n; ++i) max_deg = std::max(max_deg, h_adj_off[i + 1] - h_adj_off[i]);
    ell_width = std::min(max_deg, kEllMax);
    ell_pitch = (n + kTile - 1) / kTile * kTile;
    ell_overflow = max_deg > kEllMax ? 1 : 0;
    std::vector<int2> h_ell(static_cast<size_t>(std::max(ell_width, 1)) * std::max(ell_pitch, 1), make_int2(-1, 0));
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < std::min(h_adj_off[i + 1] - h_adj_off[i], ell_width); ++j)
            h_ell[static_cast<size_t>(j) * ell_pitch + i] = h_adj[h_adj_off[i] + j];
    // 6-byte slots (ClothView::ell16, ell_rest) for a cloth on the direct
    // stretch whose neighbours all lie within 16383 indices; the 8-byte table
    // is then not needed on the device
    bool compact_ell = !cfg.stretch_optimization_enabled && ell_width > 0;
    std::vector<short> h_ell16;
    std::vector<float> h_ell_rest;
    if (compact_ell) {
        h_ell16.assign(h_ell.size(), 0);
        h_ell_rest.assign(h_ell.size(), 0.0f);
        for (size_t k = 0; k < h_ell.size() && compact_ell; ++k) {
            const int2 a = h_ell[k];
            if (a.x == -1) continue;
            const int v = static_cast<int>(k % static_cast<size_t>(ell_pitch));
            const int d = (a.x & 0x7fffffff) - v;
            if (d < -16384 || d > 16383) { compact_ell = false; break; }
            h_ell16[k] = static_cast<short>(d * 2 + (a.x < 0 ? 1 : 0));
            std::memcpy(&h_ell_rest[k], &a.y, 4);
        }
        if (!compact_ell) {
            h_ell16.clear();
            h_ell_rest.clear();
        }
    }
    // the widest neighbour distance of the 6-byte slots decides whether the
    // stretch gather stages its tile's neighbourhood in LDS
    stage_halo = 0;
    if (compact_ell) {
        int halo = 0;
        for (short code : h_ell16) halo = std::max(halo, std::abs(code >> 1));
        stage_halo = halo <= kStageHaloMax ? halo : 0;
    }
    if (compact_ell) h_ell.clear();
    edge_colors.clear();
    if (cfg.edge_color_count > 0) {
        if (!mesh->edge_color_offsets) return fail("ABI41 edge color offsets are required");
        edge_colors.assign(mesh->edge_color_offsets, mesh->edge_color_offsets + cfg.edge_color_count + 1);
    }
    h_sedges.resize(cfg.edge_count);
    for (int e = 0; e < cfg.edge_count; ++e) {
        int rb;
        std::memcpy(&rb, &edge_rest[e], 4);
        h_sedges[e] = make_int4(edges[2 * e], edges[2 * e + 1], rb, 0);
    }
    edge_coloring_valid = !edge_colors.empty();
    if (edge_coloring_valid) {
        std::vector<int> stamp(n, -1);
        for (int c = 0; c < cfg.edge_color_count && edge_coloring_valid; ++c) {
            const int lo = std::max(edge_colors[c], 0), hi = std::min(edge_colors[c + 1], cfg.edge_count);
            for (int e = lo; e < hi; ++e) {
                const int i = edges[2 * e], j = edges[2 * e + 1];
                if (i >= n || j >= n) continue;
                if (stamp[i] == c || stamp[j] == c) { edge_coloring_valid = false; break; }
                stamp[i] = c; stamp[j] = c;
            }
        }
    }
    if (cfg.bend_count > 0 && (!mesh->bends || !mesh->bend_rest_lengths))
        return fail("ABI41 bend constraints require bend pairs and rest lengths");
    if (cfg.lra_count > 0 && (!mesh->lra_edges || !mesh->lra_rest_lengths))
        return fail("ABI41 LRA constraints require edge pairs and rest lengths");
    bend_colors.clear();
    if (cfg.bend_color_count > 0) {
        if (!mesh->bend_color_offsets) return fail("ABI41 bend color offsets are required");
        bend_colors.assign(mesh->bend_color_offsets, mesh->bend_color_offsets + cfg.bend_color_count + 1);
    }
    std::vector<int2> h_bends(cfg.bend_count);
    std::vector<float> h_bend_rest(cfg.bend_count);
    for (int b = 0; b < cfg.bend_count; ++b) {
        h_bends[b] = make_int2(mesh->bends[2 * b], mesh->bends[2 * b + 1]);
        h_bend_rest[b] = mesh->bend_rest_lengths[b];
    }
    // The coloured bend pass is a Gauss-Seidel sweep only when no colour
    // touches a vertex twice; otherwise every colour range is treated as one.
    bend_coloring_valid = !bend_colors.empty();
    if (bend_coloring_valid) {
        std::vector<int> stamp(n, -1);
        for (int c = 0; c < cfg.bend_color_count && bend_coloring_valid; ++c) {
            const int lo = std::max(bend_colors[c], 0), hi = std::min(bend_colors[c + 1], cfg.bend_count);
            for (int b = lo; b < hi; ++b) {
                const int i = h_bends[b].x, j = h_bends[b].y;
                if (i < 0 || j < 0 || i >= n || j >= n) continue;
                if (stamp[i] == c || stamp[j] == c) { bend_coloring_valid = false; break; }
                stamp[i] = c; stamp[j] = c;
            }
        }
    }
    // without a valid colouring the bends run as one Jacobi sweep, gathered
    // per vertex in ascending bend order (A41:2805-2884 made deterministic)
    std::vector<int> h_bend_off(n + 1, 0);
    std::vector<int4> h_bend_adj;
    if (cfg.bend_count > 0 && !bend_coloring_valid) {
        for (int b = 0; b < cfg.bend_count; ++b) {
            const int i = h_bends[b].x, j = h_bends[b].y;
            if (i < 0 || j < 0 || i >= n || j >= n || i == j) continue;
            h_bend_off[i + 1]++; h_bend_off[j + 1]++;
        }
        for (int i = 0; i < n; ++i) h_bend_off[i + 1] += h_bend_off[i];
        h_bend_adj.resize(h_bend_off[n]);
        std::vector<int> cur(h_bend_off.begin(), h_bend_off.end() - 1);
        for (int b = 0; b < cfg.bend_count; ++b) {
            const int i = h_bends[b].x, j = h_bends[b].y;
            if (i < 0 || j < 0 || i >= n || j >= n || i == j) continue;
            int rb;
            std::memcpy(&rb, &h_bend_rest[b], 4);
            h_bend_adj[cur[i]++] = make_int4(static_cast<int>(static_cast<unsigned>(j) | 0x80000000u), rb, b, 0);
            h_bend_adj[cur[j]++] = make_int4(i, rb, b, 0);
        }
    }
    bend_gather = !h_bend_adj.empty();
    h_bend_segs.clear();
    bend_seg_first.clear();
    bend_seg_count.clear();
    for (int c = 0; c < (bend_coloring_valid ? cfg.bend_color_count : 0); ++c) {
        const int lo = std::max(bend_colors[c], 0), hi = std::min(bend_colors[c + 1], cfg.bend_count);
        bend_seg_first.push_back(static_cast<int>(h_bend_segs.size()));
        for (int b = lo; b < hi; b += kTile) h_bend_segs.push_back(make_int4(0, b, std::min(hi, b + kTile), 0));
        bend_seg_count.push_back(static_cast<int>(h_bend_segs.size()) - bend_seg_first.back());
    }
    // LRA incidences, ascending constraint order
    std::vector<int> h_lra_off(n + 1, 0);
    std::vector<int4> h_lra;
    lra_invalid = 0;
    for (int k = 0; k < cfg.lra_count; ++k) {
        const int i = mesh->lra_edges[2 * k], j = mesh->lra_edges[2 * k + 1];
        if (i < 0 || j < 0 || i >= n || j >= n || i == j) { lra_invalid++; continue; }
        h_lra_off[i + 1]++; h_lra_off[j + 1]++;
    }
    for (int i = 0; i < n; ++i) h_lra_off[i + 1] += h_lra_off[i];
    h_lra.resize(h_lra_off[n]);
    {
        std::vector<int> cur(h_lra_off.begin(), h_lra_off.end() - 1);
        for (int k = 0; k < cfg.lra_count; ++k) {
            const int i = mesh->lra_edges[2 * k], j = mesh->lra_edges[2 * k + 1];
            if (i < 0 || j < 0 || i >= n || j >= n || i == j) continue;
            int rb;
            std::memcpy(&rb, &mesh->lra_rest_lengths[k], 4);
            h_lra[cur[i]++] = make_int4(static_cast<int>(static_cast<unsigned>(j) | 0x80000000u), rb, k, 0);
            h_lra[cur[j]++] = make_int4(i, rb, k, 0);
        }
    }
    // triangles and the incident-triangle lists (A41:7051-7109)
    std::vector<int> tri(3 * static_cast<size_t>(cfg.triangle_count));
    for (int t = 0; t < 3 * cfg.triangle_count; ++t) tri[t] = mesh->triangles ? std::max(mesh->triangles[t], 0) : 0;
    std::vector<int4> h_tris(cfg.triangle_count);
    for (int t = 0; t < cfg.triangle_count; ++t) h_tris[t] = make_int4(tri[3 * t], tri[3 * t + 1], tri[3 * t + 2], 0);
    std::vector<int> h_tri_off(n + 1, 0);
    auto tri_ok = [&](int t) { return tri[3 * t] < n && tri[3 * t + 1] < n && tri[3 * t + 2] < n; };
    for (int t = 0; t < cfg.triangle_count; ++t)
        if (tri_ok(t)) for (int k = 0; k < 3; ++k) h_tri_off[tri[3 * t + k] + 1]++;
    for (int i = 0; i < n; ++i) h_tri_off[i + 1] += h_tri_off[i];
    std::vector<int> h_tri_inc(h_tri_off[n]);
    {
        std::vector<int> cur(h_tri_off.begin(), h_tri_off.end() - 1);
        for (int t = 0; t < cfg.triangle_count; ++t)
            if (tri_ok(t)) for (int k = 0; k < 3; ++k) h_tri_inc[cur[tri[3 * t + k]]++] = t;
    }
    std::vector<float> h_share, h_scale;
    pressure_setup(*this, P, tri, edges, edge_rest, h_share, h_scale);
    const std::vector<int> h_tri_order = spatial_order(h_pos, tri, cfg.triangle_count);

    ntiles = (n + kTile - 1) / kTile;
    std::vector<Tile> h_tiles(ntiles);
    for (int t = 0; t < ntiles; ++t) h_tiles[t] = Tile{0, t * kTile, std::min(n, (t + 1) * kTile), 0};
    std::vector<int> h_pin_off(n + 1, 0);

    if (!sync_stream()) return false;
    bool ok = pos0.upload(h_pos, "position allocation") && pos1.upload(h_pos, "position allocation") &&
              prev.upload(h_prev, "previous position allocation") && vel.upload(h_zero, "velocity allocation") &&
              rest.upload(h_rest, "rest position allocation") && adj_off.upload(h_adj_off, "edge incidence allocation") &&
              adj.upload(h_adj, "edge incidence allocation") && ell.upload(h_ell, "edge incidence allocation") &&
              ell16.upload(h_ell16, "edge incidence allocation") && ell_rest.upload(h_ell_rest, "edge incidence allocation") && lra_off.upload(h_lra_off, "LRA incidence allocation") &&
              lra_adj.upload(h_lra, "LRA incidence allocation") && tri_off.upload(h_tri_off, "triangle incidence allocation") &&
              tri_inc.upload(h_tri_inc, "triangle incidence allocation") && tris.upload(h_tris, "triangle allocation") && tri_order.upload(h_tri_order, "triangle allocation") &&
              bends.upload(h_bends, "bend allocation") && bend_rest.upload(h_bend_rest, "bend rest allocation") &&
              bend_off.upload(h_bend_off, "bend incidence allocation") && bend_adj.upload(h_bend_adj, "bend incidence allocation") &&
              bend_segs.upload(h_bend_segs, "bend segment allocation") && area_share.upload(h_share, "pressure share allocation") &&
              rest_scale.upload(h_scale, "pressure scale allocation") && pin_off.upload(h_pin_off, "pin allocation") &&
              counts.alloc(kCountSlots, "diagnostic allocation") && max_smoothed.alloc(1, "diagnostic allocation") &&
              d_view.alloc(1, "view allocation") && d_tiles.upload(h_tiles, "tile allocation") &&
              tile_work.alloc(dynamic_work_ints(std::max(ntiles, 1)), "tile allocation") &&
              tile_order.alloc(dynamic_order_ints(std::max(ntiles, 1)), "tile allocation") && dyn_best.alloc(n, "dynamic contact allocation");
    if (!ok || !staging.reserve(sizeof(ClothView))) return false;
    if (!hip_ok(hipMemset(counts.p, 0, sizeof(unsigned long long) * kCountSlots), "diagnostic reset")) return false;
    if (!hip_ok(hipMemset(tile_work.p, 0, sizeof(int) * tile_work.n), "tile reset")) return false;
    if (n > 0 && !hip_ok(hipMemset(dyn_best.p, 0, sizeof(unsigned long long) * dyn_best.n), "dynamic contact reset")) return false;
    if (!hip_ok(hipMemset(max_smoothed.p, 0, sizeof(float)), "diagnostic reset")) return false;
    if (cfg.self_collision) {
        // the self-collision hashes visit vertices, triangles and edges in
        // Morton order of their rest positions (a vertex as a degenerate
        // triangle, an edge as the triangle of its ends)
        std::vector<int> as_tri(3 * static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) as_tri[3 * i] = as_tri[3 * i + 1] = as_tri[3 * i + 2] = i;
        const std::vector<int> v_order = spatial_order(h_pos, as_tri, n);
        const int ne = static_cast<int>(edges.size() / 2);
        as_tri.assign(3 * static_cast<size_t>(ne), 0);
        for (int e = 0; e < ne; ++e) {
            as_tri[3 * e] = edges[2 * e];
            as_tri[3 * e + 1] = as_tri[3 * e + 2] = edges[2 * e + 1];
        }
        const std::vector<int> e_order = spatial_order(h_pos, as_tri, ne);
        if (!self.prepare(cfg, edges, tri, n, counts.p, v_order, h_tri_order, e_order)) return false;
    }
    if (cfg.stretch_optimization_enabled && !pcg.prepare(cfg, edges, edge_rest, edge_colors, n, h_adj_off[n], ntiles, ell_width, ell_pitch,
                                                           ell_overflow, counts.p))
        return false;
    if (pcg.enabled() && !pcg_group.build({PcgMember{0, &pcg, n}}, pcg_cadence)) return false;
    // A41:7193-7302: bending "wings" (an edge met again with another opposite
    // vertex); the reference builds them for its diagnostics only.
    wing_count = 0;
    if (cfg.triangle_count > 1) {
        std::unordered_map<uint64_t, int> owner;
        for (int t = 0; t < cfg.triangle_count; ++t) {
            const int ids[3] = {tri[3 * t], tri[3 * t + 1], tri[3 * t + 2]};
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
    diag.finite_flag = 1;
    const int stc = std::max(cfg.static_triangle_count, 0);
    if (stc > 0 && !mesh->static_triangles) return fail("static triangle data is required when triangle count is nonzero");
    return build_sdf(mesh->static_triangles, stc);
}

// A41:6026-6063 closed-soup test on a 1e-6 welding grid.
struct QKey { long long x, y, z; bool operator==(const QKey& o) const { return x == o.x && y == o.y && z == o.z; } };
struct QKeyHash {
    size_t operator()(const QKey& k) const {
        uint64_t h = static_cast<uint64_t>(k.x) * 73856093ull ^ static_cast<uint64_t>(k.y) * 19349663ull ^ static_cast<uint64_t>(k.z) * 83492791ull;
        h ^= h >> 33; h *= 0xff51afd7ed558ccdull; h ^= h >> 33;
        return static_cast<size_t>(h);
    }
};
static bool closed_soup(const float* t, int count) {
    if (!t || count <= 0) return false;
    std::unordered_map<QKey, int, QKeyHash> ids;
    std::unordered_map<uint64_t, int> edge_uses;
    auto vid = [&](const float* v) {
        const QKey k{std::llround(double(v[0]) * 1e6), std::llround(double(v[1]) * 1e6), std::llround(double(v[2]) * 1e6)};
        auto f = ids.find(k);
        if (f != ids.end()) return f->second;
        const int id = static_cast<int>(ids.size());
        ids.emplace(k, id);
        return id;
    };
    auto key = [](int a, int b) { return (uint64_t(unsigned(std::min(a, b))) << 32) | uint64_t(unsigned(std::max(a, b))); };
    for (int i = 0; i < count; ++i) {
        const float* a = t + 9 * i;
        for (int k = 0; k < 9; ++k) if (!std::isfinite(a[k])) return false;
        const int ia = vid(a), ib = vid(a + 3), ic = vid(a + 6);
        if (ia == ib || ib == ic || ic == ia) continue;
        edge_uses[key(ia, ib)]++; edge_uses[key(ib, ic)]++; edge_uses[key(ic, ia)]++;
    }
    if (edge_uses.empty()) return false;
    for (auto& kv : edge_uses) if (kv.second != 2) return false;
    return true;
}

// A41:6123-6218 grid set-up on the host, distances on the device.
bool Solver::build_sdf(const float* triangles, int count) {
    sdf_ready = 0;
    sdf_dim[0] = sdf_dim[1] = sdf_dim[2] = 0;
    sdf_signed = 0;
    nstatic = 0;
    cfg.static_triangle_count = count;
    if (count <= 0) { populate_sdf_diag(); return true; }
    const auto t0 = std::chrono::steady_clock::now();
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < 3 * count; ++i) {
        const float* v = triangles + 3 * i;
        if (!std::isfinite(v[0]) || !std::isfinite(v[1]) || !std::isfinite(v[2]))
            return fail("static triangle data contains non-finite coordinates");
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], v[k]); hi[k] = std::max(hi[k], v[k]); }
    }
    float voxel = (std::isfinite(cfg.static_sdf_voxel_size) && cfg.static_sdf_voxel_size > 0.0f)
                      ? cfg.static_sdf_voxel_size : std::max(std::max(cfg.collision_margin, 0.0f) * 0.5f, 0.002f);
    const int band = std::max(cfg.static_sdf_band_voxels > 0 ? cfg.static_sdf_band_voxels : 4, 1);
    const int max_res = std::max(cfg.static_sdf_max_resolution > 0 ? cfg.static_sdf_max_resolution : 160, 16);
    const float expand = std::max(cfg.collision_margin, 0.0f) + static_cast<float>(band) * voxel + voxel * 2.0f;
    float ext[3];
    for (int k = 0; k < 3; ++k) { lo[k] -= expand; hi[k] += expand; ext[k] = std::max(hi[k] - lo[k], voxel); }
    const float largest = std::max(ext[0], std::max(ext[1], ext[2]));
    if (largest / voxel + 1.0f > static_cast<float>(max_res)) voxel = largest / static_cast<float>(std::max(max_res - 1, 1));
    voxel = std::max(voxel, 1.0e-5f);
    int dim[3];
    for (int k = 0; k < 3; ++k) dim[k] = std::min(std::max(static_cast<int>(std::ceil(ext[k] / voxel)) + 1, 2), max_res);
    const long long voxels = 1LL * dim[0] * dim[1] * dim[2];
    if (voxels <= 0 || voxels > std::numeric_limits<int>::max()) return fail("static SDF grid is too large");
    const int sgn = closed_soup(triangles, count) ? 1 : 0;
    if (!sgn) sdf_unsigned++;
    if (!sync_stream()) return false;
    std::vector<float> h_tris(triangles, triangles + 9 * static_cast<size_t>(count));
    if (!static_tris.upload(h_tris, "static triangle upload")) return false;
    if (sdf.n < static_cast<size_t>(voxels) && !sdf.alloc(static_cast<size_t>(voxels), "static SDF allocation")) return false;
    launch_sdf_build(lib_stream(), sdf.p, dim[0], dim[1], dim[2], lo[0], lo[1], lo[2], voxel, static_tris.p, count, sgn);
    if (!hip_ok(hipGetLastError(), "build static SDF") || !sync_stream()) return false;
    for (int k = 0; k < 3; ++k) { sdf_dim[k] = dim[k]; sdf_origin[k] = lo[k]; }
    sdf_voxel = voxel;
    sdf_signed = sgn;
    nstatic = count;
    sdf_ready = 1;
    sdf_rebuilds++;
    sdf_build_ms = static_cast<float>(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    populate_sdf_diag();
    return true;
}

// A static SDF given as a grid (the reference builds its grid from static
// triangles only, A41:6123-6218).  Sampling, normals and contacts are those
// of a built grid; static triangles uploaded before stay for the normal
// fallback of A41:1749-1769.
bool Solver::set_sdf_grid(const int* dims, const float* ov, const float* values, int count) {
    if (!dims || !ov || !values) return fail("static SDF grid, origin and values are required");
    const long long voxels = 1LL * dims[0] * dims[1] * dims[2];
    if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2 || voxels > std::numeric_limits<int>::max() || count != voxels)
        return fail("static SDF grid needs at least 2 voxels per axis and dx*dy*dz values");
    if (!std::isfinite(ov[0]) || !std::isfinite(ov[1]) || !std::isfinite(ov[2]) || !std::isfinite(ov[3]) || ov[3] <= 0.0f)
        return fail("static SDF origin and voxel size must be finite, the voxel size positive");
    const auto t0 = std::chrono::steady_clock::now();
    if (!sync_stream()) return false;
    std::vector<float> h(values, values + voxels);
    if (!sdf.upload(h, "static SDF upload")) return false;
    for (int k = 0; k < 3; ++k) { sdf_dim[k] = dims[k]; sdf_origin[k] = ov[k]; }
    sdf_voxel = ov[3];
    sdf_signed = 1;
    sdf_ready = 1;
    sdf_rebuilds++;
    sdf_build_ms = static_cast<float>(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    populate_sdf_diag();
    return true;
}

// A41:6996-7049.  Pins are kept on the host and turned into the per-vertex
// device lists with their relaxation factor (A41:3466-3477), which depends on
// the iteration count of the step.
bool Solver::upload_pins(const int* idx, const float* p, const float* w, int count) {
    if (count < 0) return fail("invalid pin count");
    if (count > 0 && (!idx || !p)) return fail("pin arrays are required when pin count is nonzero");
    pins.clear();
    for (int k = 0; k < count; ++k) pins.push_back(HostPin{idx[k], {p[3 * k], p[3 * k + 1], p[3 * k + 2]}, w ? w[k] : 1.0f});
    pins_iterations = -1;
    // the device table is sized here, so that refreshing it in a step (a new
    // iteration count changes the relaxation) allocates nothing
    size_t in_range = 0;
    for (const HostPin& pin : pins) in_range += (pin.index >= 0 && pin.index < nv) ? 1 : 0;
    if (pin_off.n == static_cast<size_t>(nv) + 1 && pin_ids.n == in_range && d_pins.n == in_range) return true;
    if (!sync_stream()) return false;
    return pin_off.alloc(static_cast<size_t>(nv) + 1, "pin allocation") && pin_ids.alloc(in_range, "pin allocation") &&
           d_pins.alloc(in_range, "pin allocation");
}

bool Solver::refresh_pins(int iterations) {
    if (pins_iterations == iterations) return true;
    const float exponent = 1.0f / static_cast<float>(std::max(iterations + 1, 1));
    std::vector<int> off(nv + 1, 0), ids;
    std::vector<float4> h(pins.size());
    for (size_t k = 0; k < pins.size(); ++k) {
        const HostPin& pin = pins[k];
        const float w = pin.weight;
        float relax;
        if (!std::isfinite(w) || w < 0.05f) relax = -1.0f;
        else if (w >= 0.75f) relax = 2.0f;
        else {
            const float total = std::min(std::max(w / 0.75f, 0.0f), 1.0f);
            if (total >= 1.0f) relax = 1.0f;
            else {
                const float remaining = std::max(1.0f - total, 0.0f);
                const float r = 1.0f - std::pow(remaining, std::max(exponent, 1.0e-4f));
                relax = std::isfinite(r) ? std::min(std::max(r, 0.0f), 1.0f) : total;
            }
        }
        h[k] = make_float4(pin.target[0], pin.target[1], pin.target[2], relax);
        if (pin.index >= 0 && pin.index < nv) off[pin.index + 1]++;
    }
    for (int i = 0; i < nv; ++i) off[i + 1] += off[i];
    ids.assign(off[nv], 0);
    std::vector<int> cur(off.begin(), off.end() - 1);
    for (size_t k = 0; k < pins.size(); ++k)
        if (pins[k].index >= 0 && pins[k].index < nv) ids[cur[pins[k].index]++] = static_cast<int>(k);
    // the device records in the same vertex order as ids, so a vertex's pins
    // are records pin_off[v] .. pin_off[v + 1] with no index between
    std::vector<float4> by_vertex(ids.size());
    for (size_t j = 0; j < ids.size(); ++j) by_vertex[j] = h[static_cast<size_t>(ids[j])];
    if (!sync_stream()) return false;
    if (!pin_off.upload(off, "pin upload") || !pin_ids.upload(ids, "pin upload") || !d_pins.upload(by_vertex, "pin upload")) return false;
    pins_iterations = iterations;
    return true;
}

bool Solver::upload_fields(const SsblXpbdForceField* f, int count, int unsupported) {  // A41:5920-5964
    if (count < 0 || count > kMaxForceFields) return fail("force field count exceeds SSBL maximum");
    unsupported_fields = std::max(unsupported, 0);
    if (count > 0 && !f) return fail("missing force field data");
    fields.assign(f, f + count);
    if (!sync_stream()) return false;
    return d_fields.upload(fields, "force field upload");
}

// Mode 1 (coloured Gauss-Seidel) needs a valid edge colouring: its one
// launch per colour updates both endpoints in place.
bool Solver::set_stretch_mode(int mode) {
    if (mode == SSBL_EXT_STRETCH_COLORED) {
        if (!edge_coloring_valid)
            return fail("coloured stretch needs a valid edge colouring (edge color offsets, no vertex twice in one colour)");
        // cloths on the hard-stretch solve keep the mode but run no
        // direct stretch: no segments, so no colour launch reaches them
        if (!cfg.stretch_optimization_enabled && sedges.n != h_sedges.size()) {
            h_stretch_segs.clear();
            stretch_seg_first.clear();
            stretch_seg_count.clear();
            for (int c = 0; c < cfg.edge_color_count; ++c) {
                const int lo = std::max(edge_colors[c], 0), hi = std::min(edge_colors[c + 1], cfg.edge_count);
                stretch_seg_first.push_back(static_cast<int>(h_stretch_segs.size()));
                for (int e = lo; e < hi; e += kTile) h_stretch_segs.push_back(make_int4(0, e, std::min(hi, e + kTile), 0));
                stretch_seg_count.push_back(static_cast<int>(h_stretch_segs.size()) - stretch_seg_first.back());
            }
            if (!sync_stream() || !sedges.upload(h_sedges, "stretch edge allocation") ||
                !stretch_segs.upload(h_stretch_segs, "stretch segment allocation"))
                return false;
        }
    }
    stretch_mode = mode;
    return true;
}

void Solver::fill_view(int substeps) {
    ClothView& v = view;
    v = ClothView{};
    v.nv = nv; v.ne = cfg.edge_count; v.nt = cfg.triangle_count; v.nbend = cfg.bend_count; v.nlra = cfg.lra_count;
    v.npin = static_cast<int>(pins.size()); v.nfield = static_cast<int>(fields.size());
    v.pos0 = pos0.p; v.pos1 = pos1.p; v.prev = prev.p; v.vel = vel.p; v.rest = rest.p;
    v.adj_off = adj_off.p; v.adj = adj.p;
    v.ell = ell.p; v.ell16 = ell16.p; v.ell_rest = ell_rest.p; v.ell_width = ell_width; v.ell_pitch = ell_pitch; v.ell_overflow = ell_overflow; v.stage_halo = stage_halo; v.lra_off = lra_off.p; v.lra_adj = lra_adj.p;
    v.tri_off = tri_off.p; v.tri_inc = tri_inc.p; v.tris = tris.p; v.bends = bends.p; v.bend_rest = bend_rest.p;
    v.bend_off = bend_off.p; v.bend_adj = bend_adj.p;
    v.pin_off = pin_off.p; v.pin_ids = pin_ids.p; v.pins = d_pins.p;
    v.area_share = area_share.p; v.rest_scale = rest_scale.p; v.fields = d_fields.p;
    v.sdf = sdf.p; v.static_tris = static_tris.p; v.nstatic = nstatic;
    v.counts = counts.p; v.max_smoothed = max_smoothed.p;
    const float dt = cfg.dt / static_cast<float>(substeps);
    v.dt_sub = dt;
    v.damping = cfg.damping;
    for (int k = 0; k < 3; ++k) v.gravity[k] = cfg.gravity[k];
    v.stretch_alpha = std::max(cfg.stretch_compliance, 0.0f) / std::max(dt * dt, 1.0e-7f);
    v.bend_alpha = std::max(cfg.bend_compliance, 0.0f) / std::max(dt * dt, 1.0e-7f);
    v.lra_compliance = cfg.lra_compliance;
    v.collision_margin = cfg.collision_margin;
    v.thickness = cfg.cloth_thickness;
    v.friction = cfg.contact_friction;
    v.tangent_damping = cfg.contact_tangent_damping;
    v.use_pressure = (cfg.use_volume_pressure && cfg.pressure_strength > 0.0f) ? 1 : 0;
    v.pressure_strength = cfg.pressure_strength;
    v.pressure_accel_scale = pressure_accel_scale;
    v.pressure_impulse_rest_scale = pressure_impulse_rest_scale;
    v.pressure_velocity_rest_scale = pressure_velocity_rest_scale;
    v.pressure_velocity_length_scale = pressure_velocity_length_scale;
    v.colliders = colliders;
    v.sdf_ready = sdf_ready;
    for (int k = 0; k < 3; ++k) { v.sdf_dim[k] = sdf_dim[k]; v.sdf_origin[k] = sdf_origin[k]; }
    v.sdf_voxel = sdf_voxel;
    v.sdf_inv = sdf_voxel > 0.0f ? 1.0f / sdf_voxel : 0.0f;
    v.sdf_inv_clamped = 1.0f / std::max(sdf_voxel, 1.0e-7f);   // the kernels' eps
    for (int k = 0; k < 3; ++k) v.sdf_top[k] = static_cast<float>(sdf_dim[k] - 1);
    v.need_normal = (v.use_pressure || !fields.empty()) ? 1 : 0;
    v.direct_stretch = cfg.stretch_optimization_enabled ? 0 : 1;
    v.contacts_on = (analytic_on() || sdf_ready) ? 1 : 0;
    // a signed own set replaces the unsigned contacts (A41:9240-9252); an
    // attached scene is never signed (A41:6561)
    const bool own_signed = dyn.is_signed() && (dyn.triangles() > 0 || dyn.particles() > 0);
    v.dyn = (dyn_on() && (use_scene() || !dyn.is_signed())) ? dyn_set().device_view() : nullptr;
    v.sdyn = own_signed ? dyn.device_view() : nullptr;
    v.dyn_best = dyn_best.p;
    v.collider_dt = (std::isfinite(cfg.dt) && cfg.dt > 1.0e-7f) ? cfg.dt : std::max(v.dt_sub, 1.0e-7f);   // A41:4513-4518
    v.dyn_mode = use_scene() ? cross_mode : 0;
    v.dyn_target_src = target_src;
    v.dyn_target_layer = target_layer;
    v.sedges = colored_stretch() ? sedges.p : nullptr;
    v.pcg = pcg.enabled() ? pcg.view() : PcgView{};
}

// ------------------------------------------------------------ schedule
// One step (A41:9165-9378) of every member of `members`; one launch per
// phase covers all of them.  `ctx` addresses their views and tiles;
// `dviews[m]` is member m's view inside that table.
struct StepPlan {
    bool any_pins = false, any_edges = false, any_lra = false, any_bend = false, any_self = false, any_pcg = false;
    bool any_bend_gather = false, any_forces = false, any_dyn = false, any_signed = false;
    bool colored = false;   // direct stretch as coloured Gauss-Seidel (every member shares the mode)
};

static StepPlan plan_of(const std::vector<Solver*>& members) {
    StepPlan plan;
    for (const Solver* s : members) {
        plan.any_pins |= !s->pins.empty();
        plan.any_edges |= s->cfg.edge_count > 0 && !s->cfg.stretch_optimization_enabled;
        plan.any_lra |= s->cfg.lra_count > 0;
        plan.any_bend |= s->cfg.bend_count > 0;
        plan.any_self |= s->cfg.self_collision && s->nv > 0 && s->self_radius() > 0.0f;
        plan.any_pcg |= s->pcg.enabled();
        plan.any_bend_gather |= s->bend_gather;
        plan.any_forces |= s->view.need_normal != 0 || s->view.nfield > 0;
        plan.any_dyn |= s->view.dyn != nullptr;
        plan.any_signed |= s->view.sdyn != nullptr;
        plan.colored |= s->colored_stretch();
    }
    return plan;
}

// The graph key of a step: the launch sequence of run_schedule is a function
// of these values alone, for plans without self collision or the hard-stretch
// solve (whose launches depend on more host state and are never captured).
static bool graph_key(const StepPlan& plan, int substeps, int iterations, bool count, const PoolTuner& tuner,
                      std::vector<long long>* key) {
    if (g_graph_mode.load(std::memory_order_relaxed) == 0 || timing_enabled()) return false;
    if (plan.any_dyn && tuner.tuning()) return false;
    // the self-collision and hard-stretch launches are functions of the
    // members' fixed set-up (intervals, passes, cadences, tile tables) and of
    // the flags below
    *key = {substeps, iterations, count ? 1 : 0, plan.any_pins, plan.any_edges, plan.any_lra, plan.any_bend,
            plan.any_bend_gather, plan.any_forces, plan.any_dyn, plan.any_signed, plan.colored, plan.any_self, plan.any_pcg,
            tuner.value()};
    return true;
}

// Per-colour segment tables ({cloth, begin, end}) of a handle or batch.
struct ColorSegs {
    const int4* dev;
    const std::vector<int>* first;
    const std::vector<int>* count;
};

static bool run_schedule(std::vector<Solver*>& members, const LaunchCtx& ctx, const ClothView* dviews, const ColorSegs& bends,
                         const ColorSegs& stretch_colors, const std::vector<const PcgGroup*>& pcg_groups, SelfBatch* self_batch,
                         int substeps, int iterations, bool count, PoolTuner& tuner, bool tune) {
    const StepPlan plan = plan_of(members);
    const unsigned cnt = count ? kOpCount : 0u;
    const unsigned pin_ops = plan.any_pins ? (kOpPin | kOpPinPolish) : 0u;
    const unsigned stretch = plan.any_edges ? kOpStretch : 0u;
    hipStream_t st = ctx.stream;
    int parity = 0;
    // The direct stretch projection followed by the vertex-local work `ops`:
    // the Jacobi gather fused with them in one pass (src -> dst), or, in the
    // coloured mode, one in-place launch per colour and then `ops` in place.
    auto stretch_then = [&](unsigned ops) {
        if (plan.colored && plan.any_edges) {
            for (size_t c = 0; c < stretch_colors.first->size(); ++c)
                launch_stretch_color(ctx, stretch_colors.dev + (*stretch_colors.first)[c], (*stretch_colors.count)[c], parity, cnt);
            if (ops & ~kOpCount) launch_vertex_pass(ctx, parity, parity, ops);
            return;
        }
        launch_vertex_pass(ctx, parity, parity ^ 1, stretch | ops);
        parity ^= 1;
    };
    // the dynamic contact pass takes the tiles that searched most in the step
    // before first
    if (plan.any_dyn) launch_order_tiles(ctx);
    // the hard-stretch scalars of the step start at zero (A41:7586-7624)
    if (plan.any_pcg)
        for (const PcgGroup* g : pcg_groups) g->reset(st, dviews);
    bool integrated = false;   // the previous substep's last launch integrated this one
    for (int sub = 0; sub < substeps; ++sub) {
        if (!integrated) {
            launch_integrate(ctx, parity, plan.any_pins ? kOpPin : 0u, plan.any_forces);
            parity ^= 1;
        }
        integrated = false;
        for (int it = 0; it < iterations; ++it) {
            const bool last = it == iterations - 1;
            const bool final_c = last && sub == substeps - 1;
            const unsigned last_op = last ? kOpLastIter : 0u;
            const bool dyn_now = last && plan.any_dyn;   // A41:9230-9252
            const bool signed_now = last && plan.any_signed;   // A41:9322-9359
            bool self_now = false;
            if (plan.any_self && last) {
                for (Solver* m : members) {
                    const int interval = std::max(m->cfg.self_collision_interval, 1);
                    if (m->cfg.self_collision && (((sub + 1) % interval) == 0 || sub == substeps - 1)) self_now = true;
                }
            }
            // A41:9158-9177: each group's solve runs on every cadence-th
            // substep and always on the last
            bool run_pcg = false;
            if (plan.any_pcg && last)
                for (const PcgGroup* g : pcg_groups) run_pcg |= g->due(sub, substeps);
            const bool split = run_pcg || self_now || dyn_now || signed_now || (final_c && (plan.any_lra || plan.any_bend));
            if (!split) {
                stretch_then(pin_ops | kOpAnalytic | kOpSdf | last_op | (last ? kOpVelocity : 0u) | cnt);
                continue;
            }
            // A41:9178-9369 order: stretch (direct or PCG + polish + caps),
            // final LRA and bends, pins, analytic and SDF contacts, self
            // collision, final caps, second pin projection
            const bool tail = self_now || dyn_now || signed_now || (run_pcg && final_c);
            const unsigned contact_ops =
                (tail ? (pin_ops & ~kOpPinPolish) : (pin_ops | last_op | kOpVelocity)) | kOpAnalytic | kOpSdf | cnt;
            if (!run_pcg && !(final_c && (plan.any_lra || plan.any_bend))) {
                // nothing runs between the stretch gather and the pins and
                // contacts: one pass does both, as the unsplit iteration does
                stretch_then(contact_ops);
            } else {
                stretch_then(cnt);
                if (run_pcg) {
                    for (const PcgGroup* g : pcg_groups) {
                        if (!g->due(sub, substeps)) continue;
                        g->solve(st, dviews, parity);
                        g->polish(st, dviews, parity);
                    }
                }
                if (final_c && plan.any_lra) {
                    launch_vertex_pass(ctx, parity, parity ^ 1, kOpLra | cnt);
                    parity ^= 1;
                }
                if (final_c && plan.any_bend) {
                    for (size_t c = 0; c < bends.first->size(); ++c)
                        launch_bend_color(ctx, bends.dev + (*bends.first)[c], (*bends.count)[c], parity, cnt);
                    if (plan.any_bend_gather) {
                        launch_vertex_pass(ctx, parity, parity ^ 1, kOpBend | cnt);
                        parity ^= 1;
                    }
                }
                launch_vertex_pass(ctx, parity, parity, contact_ops);
            }
            const unsigned tail_ops = (pin_ops & kOpPinPolish) | last_op | kOpVelocity | cnt;
            // the dynamic contacts end the iteration: their launch
            // applies the closing steps itself, and (gravity only) the next
            // substep's integration as well
            const bool tail_fused = dyn_now && !self_now && !signed_now && !(run_pcg && final_c);
            const bool integrate_next = tail_fused && sub + 1 < substeps && !plan.any_forces;
            const unsigned fused_ops =
                tail_ops | (integrate_next ? (kOpIntegrate | (plan.any_pins ? kOpIntegratePin : 0u)) : 0u);
            if (dyn_now) {
                hipEvent_t tb = nullptr, te = nullptr;
                const int pool = tune ? tuner.candidate(sub, &tb, &te) : tuner.value();
                launch_dynamic_pass(ctx, parity, cnt, tail_fused ? fused_ops : 0u, pool, tb, te);
            }
            if (integrate_next) {
                parity ^= 1;
                integrated = true;
            }
            if (self_now) {
                if (!self_batch || !self_batch->run(ctx, parity, sub, substeps, count)) return false;
            }
            if (signed_now) launch_signed_pass(ctx, parity, false, cnt);
            if (run_pcg && final_c)
                for (const PcgGroup* g : pcg_groups) g->final_caps(st, dviews, parity);
            if (signed_now && final_c) launch_signed_pass(ctx, parity, true, cnt);
            if (tail && !tail_fused) launch_vertex_pass(ctx, parity, parity, tail_ops);
        }
    }
    if (parity != 0) launch_vertex_pass(ctx, 1, 0, 0u);
    return hip_ok(hipGetLastError(), "launch ABI41 substep");
}

static void reset_diag(Solver& s) {
    s.diag = SsblXpbdDiagnostics{};
    s.diag.finite_flag = 1;
    s.diag.force_field_count = static_cast<long long>(s.fields.size());
    s.diag.unsupported_force_field_count = s.unsupported_fields;
    s.diag.abi41_bending_texture_ready = s.wing_count > 0 ? 1 : 0;
    s.populate_sdf_diag();
    s.diag.dynamic_triangle_count = s.dyn_set().triangles();   // A41:9103-9122
    s.diag.dynamic_particle_count = s.dyn_set().particles();
    if (s.use_scene()) {
        s.diag.global_dynamic_particle_count = s.scene->set.particles();
        s.diag.global_dynamic_triangle_count = s.scene->set.triangles();
        s.diag.global_dynamic_hash_overflow = s.scene->hash_overflow;
    }
}

// Counters of a scene build that ran asynchronously (the stream is synchronised).
static void settle_scene_overflow(GlobalScene& sc) {
    if (!sc.overflow_pending) return;
    unsigned long long c[kCountSlots] = {};
    if (hip_ok(hipMemcpy(c, sc.counts.p, sizeof(c), hipMemcpyDeviceToHost), "global scene counters"))
        sc.hash_overflow = static_cast<long long>(c[kCountDynParticleOverflow] + c[kCountDynTriBucketOverflow]);
    sc.overflow_pending = false;
}

static bool fetch_counts(Solver& s) {  // A41:8059-8179
    unsigned long long c[kCountSlots] = {};
    float ms = 0.0f;
    if (!hip_ok(hipMemcpy(c, s.counts.p, sizeof(c), hipMemcpyDeviceToHost), "fetch ABI41 diagnostics")) return false;
    if (!hip_ok(hipMemcpy(&ms, s.max_smoothed.p, sizeof(float), hipMemcpyDeviceToHost), "fetch ABI41 diagnostics")) return false;
    SsblXpbdDiagnostics& d = s.diag;
    d.abi41_soft_contact_count = static_cast<long long>(c[kCountSoft]);
    d.abi41_exact_impulse_contact_count = static_cast<long long>(c[kCountExactImpulse]);
    d.abi41_edge_edge_contact_count = static_cast<long long>(c[kCountEdgeEdge]);
    d.abi41_hard_projection_fallbacks = static_cast<long long>(c[kCountHardFallbacks]);
    d.abi41_max_smoothed_delta = ms;
    d.fast_soft_repulsion_candidates = static_cast<long long>(c[kCountSelfCandidates]);
    d.fast_soft_repulsion_applied = static_cast<long long>(c[kCountSoft]);
    d.external_friction_corrections = static_cast<long long>(c[kCountFriction]);
    d.self_candidate_count = static_cast<long long>(c[kCountSelfCandidates]);
    d.static_sdf_contact_count = static_cast<long long>(c[kCountStaticSdf]);
    d.candidate_count = static_cast<long long>(c[kCountDynParticleCandidates] + c[kCountDynTriCandidates] +
                                               c[kCountSelfCandidates] + c[kCountStaticSdf]);
    // A41:8078-8110
    d.dynamic_particle_candidate_count = static_cast<long long>(c[kCountDynParticleCandidates]);
    d.dynamic_particle_contacts = static_cast<long long>(c[kCountDynParticleContacts]);
    d.dynamic_particle_overflow = 0;
    d.dynamic_triangle_candidate_count = static_cast<long long>(c[kCountDynTriCandidates]);
    d.dynamic_triangle_bucket_overflow = static_cast<long long>(c[kCountDynTriBucketOverflow]);
    d.dynamic_triangle_large_primitive_count = static_cast<long long>(c[kCountDynTriLarge]);
    d.dynamic_triangle_aabb_reject_count = static_cast<long long>(c[kCountDynTriAabbRejects]);
    d.dynamic_triangle_max_bucket_occupancy = 0;
    if (s.dyn.triangles_allocated()) {
        d.dynamic_triangle_large_primitive_count =
            std::max(d.dynamic_triangle_large_primitive_count, static_cast<long long>(s.dyn.large_count()));
        d.dynamic_triangle_max_bucket_occupancy = s.dyn.max_occupancy();
    }
    d.fast_triangle_pair_candidates = d.dynamic_triangle_candidate_count;
    d.fast_triangle_pair_contacts = static_cast<long long>(c[kCountTrianglePairs]);
    d.fast_triangle_pair_skipped_rest = d.dynamic_triangle_aabb_reject_count;
    if (s.use_scene()) {
        settle_scene_overflow(*s.scene);
        d.global_dynamic_hash_overflow = s.scene->hash_overflow;
    }
    d.abi41_lra_tack_count = static_cast<long long>(c[kCountLraTacks]);
    d.abi41_bending_wing_count = c[kCountBendingWings] > 0 ? static_cast<long long>(c[kCountBendingWings]) : s.wing_count;
    d.abi41_tack_jitter_guarded = static_cast<long long>(c[kCountTackGuards]) + (s.cfg.lra_count > 0 ? s.lra_invalid : 0);
    d.abi41_bending_guarded = static_cast<long long>(c[kCountBendingGuards]);
    // A41:8128-8160
    d.fast_overlap_island_candidates = static_cast<long long>(c[kCountIslandCandidates]);
    d.fast_overlap_island_clusters = static_cast<long long>(c[kCountIslandClusters]);
    d.fast_overlap_island_vertex_refs = static_cast<long long>(c[kCountIslandVertexRefs]);
    d.fast_overlap_island_applied_vertices = static_cast<long long>(c[kCountIslandApplied]);
    d.fast_overlap_island_guarded = static_cast<long long>(c[kCountIslandGuarded]);
    d.fast_overlap_island_max_delta = static_cast<float>(c[kCountIslandMaxDeltaMicro]) * 1.0e-6f;
    d.self_filter_seen = static_cast<long long>(c[kCountFilterSeen]);
    d.self_filter_accepted_vv = static_cast<long long>(c[kCountFilterAcceptedVv]);
    d.self_filter_accepted_vt = static_cast<long long>(c[kCountFilterAcceptedVt]);
    d.self_filter_accepted_ee = static_cast<long long>(c[kCountFilterAcceptedEe]);
    d.self_filter_rejected_rest = static_cast<long long>(c[kCountFilterRejectedRest]);
    d.self_filter_rejected_duplicate = static_cast<long long>(c[kCountFilterRejectedDuplicate]);
    d.self_filter_rejected_ownership = static_cast<long long>(c[kCountFilterRejectedOwnership]);
    d.self_filter_cache_hits = static_cast<long long>(c[kCountFilterCacheHits]);
    d.self_filter_cache_misses = static_cast<long long>(c[kCountFilterCacheMisses]);
    d.self_cluster_count = static_cast<long long>(c[kCountClusterCount]);
    d.self_cluster_owned_contacts = static_cast<long long>(c[kCountClusterOwnedContacts]);
    return s.pcg.fetch(&d);
}

static bool prepare_member(Solver& s, int substeps, int iterations, bool count) {
    if (!s.pins.empty() && !s.refresh_pins(iterations)) return false;
    s.fill_view(substeps);
    if (count) {
        launch_fill_u32(lib_stream(), reinterpret_cast<unsigned*>(s.counts.p), 0u, 2 * static_cast<size_t>(kCountSlots));
        launch_fill_u32(lib_stream(), reinterpret_cast<unsigned*>(s.max_smoothed.p), 0u, 1);
    }
    // the cloth's own dynamic colliders are re-hashed every step (A41:9129-9136)
    const bool own = (!s.use_scene() || s.dyn.is_signed()) && (s.dyn.triangles() > 0 || s.dyn.particles() > 0);
    if (own && !s.dyn.build(lib_stream(), s.counts.p, count, count)) return false;
    return true;
}

// no_alloc: the allocation-free variant -- launches directly (no graph
// capture or instantiation) and never times launch candidates.
static bool step_one(Solver& s, int substeps, int iterations, int fetch, int force_sync, bool no_alloc = false) {
    const auto t0 = std::chrono::steady_clock::now();
    reset_diag(s);
    substeps = std::max(substeps, 1);
    iterations = std::max(iterations, 1);
    const bool count = fetch != 0;
    if (!prepare_member(s, substeps, iterations, count)) return false;
    if (!s.staging.upload(s.d_view.p, &s.view, sizeof(ClothView), lib_stream())) return false;
    if (s.nv > 0) {
        LaunchCtx ctx{lib_stream(), s.d_view.p, s.d_tiles.p, s.ntiles, s.tile_work.p, s.tile_order.p};
        std::vector<Solver*> one{&s};
        SelfBatch* sb = s.cfg.self_collision ? s.self.single_batch(s.view, s.d_view.p) : nullptr;
        const StepPlan plan = plan_of(one);
        if (plan.any_dyn && !no_alloc && !s.tuner.poll(s.tuner.tuning())) return false;
        const bool tune = plan.any_dyn && !no_alloc && s.tuner.tuning();
        if (tune && !s.tuner.begin_step(substeps)) return false;
        const ColorSegs bends{s.bend_segs.p, &s.bend_seg_first, &s.bend_seg_count};
        const ColorSegs stretch_colors{s.stretch_segs.p, &s.stretch_seg_first, &s.stretch_seg_count};
        std::vector<const PcgGroup*> groups;
        if (s.pcg.enabled()) groups.push_back(&s.pcg_group);
        auto record = [&] {
            return run_schedule(one, ctx, s.d_view.p, bends, stretch_colors, groups, sb, substeps, iterations, count, s.tuner, tune);
        };
        std::vector<long long> key;
        const bool graphed = !no_alloc && graph_key(plan, substeps, iterations, count, s.tuner, &key);
        if (!(graphed ? s.graph.run(lib_stream(), key, record) : record())) return false;
        if (tune) s.tuner.end_step();
        if (plan.any_dyn && !no_alloc) ++s.tuner.step;
    }
    if (!s.dyn.clear_motion(lib_stream())) return false;
    if (force_sync || fetch) {
        const auto ts = std::chrono::steady_clock::now();
        if (!sync_stream()) return false;
        s.diag.sync_ms = static_cast<float>(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count());
    }
    if (fetch && !fetch_counts(s)) return false;
    s.diag.step_ms = static_cast<float>(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    s.diag.constraints_ms = s.diag.step_ms;
    s.diag.resolved_contacts = s.diag.abi41_soft_contact_count + s.diag.abi41_exact_impulse_contact_count +
                               s.diag.abi41_edge_edge_contact_count + s.diag.static_sdf_contact_count;
    return true;
}

// ----------------------------------------------------------------- batch
struct Batch {
    std::vector<Solver*> members;
    int stretch_mode = SSBL_EXT_STRETCH_JACOBI;   // shared by every member
    Dev<ClothView> d_views;
    Dev<Tile> d_tiles;
    Dev<int> tile_work, tile_order;   // the dynamic contact pass's tile order (LaunchCtx)
    Dev<int4> d_bend_segs, d_stretch_segs;
    std::vector<int4> h_bend_segs, h_stretch_segs;
    std::vector<int> seg_first, seg_count, sseg_first, sseg_count;
    std::vector<ClothView> h_views;
    Staging staging;
    int ntiles = 0;
    int stepped_substeps = 0;   // substeps of the last step since the tables were built (0: none; d_views unfilled)
    SelfBatchHolder self;
    std::vector<PcgGroup> pcg_groups;   // members on the hard-stretch solve, one group per cadence
    StepGraph graph;
    PoolTuner tuner;
    // Two halves of the members (by vertex count), stepped concurrently on
    // two streams: their cloths share nothing within a step, so one half's
    // launch fills the GPU while the other's waits for memory or ends on a
    // few slow waves.  Empty: the batch steps as one table.
    std::vector<std::unique_ptr<Batch>> parts;
    bool own_tables = false;   // the one-table form's tables below are built (always, unless split)
    hipEvent_t fork = nullptr, join = nullptr;
    Batch() = default;
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    ~Batch() {
        if (fork) hipEventDestroy(fork);
        if (join) hipEventDestroy(join);
    }
};

std::atomic<int> g_batch_parts{2};   // ssbl_ext_set_batch_parts: 1 or 2
constexpr int kMinSplitTiles = 64;   // a batch with fewer tiles steps as one table

// Colour c of the batch: colour c's segments of every member, in member order,
// with the member index in the segment's cloth field.
static void merge_color_segs(const std::vector<Solver*>& members, std::vector<int4> Solver::*segs, std::vector<int> Solver::*first,
                             std::vector<int> Solver::*count, std::vector<int4>& out, std::vector<int>& out_first,
                             std::vector<int>& out_count) {
    size_t colors = 0;
    for (const Solver* s : members) colors = std::max(colors, (s->*first).size());
    out.clear();
    out_first.clear();
    out_count.clear();
    for (size_t c = 0; c < colors; ++c) {
        out_first.push_back(static_cast<int>(out.size()));
        for (size_t m = 0; m < members.size(); ++m) {
            const Solver* s = members[m];
            if (c >= (s->*first).size()) continue;
            for (int k = 0; k < (s->*count)[c]; ++k) {
                int4 seg = (s->*segs)[(s->*first)[c] + k];
                seg.x = static_cast<int>(m);
                out.push_back(seg);
            }
        }
        out_count.push_back(static_cast<int>(out.size()) - out_first.back());
    }
}

static bool build_tables(Batch& b) {
    std::vector<Tile> tiles;
    b.stretch_mode = b.members[0]->stretch_mode;
    for (size_t m = 0; m < b.members.size(); ++m) {
        Solver* s = b.members[m];
        for (int t = 0; t < s->ntiles; ++t) tiles.push_back(Tile{static_cast<int>(m), t * kTile, std::min(s->nv, (t + 1) * kTile), 0});
        if (s->stretch_mode != b.stretch_mode) return fail("batch members must share the stretch mode");
    }
    merge_color_segs(b.members, &Solver::h_bend_segs, &Solver::bend_seg_first, &Solver::bend_seg_count, b.h_bend_segs, b.seg_first,
                     b.seg_count);
    if (b.stretch_mode == SSBL_EXT_STRETCH_COLORED)
        merge_color_segs(b.members, &Solver::h_stretch_segs, &Solver::stretch_seg_first, &Solver::stretch_seg_count, b.h_stretch_segs,
                         b.sseg_first, b.sseg_count);
    b.ntiles = static_cast<int>(tiles.size());
    b.stepped_substeps = 0;
    b.h_views.resize(b.members.size());
    // members on the hard-stretch solve, grouped by substep cadence
    std::map<int, std::vector<PcgMember>> by_cadence;
    for (size_t m = 0; m < b.members.size(); ++m) {
        Solver* s = b.members[m];
        if (s->pcg.enabled()) by_cadence[s->pcg_cadence].push_back(PcgMember{static_cast<int>(m), &s->pcg, s->nv});
    }
    b.pcg_groups.resize(by_cadence.size());   // existing groups keep their (grow-only) tables
    size_t g = 0;
    for (auto& kv : by_cadence)
        if (!b.pcg_groups[g++].build(kv.second, kv.first)) return false;
    // grow-only: a batch given new members keeps its device tables and staging
    const int nt = std::max(static_cast<int>(tiles.size()), 1);
    if (!b.tile_work.reserve(dynamic_work_ints(nt), "batch tile allocation") ||
        !b.tile_order.reserve(dynamic_order_ints(nt), "batch tile allocation") ||
        !hip_ok(hipMemset(b.tile_work.p, 0, sizeof(int) * b.tile_work.n), "batch tile reset"))
        return false;
    return b.d_views.reserve(b.members.size(), "batch view allocation") && b.d_tiles.upload_grow(tiles, "batch tile allocation") &&
           b.d_bend_segs.upload_grow(b.h_bend_segs, "batch bend allocation") &&
           b.d_stretch_segs.upload_grow(b.h_stretch_segs, "batch stretch allocation") &&
           b.staging.reserve(sizeof(ClothView) * b.members.size());
}

// The batch's own tables, and its two parts when it is large enough: the
// members are dealt largest first to the part with fewer vertices so far,
// each part keeping the members' relative order.
static bool build_batch(Batch& b) {
    b.own_tables = false;
    int tiles = 0;
    for (const Solver* s : b.members) tiles += s->ntiles;
    if (g_batch_parts.load(std::memory_order_relaxed) < 2 || b.members.size() < 2 || tiles < kMinSplitTiles || !part_stream()) {
        b.parts.clear();
        if (!build_tables(b)) return false;
        b.own_tables = true;
        return true;
    }
    // split: the batch's own one-table form is built only when asked for
    // (launch tuning, one-table pass timing)
    b.stretch_mode = b.members[0]->stretch_mode;
    for (const Solver* s : b.members)
        if (s->stretch_mode != b.stretch_mode) return fail("batch members must share the stretch mode");
    b.ntiles = tiles;
    b.stepped_substeps = 0;
    std::vector<size_t> idx(b.members.size());
    for (size_t m = 0; m < idx.size(); ++m) idx[m] = m;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) { return b.members[x]->nv > b.members[y]->nv; });
    std::vector<char> side(b.members.size(), 0);
    long long load[2] = {0, 0};
    for (size_t m : idx) {
        const int k = load[1] < load[0] ? 1 : 0;
        side[m] = static_cast<char>(k);
        load[k] += b.members[m]->nv;
    }
    // a batch given new members keeps its parts' (grow-only) tables
    b.parts.resize(2);
    bool degenerate = false;
    for (int k = 0; k < 2; ++k) {
        if (!b.parts[k]) b.parts[k] = std::make_unique<Batch>();
        Batch& part = *b.parts[k];
        part.members.clear();
        for (size_t m = 0; m < b.members.size(); ++m)
            if (side[m] == k) part.members.push_back(b.members[m]);
        if (part.members.empty()) { degenerate = true; break; }
        if (!build_tables(part)) return false;
    }
    if (degenerate) {
        b.parts.clear();
        if (!build_tables(b)) return false;
        b.own_tables = true;
        return true;
    }
    if (!b.fork && !hip_ok(hipEventCreateWithFlags(&b.fork, hipEventDisableTiming), "batch event")) return false;
    if (!b.join && !hip_ok(hipEventCreateWithFlags(&b.join, hipEventDisableTiming), "batch event")) return false;
    return true;
}

// The one-table form of a split batch, built on first use.
static bool ensure_own_tables(Batch& b) {
    if (b.own_tables) return true;
    const int stepped = b.stepped_substeps;
    if (!build_tables(b)) return false;
    b.stepped_substeps = stepped;
    b.own_tables = true;
    return true;
}

// New members for an existing batch: its tables are rebuilt in place (no
// allocation unless they grow) and its graphs dropped.  Waits for the
// library's queued work, which may still read the old tables.
static bool retarget_batch(Batch& b, const std::vector<Solver*>& members) {
    if (!sync_stream() || !b.graph.clear(lib_stream())) return false;
    b.members = members;
    return build_batch(b);
}

// fetch: count diagnostics and read every member's counters afterwards (as
// ssbl_step_solver_ex with fetch_diagnostics does for one handle).
static bool step_batch(Batch& b, int substeps, int iterations, int force_sync, bool fetch = false) {
    const auto t0 = std::chrono::steady_clock::now();
    substeps = std::max(substeps, 1);
    iterations = std::max(iterations, 1);
    for (const Solver* s : b.members)
        if (s->stretch_mode != b.stretch_mode) return fail("a batch member changed its stretch mode; create the batch again");
    for (size_t m = 0; m < b.members.size(); ++m) {
        Solver& s = *b.members[m];
        reset_diag(s);
        if (!prepare_member(s, substeps, iterations, fetch)) return false;
    }
    // the launch-candidate timing of the tuner runs on the whole table only
    const StepPlan plan = plan_of(b.members);
    const bool split = !b.parts.empty() && !(plan.any_dyn && b.tuner.tuning());
    if (!split) {
        if (!ensure_own_tables(b)) return false;
        for (size_t m = 0; m < b.members.size(); ++m) b.h_views[m] = b.members[m]->view;
        if (!b.staging.upload(b.d_views.p, b.h_views.data(), sizeof(ClothView) * b.h_views.size(), lib_stream())) return false;
    }
    b.stepped_substeps = substeps;
    std::vector<long long> key;
    if (split) {
        const bool graphed = graph_key(plan, substeps, iterations, fetch, b.tuner, &key);
        key.push_back(1);   // the two-part launch sequence
        // each part's views, then its schedule on its own stream between a
        // fork from and a join back to the library stream
        struct PartRun {
            LaunchCtx ctx;
            SelfBatch* sb;
            std::vector<const PcgGroup*> groups;
        };
        std::vector<PartRun> runs;
        for (size_t k = 0; k < b.parts.size(); ++k) {
            Batch& p = *b.parts[k];
            for (size_t m = 0; m < p.members.size(); ++m) p.h_views[m] = p.members[m]->view;
            if (!p.staging.upload(p.d_views.p, p.h_views.data(), sizeof(ClothView) * p.h_views.size(), lib_stream())) return false;
            p.stepped_substeps = substeps;
            PartRun r{LaunchCtx{k == 0 ? lib_stream() : part_stream(), p.d_views.p, p.d_tiles.p, p.ntiles, p.tile_work.p, p.tile_order.p},
                      p.self.get(p.members, p.h_views, p.d_views.p), {}};
            for (const PcgGroup& g : p.pcg_groups) r.groups.push_back(&g);
            runs.push_back(std::move(r));
        }
        auto record = [&] {
            if (!hip_ok(hipEventRecord(b.fork, lib_stream()), "batch fork") || !hip_ok(hipStreamWaitEvent(part_stream(), b.fork, 0), "batch fork"))
                return false;
            for (size_t k = 0; k < b.parts.size(); ++k) {
                Batch& p = *b.parts[k];
                const ColorSegs bends{p.d_bend_segs.p, &p.seg_first, &p.seg_count};
                const ColorSegs stretch_colors{p.d_stretch_segs.p, &p.sseg_first, &p.sseg_count};
                // the batch's tuned pooling limit (never timed here)
                if (!run_schedule(p.members, runs[k].ctx, p.d_views.p, bends, stretch_colors, runs[k].groups, runs[k].sb, substeps, iterations,
                                  fetch, b.tuner, false))
                    return false;
            }
            return hip_ok(hipEventRecord(b.join, part_stream()), "batch join") && hip_ok(hipStreamWaitEvent(lib_stream(), b.join, 0), "batch join");
        };
        if (!(graphed ? b.graph.run(lib_stream(), key, record) : record())) return false;
    } else {
        LaunchCtx ctx{lib_stream(), b.d_views.p, b.d_tiles.p, b.ntiles, b.tile_work.p, b.tile_order.p};
        SelfBatch* sb = b.self.get(b.members, b.h_views, b.d_views.p);
        if (plan.any_dyn && !b.tuner.poll(b.tuner.tuning())) return false;
        const bool tune = plan.any_dyn && b.tuner.tuning();
        if (tune && !b.tuner.begin_step(substeps)) return false;
        const bool graphed = graph_key(plan, substeps, iterations, fetch, b.tuner, &key);
        key.push_back(0);
        const ColorSegs bends{b.d_bend_segs.p, &b.seg_first, &b.seg_count};
        const ColorSegs stretch_colors{b.d_stretch_segs.p, &b.sseg_first, &b.sseg_count};
        std::vector<const PcgGroup*> groups;
        for (const PcgGroup& g : b.pcg_groups) groups.push_back(&g);
        auto record = [&] {
            return run_schedule(b.members, ctx, b.d_views.p, bends, stretch_colors, groups, sb, substeps, iterations, fetch, b.tuner, tune);
        };
        if (!(graphed ? b.graph.run(lib_stream(), key, record) : record())) return false;
        if (tune) b.tuner.end_step();
    }
    if (plan.any_dyn) ++b.tuner.step;
    for (Solver* s : b.members)
        if (!s->dyn.clear_motion(lib_stream())) return false;
    if (!fetch) return force_sync ? sync_stream() : true;
    if (!sync_stream()) return false;
    const float ms = static_cast<float>(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    for (Solver* s : b.members) {
        if (!fetch_counts(*s)) return false;
        s->diag.step_ms = s->diag.constraints_ms = ms;
        s->diag.resolved_contacts = s->diag.abi41_soft_contact_count + s->diag.abi41_exact_impulse_contact_count +
                                    s->diag.abi41_edge_edge_contact_count + s->diag.static_sdf_contact_count;
    }
    return true;
}

// Runs `launches` (even) back-to-back fused constraint passes of a batch --
// the pass of every iteration but a substep's last: stretch gather, pins,
// analytic and SDF contacts -- between two events on the library stream, and
// returns the mean duration.  Consecutive launches on one stream: the mean is
// the kernel time plus the dispatch gap, with no per-launch instrumentation.
// The members' positions, previous positions and velocities are copied aside
// before and back afterwards, so the measurement leaves the state as it was.
// The batch must have been stepped since its tables were built: the members'
// views (refilled here, with that step's substep count) describe a prepared
// step only then.
// With `parts`, a batch stepped as two parts runs each iteration as its two
// parts' launches on their two streams (as its steps do), and the mean is
// the time of one iteration over all members.
static bool time_batch_pass(Batch& b, int launches, float* avg_ms, bool parts = false) {
    if (b.stepped_substeps <= 0 || b.ntiles <= 0) return fail("time a batch's pass after stepping it once");
    if (!ensure_own_tables(b)) return false;
    for (size_t m = 0; m < b.members.size(); ++m) {
        b.members[m]->fill_view(b.stepped_substeps);
        b.h_views[m] = b.members[m]->view;
    }
    if (!b.staging.upload(b.d_views.p, b.h_views.data(), sizeof(ClothView) * b.h_views.size(), lib_stream())) return false;
    parts = parts && !b.parts.empty();
    if (parts) {
        for (auto& part : b.parts) {
            for (size_t m = 0; m < part->members.size(); ++m) part->h_views[m] = part->members[m]->view;
            if (!part->staging.upload(part->d_views.p, part->h_views.data(), sizeof(ClothView) * part->h_views.size(), lib_stream()))
                return false;
        }
    }
    launches = std::max(2, launches + (launches & 1));
    size_t total = 0;
    for (Solver* s : b.members) total += 4 * static_cast<size_t>(s->nv);
    Dev<float4> saved;
    if (!sync_stream() || !saved.alloc(std::max<size_t>(total, 1), "timing snapshot")) return false;
    auto copy_state = [&](bool save) {
        size_t at = 0;
        for (Solver* s : b.members) {
            for (float4* a : {s->pos0.p, s->pos1.p, s->prev.p, s->vel.p}) {
                const size_t n = static_cast<size_t>(s->nv);
                if (n == 0) continue;
                if (!hip_ok(hipMemcpyAsync(save ? saved.p + at : a, save ? a : saved.p + at, sizeof(float4) * n, hipMemcpyDeviceToDevice,
                                           lib_stream()), "timing snapshot"))
                    return false;
                at += n;
            }
        }
        return true;
    };
    if (!copy_state(true)) return false;
    bool pins = false, edges = false;
    for (Solver* s : b.members) {
        pins |= !s->pins.empty();
        edges |= s->cfg.edge_count > 0 && !s->cfg.stretch_optimization_enabled;
    }
    const unsigned ops = (edges ? kOpStretch : 0u) | (pins ? (kOpPin | kOpPinPolish) : 0u) | kOpAnalytic | kOpSdf;
    LaunchCtx ctx{lib_stream(), b.d_views.p, b.d_tiles.p, b.ntiles};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (!hip_ok(hipEventCreate(&e0), "timing event") || !hip_ok(hipEventCreate(&e1), "timing event")) return false;
    bool ok = hip_ok(hipEventRecord(e0, ctx.stream), "timing event");
    if (parts) {
        ok = ok && hip_ok(hipStreamWaitEvent(part_stream(), e0, 0), "timing event");
        for (int k = 0; ok && k < launches; ++k)
            for (size_t q = 0; q < b.parts.size(); ++q) {
                const Batch& p = *b.parts[q];
                launch_vertex_pass(LaunchCtx{q == 0 ? lib_stream() : part_stream(), p.d_views.p, p.d_tiles.p, p.ntiles}, k & 1, (k & 1) ^ 1,
                                   ops);
            }
        ok = ok && hip_ok(hipEventRecord(b.join, part_stream()), "timing event") && hip_ok(hipStreamWaitEvent(lib_stream(), b.join, 0), "timing event");
    } else {
        for (int k = 0; ok && k < launches; ++k) launch_vertex_pass(ctx, k & 1, (k & 1) ^ 1, ops);
    }
    ok = ok && hip_ok(hipEventRecord(e1, ctx.stream), "timing event") && hip_ok(hipEventSynchronize(e1), "timing event");
    float ms = 0.0f;
    ok = ok && hip_ok(hipEventElapsedTime(&ms, e0, e1), "timing event");
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    ok = ok && copy_state(false) && sync_stream();
    if (ok) *avg_ms = ms / static_cast<float>(launches);
    return ok;
}

// Adds the device memory allocated (or freed) during its scope to a solver.
struct ByteScope {
    Solver* s;
    long long before;
    explicit ByteScope(Solver* solver) : s(solver), before(device_bytes_in_use()) {}
    ~ByteScope() { if (s) s->dev_bytes += device_bytes_in_use() - before; }
};

}  // namespace ssbl

using namespace ssbl;

// =================================================================== ABI41
extern "C" {

SSBL_API void* ssbl_create_solver(const SsblXpbdConfig* config, const SsblXpbdMesh* mesh) {
    g_error.clear();
    if (!config || config->vertex_count < 0 || config->edge_count < 0 || config->triangle_count < 0 || !std::isfinite(config->dt) ||
        config->dt <= 0.0f || !mesh || !mesh->positions || !mesh->inv_mass) {
        fail("invalid ABI41 solver create request");
        return nullptr;
    }
    Solver* s = new Solver();
    bool created;
    {
        ByteScope bytes(s);
        created = s->create(config, mesh);
    }
    if (!created) {
        const std::string err = g_error;
        delete s;
        g_error = err;
        return nullptr;
    }
    return s;
}

SSBL_API int ssbl_destroy_solver(void* handle) {
    g_error.clear();
    if (handle) { sync_stream(); delete static_cast<Solver*>(handle); }
    return 1;
}

SSBL_API int ssbl_reset_solver(void* handle) {  // A41:8975-8990
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail("invalid solver handle") ? 1 : 0;
    if (!sync_stream()) return 0;
    std::vector<float4> p(s->nv), q(s->nv), z(s->nv, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    for (int i = 0; i < s->nv; ++i) {
        p[i] = make_float4(s->h_rest[i].x, s->h_rest[i].y, s->h_rest[i].z, s->h_weff[i]);
        q[i] = make_float4(s->h_rest[i].x, s->h_rest[i].y, s->h_rest[i].z, 0.0f);
    }
    if (!s->pos0.upload(p, "reset positions") || !s->pos1.upload(p, "reset positions") || !s->prev.upload(q, "reset previous positions") ||
        !s->vel.upload(z, "reset velocities"))
        return 0;
    s->diag = SsblXpbdDiagnostics{};
    s->diag.finite_flag = 1;
    return 1;
}

SSBL_API int ssbl_update_pin_targets(void* handle, const int* indices, const float* positions, const float* weights, int count) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail("invalid solver handle") ? 1 : 0;
    ByteScope bytes(s);
    return s->upload_pins(indices, positions, weights, count) ? 1 : 0;
}

SSBL_API int ssbl_update_runtime_colliders(void* handle, const SsblXpbdRuntimeColliders* in) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !in) return fail("invalid runtime collider update") ? 1 : 0;
    s->colliders = *in;
    return 1;
}

SSBL_API int ssbl_update_positions(void* handle, const float* positions, int max_floats) {  // A41:9007-9019
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !positions || max_floats < s->nv * 3) return fail("invalid position upload") ? 1 : 0;
    if (!sync_stream()) return 0;
    std::vector<float4> p(s->nv), q(s->nv);
    for (int i = 0; i < s->nv; ++i) {
        p[i] = make_float4(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], s->h_weff[i]);
        q[i] = make_float4(positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], 0.0f);
    }
    return (s->pos0.upload(p, "upload positions") && s->prev.upload(q, "upload previous positions")) ? 1 : 0;
}

SSBL_API int ssbl_update_static_triangles(void* handle, const float* triangles, int triangle_count) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail("invalid static triangle update") ? 1 : 0;
    triangle_count = std::max(triangle_count, 0);
    if (triangle_count > 0 && !triangles) return fail("static triangle data is required when triangle count is nonzero") ? 1 : 0;
    ByteScope bytes(s);
    return s->build_sdf(triangles, triangle_count) ? 1 : 0;
}

SSBL_API int ssbl_update_dynamic_triangles(void* handle, const float* triangles, int triangle_count) {  // A41:8338-8420
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail("invalid dynamic triangle update") ? 1 : 0;
    triangle_count = std::max(triangle_count, 0);
    if (triangle_count > 0 && !triangles) return fail("dynamic triangle data is required when triangle count is nonzero") ? 1 : 0;
    ByteScope bytes(s);
    return s->dyn.set_triangles(lib_stream(), triangles, triangle_count, nullptr, nullptr, s->cfg.collision_margin,
                                s->cfg.cloth_thickness, false) ? 1 : 0;
}

SSBL_API int ssbl_update_frame_inputs(void* handle, const SsblXpbdFrameInputs* in) {  // A41:9035-9085
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !in) return fail("invalid frame input update") ? 1 : 0;
    ByteScope bytes(s);
    if (in->update_pin_targets && !s->upload_pins(in->pin_indices, in->pin_positions, in->pin_weights, in->pin_count)) return 0;
    if (in->update_runtime_colliders) s->colliders = in->runtime_colliders;
    if (in->update_dynamic_triangles) {
        const int c = std::max(in->dynamic_triangle_count, 0);
        if (c > 0 && !in->dynamic_triangles) return fail("dynamic triangle data is required when triangle count is nonzero") ? 1 : 0;
        if (!s->dyn.set_triangles(lib_stream(), in->dynamic_triangles, c, nullptr, nullptr, s->cfg.collision_margin,
                                  s->cfg.cloth_thickness, false))
            return 0;
    }
    if (in->update_dynamic_indexed_triangles) {   // A41:8433-8590
        const int nv = std::max(in->dynamic_triangle_vertex_count, 0), nt = std::max(in->dynamic_indexed_triangle_count, 0);
        if (nv > 0 && nt > 0 && (!in->dynamic_triangle_vertices || !in->dynamic_triangle_indices))
            return fail("dynamic indexed triangle positions and indices are required when count is nonzero") ? 1 : 0;
        if (!s->dyn.set_indexed(lib_stream(), in->dynamic_triangle_vertices, nv, in->dynamic_triangle_indices, nt,
                                in->dynamic_triangle_topology_changed != 0, s->cfg.collision_margin, s->cfg.cloth_thickness))
            return 0;
    }
    if (in->update_dynamic_particles) {   // A41:8575-8697
        const int c = std::max(in->dynamic_particle_count, 0);
        if (c > 0 && (!in->dynamic_particle_positions || !in->dynamic_particle_radii))
            return fail("dynamic particle position/radius data is required when particle count is nonzero") ? 1 : 0;
        if (!s->dyn.set_particles(lib_stream(), in->dynamic_particle_positions, nullptr, in->dynamic_particle_radii, nullptr, nullptr, c,
                                  s->cfg.collision_margin, s->cfg.cloth_thickness))
            return 0;
    }
    if (in->update_static_triangles) {
        const int c = std::max(in->static_triangle_count, 0);
        if (c > 0 && !in->static_triangles) return fail("static triangle data is required when triangle count is nonzero") ? 1 : 0;
        if (!s->build_sdf(in->static_triangles, c)) return 0;
    }
    if (in->update_force_fields && !s->upload_fields(in->force_fields, in->force_field_count, in->unsupported_force_field_count)) return 0;
    s->diag.force_field_count = static_cast<long long>(s->fields.size());
    s->diag.unsupported_force_field_count = s->unsupported_fields;
    return 1;
}

SSBL_API int ssbl_step_solver_ex(void* handle, int substeps, int iterations, int fetch_diagnostics, int force_sync) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail("invalid solver handle") ? 1 : 0;
    return step_one(*s, substeps, iterations, fetch_diagnostics, force_sync) ? 1 : 0;
}

// The allocation-free variant of ssbl_step_solver_ex: the hard-stretch solve
// takes its scratch from the library's static array instead of the handle's
// own arrays, and the step launches directly; nothing is allocated or freed.
SSBL_API int ssbl_ext_step_solver_ex_noalloc(void* handle, int substeps, int iterations, int fetch_diagnostics, int force_sync) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail("invalid solver handle") ? 1 : 0;
    if (s->pcg.enabled()) {
        void* base = nullptr;
        size_t bytes = 0;
        if (!pcg_static_scratch(&base, &bytes) || !s->pcg.use_scratch(base, bytes)) return 0;
    }
    const bool ok = step_one(*s, substeps, iterations, fetch_diagnostics, force_sync, true);
    s->pcg.use_own();
    return ok ? 1 : 0;
}

SSBL_API int ssbl_step_solver(void* handle, int substeps, int iterations) {
    return ssbl_step_solver_ex(handle, substeps, iterations, 1, 1);
}

static bool download_float4(const Dev<float4>& src, int n, float* out) {
    std::vector<float4> h(n);
    if (n > 0 && !hip_ok(hipMemcpy(h.data(), src.p, sizeof(float4) * n, hipMemcpyDeviceToHost), "download ABI41 state")) return false;
    for (int i = 0; i < n; ++i) { out[3 * i] = h[i].x; out[3 * i + 1] = h[i].y; out[3 * i + 2] = h[i].z; }
    return true;
}

SSBL_API int ssbl_download_positions(void* handle, float* out, int max_floats) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !out || max_floats < s->nv * 3) return fail("invalid position download") ? 1 : 0;
    if (!sync_stream()) return 0;
    return download_float4(s->pos0, s->nv, out) ? 1 : 0;
}

SSBL_API int ssbl_get_diagnostics(void* handle, SsblXpbdDiagnostics* out) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !out) return fail("invalid diagnostics request") ? 1 : 0;
    *out = s->diag;
    return 1;
}

SSBL_API unsigned int ssbl_capabilities(void) {
    return SSBL_CAP_STRETCH_OPTIMIZATION | SSBL_CAP_PIN_WEIGHTS | SSBL_CAP_GLOBAL_DYNAMIC_SCENE | SSBL_CAP_HIP_BATCHED_STEP;
}

SSBL_API void* ssbl_create_global_dynamic_scene(void) {  // A41:8733-8736
    g_error.clear();
    GlobalScene* sc = new GlobalScene();
    if (!sc->counts.alloc(kCountSlots, "global dynamic scene counts allocation")) {
        delete sc;
        return nullptr;
    }
    return sc;
}

SSBL_API int ssbl_destroy_global_dynamic_scene(void* scene) {
    g_error.clear();
    if (!scene) return fail("invalid global dynamic scene handle") ? 1 : 0;
    sync_stream();
    delete static_cast<GlobalScene*>(scene);
    return 1;
}

SSBL_API int ssbl_update_global_dynamic_scene(void* scene, const SsblXpbdGlobalDynamicSceneInputs* in) {  // A41:6597-6782
    g_error.clear();
    GlobalScene* sc = static_cast<GlobalScene*>(scene);
    if (!sc || !in) return fail("invalid global dynamic scene update") ? 1 : 0;
    const int np = std::max(in->particle_count, 0), nt = std::max(in->triangle_count, 0);
    if (np > 0 && (!in->particle_positions || !in->particle_radii || !in->particle_source_ids || !in->particle_collision_layers))
        return fail("global dynamic particle arrays are incomplete") ? 1 : 0;
    if (nt > 0 && (!in->triangle_vertices || !in->triangle_source_ids || !in->triangle_collision_layers))
        return fail("global dynamic triangle arrays are incomplete") ? 1 : 0;
    sc->margin = std::max(in->collision_margin, 0.0f);
    sc->thickness = std::max(in->cloth_thickness, 0.0f);
    hipStream_t st = lib_stream();
    launch_fill_u32(st, reinterpret_cast<unsigned*>(sc->counts.p), 0u, 2 * static_cast<size_t>(kCountSlots));
    const float* prev = in->particle_prev_positions ? in->particle_prev_positions : in->particle_positions;
    if (!sc->set.set_particles(st, in->particle_positions, prev, in->particle_radii, in->particle_source_ids, in->particle_collision_layers,
                               np, sc->margin, sc->thickness) ||
        !sc->set.set_triangles(st, in->triangle_vertices, nt, in->triangle_source_ids, in->triangle_collision_layers, sc->margin,
                               sc->thickness, true) ||
        !sc->set.build(st, sc->counts.p, true, false))
        return 0;
    sc->overflow_pending = true;
    if (!sync_stream()) return 0;
    settle_scene_overflow(*sc);
    return 1;
}

SSBL_API int ssbl_attach_global_dynamic_scene(void* handle, void* scene, int target_source_id, int target_collision_layer,
                                              int collision_mode) {  // A41:8758-8776
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !scene) return fail("invalid global dynamic scene attach") ? 1 : 0;
    s->scene = static_cast<GlobalScene*>(scene);
    s->target_src = target_source_id;
    s->target_layer = target_collision_layer;
    s->cross_mode = collision_mode;
    return 1;
}

SSBL_API int ssbl_clear_global_dynamic_scene(void* handle) {  // A41:8778-8789
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail("invalid global dynamic scene clear") ? 1 : 0;
    s->scene = nullptr;
    s->target_src = 0;
    s->target_layer = 0;
    s->cross_mode = 0;
    return 1;
}

SSBL_API const char* ssbl_last_error(void) { return g_error.c_str(); }

// ============================================================== extensions
SSBL_API int ssbl_ext_set_stretch_mode(void* handle, int mode) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s || (mode != SSBL_EXT_STRETCH_JACOBI && mode != SSBL_EXT_STRETCH_COLORED)) return fail("invalid stretch mode") ? 1 : 0;
    ByteScope bytes(s);
    return s->set_stretch_mode(mode) ? 1 : 0;
}

SSBL_API int ssbl_ext_download_state(void* handle, float* p, float* q, float* v, int max_floats) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s || max_floats < s->nv * 3) return fail("invalid state download") ? 1 : 0;
    if (!sync_stream()) return 0;
    if (p && !download_float4(s->pos0, s->nv, p)) return 0;
    if (q && !download_float4(s->prev, s->nv, q)) return 0;
    if (v && !download_float4(s->vel, s->nv, v)) return 0;
    return 1;
}

SSBL_API int ssbl_ext_self_hash_cells(void* handle, int* cells, int max_ints) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !cells || max_ints < 3 * s->nv) return fail("invalid hash cell request") ? 1 : 0;
    if (!sync_stream()) return 0;
    return s->self.download_cells(cells, s->nv) ? 1 : 0;
}

SSBL_API int ssbl_ext_self_neighbors(void* handle, unsigned* counts, unsigned* indices, int max_vertices) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !counts || !indices || max_vertices < s->nv) return fail("invalid neighbour request") ? 1 : 0;
    if (!sync_stream()) return 0;
    return s->self.download_neighbors(counts, indices, s->nv) ? 1 : 0;
}

SSBL_API int ssbl_ext_static_sdf(void* handle, int* dims, float* ov, float* values, int max_values) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s || !dims || !ov) return fail("invalid SDF request") ? 1 : 0;
    if (!s->sdf_ready) return fail("no static SDF") ? 1 : 0;
    for (int k = 0; k < 3; ++k) dims[k] = s->sdf_dim[k];
    ov[0] = s->sdf_origin[0]; ov[1] = s->sdf_origin[1]; ov[2] = s->sdf_origin[2]; ov[3] = s->sdf_voxel;
    if (values) {
        const size_t n = static_cast<size_t>(s->sdf_dim[0]) * s->sdf_dim[1] * s->sdf_dim[2];
        if (max_values < static_cast<long long>(n)) return fail("SDF buffer too small") ? 1 : 0;
        if (!sync_stream()) return 0;
        if (!hip_ok(hipMemcpy(values, s->sdf.p, sizeof(float) * n, hipMemcpyDeviceToHost), "download SDF")) return 0;
    }
    return 1;
}

SSBL_API void* ssbl_ext_create_batch(void* const* handles, int count) {
    g_error.clear();
    if (!handles || count <= 0) { fail("invalid batch"); return nullptr; }
    Batch* b = new Batch();
    for (int k = 0; k < count; ++k) {
        if (!handles[k]) { delete b; fail("invalid batch member"); return nullptr; }
        b->members.push_back(static_cast<Solver*>(handles[k]));
    }
    if (!build_batch(*b)) { const std::string e = g_error; delete b; g_error = e; return nullptr; }
    return b;
}
SSBL_API int ssbl_ext_batch_set_members(void* batch, void* const* handles, int count) {
    g_error.clear();
    if (!batch || !handles || count <= 0) return fail("invalid batch member update") ? 1 : 0;
    std::vector<Solver*> members;
    for (int k = 0; k < count; ++k) {
        if (!handles[k]) return fail("invalid batch member") ? 1 : 0;
        members.push_back(static_cast<Solver*>(handles[k]));
    }
    return retarget_batch(*static_cast<Batch*>(batch), members) ? 1 : 0;
}
SSBL_API int ssbl_ext_destroy_batch(void* batch) {
    g_error.clear();
    if (batch) { sync_stream(); delete static_cast<Batch*>(batch); }
    return 1;
}
SSBL_API int ssbl_ext_step_batch(void* batch, int substeps, int iterations, int force_sync) {
    g_error.clear();
    if (!batch) return fail("invalid batch") ? 1 : 0;
    return step_batch(*static_cast<Batch*>(batch), substeps, iterations, force_sync) ? 1 : 0;
}
SSBL_API int ssbl_ext_step_batch_ex(void* batch, int substeps, int iterations, int fetch_diagnostics, int force_sync) {
    g_error.clear();
    if (!batch) return fail("invalid batch") ? 1 : 0;
    return step_batch(*static_cast<Batch*>(batch), substeps, iterations, force_sync, fetch_diagnostics != 0) ? 1 : 0;
}
SSBL_API int ssbl_ext_update_global_scene_from_handles(void* scene, void* const* handles, const int* source_ids,
                                                      const int* collision_layers, const float* particle_radii, int count,
                                                      float collision_margin, float cloth_thickness) {
    g_error.clear();
    GlobalScene* sc = static_cast<GlobalScene*>(scene);
    if (!sc || count < 0 || (count > 0 && (!handles || !source_ids || !collision_layers || !particle_radii)))
        return fail("invalid global scene gather") ? 1 : 0;
    std::vector<GatherSource> src;
    for (int k = 0; k < count; ++k) {
        Solver* s = static_cast<Solver*>(handles[k]);
        if (!s) return fail("invalid solver handle in global scene gather") ? 1 : 0;
        GatherSource g{};
        g.pos = s->pos0.p;
        g.tris = s->tris.p;
        g.order = s->tri_order.p;
        g.nv = s->nv;
        g.nt = s->cfg.triangle_count;
        g.radius = particle_radii[k];
        g.source_id = source_ids[k];
        g.layer = collision_layers[k];
        src.push_back(g);
    }
    sc->margin = std::max(collision_margin, 0.0f);
    sc->thickness = std::max(cloth_thickness, 0.0f);
    hipStream_t st = lib_stream();
    if (!sc->set.gather(st, src, sc->margin, sc->thickness)) return 0;
    launch_fill_u32(st, reinterpret_cast<unsigned*>(sc->counts.p), 0u, 2 * static_cast<size_t>(kCountSlots));
    if (!sc->set.build(st, sc->counts.p, true, false)) return 0;
    sc->overflow_pending = true;
    return 1;
}

SSBL_API long long ssbl_ext_global_scene_entries(void* scene) {
    g_error.clear();
    if (!scene) { fail("invalid global dynamic scene handle"); return -1; }
    if (!sync_stream()) return -1;
    return static_cast<GlobalScene*>(scene)->set.hash_entries();
}

SSBL_API int ssbl_ext_set_static_sdf(void* handle, const int* dims, const float* origin_and_voxel, const float* values, int count) {
    g_error.clear();
    Solver* s = static_cast<Solver*>(handle);
    if (!s) return fail("invalid solver handle") ? 1 : 0;
    ByteScope bytes(s);
    return s->set_sdf_grid(dims, origin_and_voxel, values, count) ? 1 : 0;
}

SSBL_API long long ssbl_ext_device_bytes(void* handle) {
    g_error.clear();
    if (!handle) { fail("invalid solver handle"); return -1; }
    return static_cast<Solver*>(handle)->dev_bytes;
}
SSBL_API int ssbl_ext_time_batch_pass(void* batch, int launches, float* avg_ms) {
    g_error.clear();
    if (!batch || !avg_ms) return fail("invalid batch timing request") ? 1 : 0;
    return time_batch_pass(*static_cast<Batch*>(batch), launches, avg_ms) ? 1 : 0;
}
SSBL_API int ssbl_ext_time_batch_iteration(void* batch, int iterations, float* avg_ms) {
    g_error.clear();
    if (!batch || !avg_ms) return fail("invalid batch timing request") ? 1 : 0;
    return time_batch_pass(*static_cast<Batch*>(batch), iterations, avg_ms, true) ? 1 : 0;
}
SSBL_API int ssbl_ext_synchronize(void) { g_error.clear(); return sync_stream() ? 1 : 0; }
SSBL_API int ssbl_ext_set_graph_mode(int mode) {
    g_error.clear();
    return g_graph_mode.exchange(mode != 0 ? 1 : 0);
}
SSBL_API int ssbl_ext_set_batch_parts(int parts) {
    g_error.clear();
    return g_batch_parts.exchange(parts >= 2 ? 2 : 1);
}
SSBL_API int ssbl_ext_set_autotune(int mode) {
    g_error.clear();
    return g_autotune.exchange(mode != 0 ? 1 : 0);
}
SSBL_API int ssbl_ext_tuned_pool_limit(void* handle, void* batch) {
    g_error.clear();
    if ((handle == nullptr) == (batch == nullptr)) {
        fail("invalid tuning request");
        return -2;
    }
    const PoolTuner& t = handle ? static_cast<Solver*>(handle)->tuner : static_cast<Batch*>(batch)->tuner;
    return t.chosen;
}
SSBL_API int ssbl_ext_graph_stats(void* handle, void* batch, long long* captures, long long* replays) {
    g_error.clear();
    if ((handle == nullptr) == (batch == nullptr) || !captures || !replays) return fail("invalid graph stats request") ? 1 : 0;
    const StepGraph& g = handle ? static_cast<Solver*>(handle)->graph : static_cast<Batch*>(batch)->graph;
    *captures = g.captures;
    *replays = g.replays;
    return 1;
}
SSBL_API int ssbl_ext_kernel_timing(int enable) {
    g_error.clear();
    if (!enable) {
        sync_stream();
        std::lock_guard<std::mutex> lock(g_timing_mutex);
        resolve_pending();
    }
    g_timing.store(enable != 0);
    return 1;
}
SSBL_API int ssbl_ext_kernel_stats(char* out, int cap) {
    g_error.clear();
    if (!out || cap <= 2) return fail("invalid stats buffer") ? 1 : 0;
    sync_stream();
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    resolve_pending();
    std::string js = "{";
    bool first = true;
    for (auto& kv : g_kstats) {
        char buf[256];
        std::snprintf(buf, sizeof(buf), "%s\"%s\": {\"launches\": %lld, \"total_ms\": %.6f}", first ? "" : ", ", kv.first.c_str(),
                      kv.second.launches, kv.second.total_ms);
        js += buf;
        first = false;
    }
    js += "}";
    g_kstats.clear();
    if (static_cast<int>(js.size()) + 1 > cap) return fail("stats buffer too small") ? 1 : 0;
    std::memcpy(out, js.c_str(), js.size() + 1);
    return 1;
}
SSBL_API const char* ssbl_ext_backend_name(void) {
    static std::string name = [] {
        hipDeviceProp_t prop{};
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return std::string("hip (no device)");
        return std::string("hip ") + prop.gcnArchName + " " + prop.name;
    }();
    return name.c_str();
}

}  // extern "C"
