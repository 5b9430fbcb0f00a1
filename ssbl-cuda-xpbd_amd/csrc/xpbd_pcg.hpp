// xpbd_pcg.hpp -- hard-stretch PCG, polish and caps of one cloth.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "xpbd_device.hpp"

namespace ssbl {

bool fail(const std::string& message);
bool hip_ok(hipError_t err, const char* what);

struct PcgScalars {
    double rz, dad, rznext, initial;
    float beta;
    int status, dir_ok, runs;
};

struct PcgView {
    int nv, ne, hard, pad;
    float strength_cfg;   // configured strength (polish / cap thresholds)
    float strength;       // clamp01(strength * 2.5), the system stiffness
    float thick, wb_limit;
    float4 *rhs, *sol, *res, *z, *dir, *adir;
    float *diag, *pinv, *off;   // 6 floats per block
    unsigned long long* partials;
    PcgScalars* sc;
    unsigned long long* guard;
    float* max_delta;
    const int2* edges;
    const float* edge_rest;
    unsigned long long* counts;
};

// Size of the library's static PCG scratch array (per GPU, held until the
// process ends): the iteration vectors, block diagonals, off-diagonal blocks,
// reduction partials and scalars of one hard-stretch solve, for cloths up to
// about 330k grid vertices.
constexpr size_t kPcgScratchBytes = size_t(96) << 20;

// Device address and size of the static scratch array on the current device
// (looked up once per device).
bool pcg_static_scratch(void** base, size_t* bytes);

class PcgPath {
public:
    PcgPath() = default;
    PcgPath(const PcgPath&) = delete;
    PcgPath& operator=(const PcgPath&) = delete;
    ~PcgPath();
    bool prepare(const SsblXpbdConfig& cfg, const std::vector<int>& edges, const std::vector<float>& edge_rest,
                 const std::vector<int>& colors, int nv, int nadj, int ntiles, unsigned long long* counts);
    bool enabled() const { return ready_; }
    bool reset_diagnostics(hipStream_t st);
    void solve(hipStream_t st, const ClothView* cv, const Tile* tiles, int parity);
    int polish(hipStream_t st, const ClothView* cv, const Tile* tiles, int parity, float dt);
    int final_caps(hipStream_t st, const ClothView* cv, const Tile* tiles, int parity, float dt);
    int copy(hipStream_t st, const ClothView* cv, const Tile* tiles, int parity) { return gather(st, cv, tiles, parity, -1, 0.0f); }
    bool fetch(SsblXpbdDiagnostics* d) const;
    // Bytes of scratch one solve of this cloth needs.
    size_t scratch_bytes() const;
    // Points the solve's scratch (vectors, blocks, partials, scalars) into
    // [base, base + bytes) instead of the handle's own arrays; fails when it
    // does not fit.  use_own() switches back.
    bool use_scratch(void* base, size_t bytes);
    void use_own() { pv_ = own_; }

private:
    int gather(hipStream_t st, const ClothView* cv, const Tile* tiles, int parity, int kind, float dt);
    int caps(hipStream_t st, const ClothView* cv, const Tile* tiles, int parity, int kind, int passes, bool reverse_too, float dt);
    void* alloc_bytes(size_t bytes);
    bool ready_ = false;
    bool colored_ = false;
    int nnz_ = 0, ntiles_ = 0;
    PcgView pv_{};    // the view the launches use
    PcgView own_{};   // the view over the handle's own arrays
    int4* segs_ = nullptr;
    std::vector<int> seg_first_, seg_count_;
    std::vector<std::pair<void*, size_t>> allocations_;
};

}  // namespace ssbl
