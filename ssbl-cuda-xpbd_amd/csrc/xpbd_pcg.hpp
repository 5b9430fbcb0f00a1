// xpbd_pcg.hpp -- hard-stretch PCG, polish and caps: the arrays of one cloth
// (PcgPath) and the launches over a group of cloths (PcgGroup).
#pragma once

#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "xpbd_device.hpp"

namespace ssbl {

bool fail(const std::string& message);
bool hip_ok(hipError_t err, const char* what);

// Size of the library's static PCG scratch array (per GPU, held until the
// process ends): the iteration vectors, block diagonals, off-diagonal blocks,
// reduction partials and scalars of one hard-stretch solve, for cloths up to
// about 330k grid vertices.
constexpr size_t kPcgScratchBytes = size_t(96) << 20;

// Device address and size of the static scratch array on the current device
// (looked up once per device).
bool pcg_static_scratch(void** base, size_t* bytes);

// One cloth's PCG arrays and cap colour segments.
class PcgPath {
public:
    PcgPath() = default;
    PcgPath(const PcgPath&) = delete;
    PcgPath& operator=(const PcgPath&) = delete;
    ~PcgPath();
    bool prepare(const SsblXpbdConfig& cfg, const std::vector<int>& edges, const std::vector<float>& edge_rest,
                 const std::vector<int>& colors, int nv, int nadj, int ntiles, int ell_width, int ell_pitch, int ell_overflow,
                 unsigned long long* counts);
    bool enabled() const { return ready_; }
    // polish and caps run for this cloth (strength >= 0.95, A41:7823-7857)
    bool polishes() const { return ready_ && pv_.strength_cfg >= 0.95f && pv_.ne > 0; }
    // the caps run per colour (a valid edge colouring) or as gathers
    bool colored() const { return colored_; }
    const PcgView& view() const { return pv_; }
    const std::vector<int4>& cap_segments() const { return segs_host_; }
    const std::vector<int>& cap_first() const { return seg_first_; }
    const std::vector<int>& cap_count() const { return seg_count_; }
    bool fetch(SsblXpbdDiagnostics* d) const;
    // Bytes of scratch one solve of this cloth needs.
    size_t scratch_bytes() const;
    // Points the solve's scratch (vectors, blocks, partials, scalars) into
    // [base, base + bytes) instead of the handle's own arrays; fails when it
    // does not fit.  use_own() switches back.
    bool use_scratch(void* base, size_t bytes);
    void use_own() { pv_ = own_; }

private:
    void* alloc_bytes(size_t bytes);
    bool ready_ = false;
    bool colored_ = false;
    int nnz_ = 0, ntiles_ = 0;
    size_t off_blocks_ = 0;   // off-diagonal blocks stored (slot-major slots, or the CSR incidences)
    PcgView pv_{};    // the view the launches use
    PcgView own_{};   // the view over the handle's own arrays
    std::vector<int4> segs_host_;
    std::vector<int> seg_first_, seg_count_;
    std::vector<std::pair<void*, size_t>> allocations_;
};

// The PCG launches over a group of cloths that share a substep cadence: every
// phase of the solve is one launch over all of the group's tiles, each
// reduction one launch with a workgroup per cloth, and the polish and caps
// one launch per gather or colour over the cloths that polish.  The launch
// count does not depend on the number of cloths.
struct PcgMember {
    int index;            // the cloth's entry in the step's view table
    const PcgPath* path;
    int nv;
};

class PcgGroup {
public:
    PcgGroup() = default;
    PcgGroup(const PcgGroup&) = delete;
    PcgGroup& operator=(const PcgGroup&) = delete;
    PcgGroup(PcgGroup&&) = default;
    PcgGroup& operator=(PcgGroup&&) = default;
    bool build(const std::vector<PcgMember>& members, int cadence);
    int cadence() const { return cadence_; }
    // The solve runs on every cadence-th substep and always on the last (A41:9158-9177).
    bool due(int substep, int substeps) const {
        const int c = std::max(1, std::min(cadence_, substeps));
        return ((substep + 1) % c) == 0 || substep == substeps - 1;
    }
    // Clears the scalars and the guard / max-delta counters of every member.
    void reset(hipStream_t st, const ClothView* views) const;
    void solve(hipStream_t st, const ClothView* views, int parity) const;
    // The polish sequence, then the extreme caps (A41:7823-7857, 7974-7995);
    // the final caps (A41:7997-8027).  Each returns with every member at `parity`.
    void polish(hipStream_t st, const ClothView* views, int parity) const;
    void final_caps(hipStream_t st, const ClothView* views, int parity) const;

private:
    template <class T> struct Buf {
        T* p = nullptr;
        size_t n = 0;
        Buf() = default;
        Buf(const Buf&) = delete;
        Buf& operator=(const Buf&) = delete;
        Buf(Buf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
        Buf& operator=(Buf&& o) noexcept { std::swap(p, o.p); std::swap(n, o.n); return *this; }
        ~Buf() { device_free(p, sizeof(T) * n); }
        bool upload(const std::vector<T>& h, const char* what);   // grow-only
    };
    // the tiles of the cloths that polish, coloured or not
    struct Polishing {
        Buf<Tile> tiles;
        int ntiles = 0;
        bool colored = false;
    };
    int gather(hipStream_t st, const ClothView* views, const Polishing& set, int parity, int kind) const;
    int caps(hipStream_t st, const ClothView* views, const Polishing& set, int parity, int kind, int passes, bool reverse_too) const;
    int cadence_ = 1;
    int ntiles_ = 0, nmembers_ = 0;
    Buf<Tile> tiles_;
    Polishing colored_, plain_;
    Buf<int> members_;
    Buf<int4> cap_segs_;
    std::vector<int> cap_first_, cap_count_;
};

}  // namespace ssbl
