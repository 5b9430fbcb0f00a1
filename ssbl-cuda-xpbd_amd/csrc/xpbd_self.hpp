// xpbd_self.hpp -- self-collision state of one cloth and its per-step driver.
#pragma once

#include <string>
#include <vector>

#include "xpbd_device.hpp"

namespace ssbl {

bool fail(const std::string& message);

struct Solver;

// Runs the self-collision passes of one step for a set of cloths.
class SelfBatch {
public:
    virtual ~SelfBatch() = default;
    virtual bool run(const LaunchCtx& ctx, int parity, int substep, int substeps, bool count) = 0;
};

class SelfCollision {
public:
    bool prepare(const SsblXpbdConfig&, const std::vector<int>&, const std::vector<int>&, int) {
        return fail("self collision is not available in this build yet");
    }
    SelfBatch* single_batch(const ClothView&, const ClothView*) { return nullptr; }
    bool download_cells(int*, int) { return fail("no self-collision hash"); }
    bool download_neighbors(unsigned*, unsigned*, int) { return fail("no self-collision table"); }
};

class SelfBatchHolder {
public:
    SelfBatch* get(const std::vector<Solver*>&, const std::vector<ClothView>&, const ClothView*) { return nullptr; }
};

}  // namespace ssbl
