// xpbd_self.hip -- self-collision kernels (spatial hashes, candidate pairs,
// vertex-vertex / vertex-triangle / edge-edge responses).
#include "xpbd_self.hpp"
