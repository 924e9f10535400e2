// Lane-per-row SpMM kernels for c128 vectors against c64 matrix values (see lanes.hpp).
#include "lanes.hpp"

namespace vbh {
VBH_LANES_INSTANTIATE(c128, c64)
}  // namespace vbh
