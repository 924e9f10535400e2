// Lane-per-row SpMM kernels for c64 storage (see lanes.hpp).
#include "lanes.hpp"

namespace vbh {
VBH_LANES_INSTANTIATE(c64, c64)
}  // namespace vbh
