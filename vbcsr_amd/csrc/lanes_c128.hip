// Lane-per-row SpMM kernels for c128 storage (see lanes.hpp).
#include "lanes.hpp"

namespace vbh {
VBH_LANES_INSTANTIATE(c128, c128)
}  // namespace vbh
