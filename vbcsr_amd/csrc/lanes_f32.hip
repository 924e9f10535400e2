// Lane-per-row SpMM kernels for float storage (see lanes.hpp).
#include "lanes.hpp"

namespace vbh {
VBH_LANES_INSTANTIATE(float, float)
}  // namespace vbh
