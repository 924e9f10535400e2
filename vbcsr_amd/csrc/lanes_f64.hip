// Lane-per-row SpMM kernels for double storage (see lanes.hpp).
#include "lanes.hpp"

namespace vbh {
VBH_LANES_INSTANTIATE(double, double)
}  // namespace vbh
