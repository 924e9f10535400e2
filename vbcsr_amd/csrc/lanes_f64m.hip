// Lane-per-row SpMM kernels for double vectors against float matrix values (see lanes.hpp).
#include "lanes.hpp"

namespace vbh {
VBH_LANES_INSTANTIATE(double, float)
}  // namespace vbh
