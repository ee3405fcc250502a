// lc_fused_fd1.hip -- fused 3D kernel variants for the FD1 filter.
#include "lc_fused_impl.hpp"

namespace lc {
LC_FUSED_INSTANTIATE(double, FD1)
LC_FUSED_INSTANTIATE(float, FD1)
}  // namespace lc
