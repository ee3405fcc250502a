// lc_fused_fd2.hip -- fused 3D kernel variants for the FD2 filter.
#include "lc_fused_impl.hpp"

namespace lc {
LC_FUSED_INSTANTIATE(double, FD2)
LC_FUSED_INSTANTIATE(float, FD2)
}  // namespace lc
