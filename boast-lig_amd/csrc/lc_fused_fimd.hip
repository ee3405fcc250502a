// lc_fused_fimd.hip -- fused 3D kernel variants for the FIMD filter.
#include "lc_fused_impl.hpp"

namespace lc {
LC_FUSED_INSTANTIATE(double, FIMD)
LC_FUSED_INSTANTIATE(float, FIMD)
}  // namespace lc
