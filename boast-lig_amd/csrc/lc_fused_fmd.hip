// lc_fused_fmd.hip -- fused 3D kernel variants for the FMD filter.
#include "lc_fused_impl.hpp"

namespace lc {
LC_FUSED_INSTANTIATE(double, FMD)
LC_FUSED_INSTANTIATE(float, FMD)
}  // namespace lc
