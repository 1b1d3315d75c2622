// Instantiations of the fused advection kernel for 2D fields (split by
// dimension so the two halves compile in parallel).
#include "advect_impl.hpp"

namespace pyftle {
void launch_advect_2d(const AdvectArgs &a, const LaunchSpec &sp) { launch_advect<2>(a, sp); }
}  // namespace pyftle
