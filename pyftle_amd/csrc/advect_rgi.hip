// Instantiations of the general advection kernel for the RegularGridInterpolator
// modes (GridInterpolator, interpolate.py:546-762) and the scattered-data modes
// (interpolate.py:188-294): FP64 positions and blend, FP32 or FP64 records,
// STEADY or FROZEN time, SINGLE layout.
#include "advect_impl.hpp"

namespace pyftle {

template <int DIM, int METHOD, int TMODE, int INTERP, bool STORE64>
static void rgi_leaf(const AdvectArgs &a, const LaunchSpec &sp) {
  constexpr int PREC = STORE64 ? PREC_NATIVE : PREC_F64C;
  if (sp.ftle)
    advect_kernel<DIM, METHOD, TMODE, false, STORE64, PREC, true, INTERP><<<sp.blocks, sp.threads, 0, sp.stream>>>(a);
  else
    advect_kernel<DIM, METHOD, TMODE, false, STORE64, PREC, false, INTERP><<<sp.blocks, sp.threads, 0, sp.stream>>>(a);
}

template <int DIM, int METHOD, int TMODE, int INTERP>
static void rgi_store(const AdvectArgs &a, const LaunchSpec &sp) {
  if (sp.store64) rgi_leaf<DIM, METHOD, TMODE, INTERP, true>(a, sp);
  else rgi_leaf<DIM, METHOD, TMODE, INTERP, false>(a, sp);
}

template <int DIM, int METHOD, int INTERP>
static void rgi_time(const AdvectArgs &a, const LaunchSpec &sp) {
  if (sp.tmode == PYFTLE_TIME_FROZEN) rgi_store<DIM, METHOD, PYFTLE_TIME_FROZEN, INTERP>(a, sp);
  else rgi_store<DIM, METHOD, PYFTLE_TIME_STEADY, INTERP>(a, sp);
}

template <int DIM, int INTERP>
static void rgi_method(const AdvectArgs &a, const LaunchSpec &sp) {
  if (sp.method == PYFTLE_RK4) rgi_time<DIM, PYFTLE_RK4, INTERP>(a, sp);
  else if (sp.method == PYFTLE_EULER) rgi_time<DIM, PYFTLE_EULER, INTERP>(a, sp);
  else rgi_time<DIM, PYFTLE_AB2, INTERP>(a, sp);
}

template <int DIM>
static void rgi_interp(int interp, const AdvectArgs &a, const LaunchSpec &sp) {
  if (interp == PYFTLE_INTERP_RGI_LINEAR) rgi_method<DIM, PYFTLE_INTERP_RGI_LINEAR>(a, sp);
  else if (interp == PYFTLE_INTERP_RGI_NEAREST) rgi_method<DIM, PYFTLE_INTERP_RGI_NEAREST>(a, sp);
  else if (interp == PYFTLE_INTERP_RGI_CUBIC) rgi_method<DIM, PYFTLE_INTERP_RGI_CUBIC>(a, sp);
  else if (interp == PYFTLE_INTERP_SCATTER_LINEAR) rgi_method<DIM, PYFTLE_INTERP_SCATTER_LINEAR>(a, sp);
  else if (interp == PYFTLE_INTERP_SCATTER_NEAREST) rgi_method<DIM, PYFTLE_INTERP_SCATTER_NEAREST>(a, sp);
  else if constexpr (DIM == 2) rgi_method<DIM, PYFTLE_INTERP_SCATTER_CUBIC>(a, sp);
}

void launch_advect_rgi(int dim, int interp, const AdvectArgs &a, const LaunchSpec &sp) {
  if (dim == 2) rgi_interp<2>(interp, a, sp);
  else rgi_interp<3>(interp, a, sp);
}

}  // namespace pyftle
