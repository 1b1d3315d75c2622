// C entry points of the fused advection kernel (kernel in advect_impl.hpp).
#include <cmath>
#include <string>

#include "advect_impl.hpp"

namespace pyftle {

static int check_opts(const pyftle_field *field, const pyftle_advect_opts *o) {
  if (!o) return fail("advect: opts is NULL");
  if (o->method != PYFTLE_EULER && o->method != PYFTLE_AB2 && o->method != PYFTLE_RK4)
    return fail("advect: unknown integrator");
  if (o->time_mode != PYFTLE_TIME_STEADY && o->time_mode != PYFTLE_TIME_FROZEN &&
      o->time_mode != PYFTLE_TIME_LINEAR)
    return fail("advect: unknown time mode");
  if (o->n_steps < 0) return fail("advect: n_steps must be >= 0");
  if (!std::isfinite(o->h)) return fail("advect: h must be finite");
  if (o->time_mode == PYFTLE_TIME_LINEAR) {
    if (!layout_pair(field->layout)) return fail("advect: LINEAR time needs the PAIR or CELL_PAIR layout");
    if (!(field->dt_snap != 0.0) || !std::isfinite(field->dt_snap) || !std::isfinite(field->t0))
      return fail("advect: LINEAR time needs finite t0 and non-zero dt_snap");
    if (!std::isfinite(o->t_start)) return fail("advect: t_start must be finite");
  } else if (o->time_mode == PYFTLE_TIME_FROZEN) {
    if (o->first_snapshot < 0 || (int64_t)o->first_snapshot + o->n_steps > field->n_snapshots)
      return fail("advect: FROZEN steps run past the last snapshot");
  } else {
    if (o->first_snapshot < 0 || o->first_snapshot >= field->n_snapshots)
      return fail("advect: STEADY snapshot index out of range");
  }
  if (o->lane_order != PYFTLE_ORDER_SEED_MAJOR && o->lane_order != PYFTLE_ORDER_NEIGHBOR_MAJOR)
    return fail("advect: unknown lane order");
  if (o->compute != PYFTLE_COMPUTE_NATIVE && o->compute != PYFTLE_COMPUTE_MIXED &&
      o->compute != PYFTLE_COMPUTE_F64)
    return fail("advect: unknown compute precision");
  if (field->dtype == PYFTLE_F64 && o->compute != PYFTLE_COMPUTE_NATIVE)
    return fail("advect: an FP64 field always computes in FP64 (compute must be NATIVE)");
  if (layout_cell(field->layout) && o->compute != PYFTLE_COMPUTE_NATIVE)
    return fail("advect: the CELL layouts are read by the FP32 kernel only (compute must be NATIVE)");
  if (field->interp != PYFTLE_INTERP_GINTERP && o->time_mode == PYFTLE_TIME_LINEAR)
    return fail("advect: RegularGridInterpolator modes run STEADY or FROZEN time");
  return PYFTLE_OK;
}

static int prec_of(const pyftle_advect_opts *o) {
  return o->compute == PYFTLE_COMPUTE_MIXED ? PREC_MIXED
         : o->compute == PYFTLE_COMPUTE_F64 ? PREC_F64C
                                            : PREC_NATIVE;
}

static void dispatch(const pyftle_field *field, const AdvectArgs &a, const LaunchSpec &sp) {
  if (field->interp != PYFTLE_INTERP_GINTERP) launch_advect_rgi(field->dim, field->interp, a, sp);
  else if (field->dim == 2) launch_advect_2d(a, sp);
  else launch_advect_3d(a, sp);
}

// Morton code of the cell of each seed group's centroid (20 bits per axis).
__device__ __forceinline__ uint64_t spread_bits(uint64_t v, int dim) {
  uint64_t r = 0;
  for (int b = 0; b < 20; ++b) r |= ((v >> b) & 1ull) << (b * dim);
  return r;
}

__global__ void __launch_bounds__(256) seed_morton_kernel(const void *pos, int pos_f64, int64_t n, int nb,
                                                          int dim, double lo0, double lo1, double lo2,
                                                          double ic0, double ic1, double ic2, int64_t *codes) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double lo[3] = {lo0, lo1, lo2}, ic[3] = {ic0, ic1, ic2};
  uint64_t code = 0;
  for (int a = 0; a < dim; ++a) {
    double c = 0.0;
    for (int j = 0; j < nb; ++j) {
      const int64_t i = ((int64_t)j * n + s) * dim + a;
      c += pos_f64 ? ((const double *)pos)[i] : (double)((const float *)pos)[i];
    }
    double q = (c / nb - lo[a]) * ic[a];
    q = (q == q) ? fmin(fmax(q, 0.0), 1048575.0) : 0.0;
    code |= spread_bits((uint64_t)q, dim) << a;
  }
  codes[s] = (int64_t)code;
}

}  // namespace pyftle

using namespace pyftle;

extern "C" {

int pyftle_seed_morton_codes(const void *positions, int32_t pos_dtype, int64_t n_groups, int32_t nb,
                             int32_t dim, const double lo[3], const double inv_cell[3], int64_t *codes,
                             void *stream) {
  if (pos_dtype != PYFTLE_F32 && pos_dtype != PYFTLE_F64) return fail("seed_morton_codes: bad dtype");
  if (dim != 2 && dim != 3) return fail("seed_morton_codes: dim must be 2 or 3");
  if (nb < 1 || n_groups < 0) return fail("seed_morton_codes: bad sizes");
  if (n_groups == 0) return PYFTLE_OK;
  if (!positions || !codes || !lo || !inv_cell) return fail("seed_morton_codes: NULL pointer");
  const int64_t blocks = (n_groups + 255) / 256;
  seed_morton_kernel<<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(
      positions, pos_dtype == PYFTLE_F64, n_groups, nb, dim, lo[0], lo[1], dim == 3 ? lo[2] : 0.0,
      inv_cell[0], inv_cell[1], dim == 3 ? inv_cell[2] : 0.0, codes);
  return check_launch("seed_morton_codes");
}

int pyftle_advect(const pyftle_field *field, const pyftle_advect_opts *opts, void *positions,
                  int32_t pos_dtype, int64_t n_particles, void *vprev, void *stream) {
  AdvectArgs a{};
  int rc = make_field_dev(field, &a.f);
  if (rc) return rc;
  rc = check_opts(field, opts);
  if (rc) return rc;
  if (pos_dtype != PYFTLE_F32 && pos_dtype != PYFTLE_F64) return fail("advect: bad positions dtype");
  if (n_particles < 0) return fail("advect: negative particle count");
  if (n_particles == 0) return PYFTLE_OK;
  if (!positions) return fail("advect: positions is NULL");
  if (opts->method == PYFTLE_AB2 && !vprev) return fail("advect: AB2 needs a vprev buffer");
  a.pos_in = positions;
  a.pos_out = positions;
  a.vprev = vprev;
  a.ftle_out = nullptr;
  a.n_groups = n_particles;
  a.nb = 1;
  a.G = 256;
  a.lane_order = PYFTLE_ORDER_SEED_MAJOR;
  a.remap = opts->xcd_remap ? 1 : 0;
  a.pos_f64 = pos_dtype == PYFTLE_F64;
  a.n_steps = opts->n_steps;
  a.first = opts->first_snapshot;
  a.have_prev = opts->ab2_have_prev ? 1 : 0;
  a.t_start = opts->t_start;
  a.h = opts->h;
  a.map_period = 1.0;
  a.order = opts->seed_order;
  const int64_t blocks = (n_particles + a.G - 1) / a.G;
  if (blocks > INT32_MAX) return fail("advect: too many particles");
  LaunchSpec sp{opts->method, opts->time_mode, layout_pair(field->layout),
                field->dtype == PYFTLE_F64, prec_of(opts), 0, (unsigned)blocks, 256u,
                as_stream(stream), layout_cell(field->layout)};
  dispatch(field, a, sp);
  return check_launch("advect");
}

}  // extern "C"

namespace pyftle {

static int advect_ftle_impl(const pyftle_field *field, const pyftle_advect_opts *opts,
                            const void *positions_in, int32_t pos_dtype, int64_t n_groups,
                            void *positions_out, double map_period, double *ftle_out,
                            unsigned long long *load_count, void *stream) {
  AdvectArgs a{};
  int rc = make_field_dev(field, &a.f);
  if (rc) return rc;
  rc = check_opts(field, opts);
  if (rc) return rc;
  if (pos_dtype != PYFTLE_F32 && pos_dtype != PYFTLE_F64) return fail("advect_ftle: bad positions dtype");
  if (n_groups < 0) return fail("advect_ftle: negative group count");
  if (n_groups == 0) return PYFTLE_OK;
  if (!positions_in || !ftle_out) return fail("advect_ftle: NULL pointer");
  if (opts->method == PYFTLE_AB2 && opts->ab2_have_prev)
    return fail("advect_ftle: a fused window starts without AB2 history");
  const int nb = 2 * field->dim;
  int G = opts->groups_per_block > 0 ? opts->groups_per_block : 64;
  if (G * nb > 512) return fail("advect_ftle: groups_per_block * 2*dim must be <= 512");
  a.pos_in = positions_in;
  a.pos_out = positions_out;
  a.vprev = nullptr;
  a.ftle_out = ftle_out;
  a.n_groups = n_groups;
  a.nb = nb;
  a.G = G;
  a.lane_order = opts->lane_order;
  a.remap = opts->xcd_remap ? 1 : 0;
  a.pos_f64 = pos_dtype == PYFTLE_F64;
  a.n_steps = opts->n_steps;
  a.first = opts->first_snapshot;
  a.have_prev = 0;
  a.t_start = opts->t_start;
  a.h = opts->h;
  a.map_period = map_period;
  a.order = opts->seed_order;
  a.load_count = load_count;
  const int64_t blocks = (n_groups + G - 1) / G;
  if (blocks > INT32_MAX) return fail("advect_ftle: too many seeds");
  int threads = G * nb;
  threads = ((threads + 63) / 64) * 64;
  LaunchSpec sp{opts->method, opts->time_mode, layout_pair(field->layout),
                field->dtype == PYFTLE_F64, prec_of(opts), 1, (unsigned)blocks, (unsigned)threads,
                as_stream(stream), layout_cell(field->layout), load_count ? 1 : 0};
  dispatch(field, a, sp);
  return check_launch("advect_ftle");
}

}  // namespace pyftle

extern "C" {

int pyftle_advect_ftle(const pyftle_field *field, const pyftle_advect_opts *opts,
                       const void *positions_in, int32_t pos_dtype, int64_t n_groups,
                       void *positions_out, double map_period, double *ftle_out, void *stream) {
  return advect_ftle_impl(field, opts, positions_in, pos_dtype, n_groups, positions_out, map_period,
                          ftle_out, nullptr, stream);
}

int pyftle_advect_ftle_count_loads(const pyftle_field *field, const pyftle_advect_opts *opts,
                                   const void *positions_in, int32_t pos_dtype, int64_t n_groups,
                                   void *positions_out, double map_period, double *ftle_out,
                                   unsigned long long *loads, void *stream) {
  if (!loads) return fail("advect_ftle_count_loads: loads is NULL");
  if (!field || !opts) return fail("advect_ftle_count_loads: NULL field or opts");
  if (!layout_cell(field->layout) || field->dtype != PYFTLE_F32 || field->interp != PYFTLE_INTERP_GINTERP ||
      opts->method != PYFTLE_RK4 || opts->compute != PYFTLE_COMPUTE_NATIVE)
    return fail("advect_ftle_count_loads: counts RK4 windows of the FP32 kernel on CELL / CELL_PAIR records");
  return advect_ftle_impl(field, opts, positions_in, pos_dtype, n_groups, positions_out, map_period,
                          ftle_out, loads, stream);
}


}  // extern "C"
