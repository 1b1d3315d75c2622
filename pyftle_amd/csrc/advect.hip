// Fused advection kernel: Euler / AB2 / RK4 (integrate.py:16-109,161-342)
// through a gridded velocity field (bi/trilinear in space, frozen or linear in
// time), positions held in registers for the whole window, with an optional
// FP64 FTLE epilogue (cauchy_green.py:15-153, ftle.py:9-233,
// ftle_solver.py:122-161) so one launch turns initial seed groups into an
// FTLE field.
//
// Work mapping: a work-group holds G seeds x nb particles (nb = 2*dim for the
// FTLE window, nb = 1 for plain particle advection).  Lane order is
// seed-major (the nb particles of a seed in adjacent lanes) or
// neighbour-major (one neighbour direction of consecutive seeds per wave).
//
// FP32 compute keeps each coordinate as a compensated pair (hi, c) with
// true = hi - c (Kahan summation of the per-step increments), initialised from
// the FP64 input, so the neighbour separations the Jacobian needs do not drown
// in the rounding of absolute positions.  FP64 compute follows the reference
// operation order (stage x + (0.5*h)*k, update x += (h/6)*(k1+2k2+2k3+k4)).
#include <cmath>
#include <string>

#include "common.hpp"

namespace pyftle {

struct AdvectArgs {
  FieldDev f;
  const void *pos_in;
  void *pos_out;
  void *vprev;
  double *ftle_out;
  int64_t n_groups;
  int nb;
  int G;
  int lane_order;
  int remap;
  int pos_f64;
  int n_steps;
  int first;
  int have_prev;
  double t_start;
  double h;
  double map_period;
};

template <typename Real>
struct PosState;

template <>
struct PosState<double> {
  double x;
  __device__ __forceinline__ void init(double v) { x = v; }
  __device__ __forceinline__ double hi() const { return x; }
  __device__ __forceinline__ void add(double inc) { x += inc; }
  __device__ __forceinline__ double value() const { return x; }
};

template <>
struct PosState<float> {
  float x, c;  // true value = x - c
  __device__ __forceinline__ void init(double v) {
    x = (float)v;
    c = (float)((double)x - v);
  }
  __device__ __forceinline__ float hi() const { return x; }
  __device__ __forceinline__ void add(float inc) {
    const float y = inc - c;
    const float t = x + y;
    c = (t - x) - y;
    x = t;
  }
  __device__ __forceinline__ double value() const { return (double)x - (double)c; }
};

template <typename Real>
__device__ __forceinline__ Real stage_pos(Real x, Real hh, Real k);
template <>
__device__ __forceinline__ double stage_pos<double>(double x, double hh, double k) {
  return x + hh * k;  // integrate.py:327-338: tmp = x; tmp += (c*h) * k
}
template <>
__device__ __forceinline__ float stage_pos<float>(float x, float hh, float k) {
  return fmaf(hh, k, x);
}

template <typename Real>
__device__ __forceinline__ Real rk4_increment(Real h6, Real k1, Real k2, Real k3, Real k4);
template <>
__device__ __forceinline__ double rk4_increment<double>(double h6, double k1, double k2, double k3,
                                                        double k4) {
  return h6 * (k1 + 2.0 * k2 + 2.0 * k3 + k4);  // integrate.py:109
}
template <>
__device__ __forceinline__ float rk4_increment<float>(float h6, float k1, float k2, float k3,
                                                      float k4) {
  return h6 * (fmaf(2.0f, k2 + k3, k1) + k4);
}

template <typename Real>
__device__ __forceinline__ Real ab2_increment(Real h, Real v, Real vp) {
  return h * ((Real)1.5 * v - (Real)0.5 * vp);  // integrate.py:78
}

template <int DIM, int METHOD, int TMODE, bool PAIR, typename Real, bool FTLE>
__global__ void __launch_bounds__(512) advect_kernel(AdvectArgs a) {
  const int64_t lb = logical_block(a.remap);
  const int t = threadIdx.x;
  const int nb = a.nb, G = a.G;
  int sl, j;
  if (a.lane_order == PYFTLE_ORDER_SEED_MAJOR) {
    sl = t / nb;
    j = t - sl * nb;
  } else {
    j = t / G;
    sl = t - j * G;
  }
  const int64_t s = lb * G + sl;
  const bool active = (t < G * nb) && (s < a.n_groups);
  const int64_t pidx = (int64_t)j * a.n_groups + s;

  double x0[DIM];
  PosState<Real> X[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    double v = 0.0;
    if (active)
      v = a.pos_f64 ? ((const double *)a.pos_in)[pidx * DIM + c]
                    : (double)((const float *)a.pos_in)[pidx * DIM + c];
    x0[c] = v;
    X[c].init(v);
  }

  const FieldDev &f = a.f;
  const Real h = (Real)a.h;
  Real vprev[DIM];
  bool have_prev = (METHOD == PYFTLE_AB2) && a.have_prev;
  if (METHOD == PYFTLE_AB2) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      Real v = (Real)0;
      if (active && have_prev && a.vprev)
        v = a.pos_f64 ? (Real)((const double *)a.vprev)[pidx * DIM + c]
                      : (Real)((const float *)a.vprev)[pidx * DIM + c];
      vprev[c] = v;
    }
  }

  const char *steady_slot = slot_ptr(f, a.first);
  for (int step = 0; step < a.n_steps; ++step) {
    const char *s0 = steady_slot;
    const char *s1 = steady_slot;
    const char *s2 = steady_slot;
    double w0 = 0.0, w1 = 0.0, w2 = 0.0;
    if (TMODE == PYFTLE_TIME_FROZEN) {
      s0 = s1 = s2 = slot_ptr(f, a.first + step);
    } else if (TMODE == PYFTLE_TIME_LINEAR) {
      const double tt = a.t_start + step * a.h;
      int k;
      time_linear_lookup(f, tt, k, w0);
      s0 = slot_ptr(f, k);
      if (METHOD == PYFTLE_RK4) {
        time_linear_lookup(f, tt + 0.5 * a.h, k, w1);
        s1 = slot_ptr(f, k);
        time_linear_lookup(f, tt + a.h, k, w2);
        s2 = slot_ptr(f, k);
      }
    }
    Real xs[DIM], k1[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) xs[c] = X[c].hi();
    sample_field<DIM, PAIR>(f, s0, (Real)w0, xs, k1);
    if (METHOD == PYFTLE_RK4) {
      Real k2[DIM], k3[DIM], k4[DIM];
      const Real hh = (Real)(0.5 * a.h);
#pragma unroll
      for (int c = 0; c < DIM; ++c) xs[c] = stage_pos<Real>(X[c].hi(), hh, k1[c]);
      sample_field<DIM, PAIR>(f, s1, (Real)w1, xs, k2);
#pragma unroll
      for (int c = 0; c < DIM; ++c) xs[c] = stage_pos<Real>(X[c].hi(), hh, k2[c]);
      sample_field<DIM, PAIR>(f, s1, (Real)w1, xs, k3);
#pragma unroll
      for (int c = 0; c < DIM; ++c) xs[c] = stage_pos<Real>(X[c].hi(), h, k3[c]);
      sample_field<DIM, PAIR>(f, s2, (Real)w2, xs, k4);
      const Real h6 = (Real)(a.h / 6.0);
#pragma unroll
      for (int c = 0; c < DIM; ++c) X[c].add(rk4_increment<Real>(h6, k1[c], k2[c], k3[c], k4[c]));
    } else if (METHOD == PYFTLE_EULER) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) X[c].add(h * k1[c]);  // integrate.py:44
    } else {
      if (have_prev) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) X[c].add(ab2_increment<Real>(h, k1[c], vprev[c]));
      } else {
#pragma unroll
        for (int c = 0; c < DIM; ++c) X[c].add(h * k1[c]);  // first AB2 step is Euler
      }
#pragma unroll
      for (int c = 0; c < DIM; ++c) vprev[c] = k1[c];
      have_prev = true;
    }
  }

  if (active && a.pos_out) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      if (a.pos_f64)
        ((double *)a.pos_out)[pidx * DIM + c] = X[c].value();
      else
        ((float *)a.pos_out)[pidx * DIM + c] = (float)X[c].value();
    }
  }
  if (METHOD == PYFTLE_AB2 && active && a.vprev && a.n_steps > 0) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      if (a.pos_f64)
        ((double *)a.vprev)[pidx * DIM + c] = (double)vprev[c];
      else
        ((float *)a.vprev)[pidx * DIM + c] = (float)vprev[c];
    }
  }

  if constexpr (FTLE) {
    // LDS exchange: final and initial positions of every particle of the
    // block, then lane sl < G forms the Jacobian of seed sl in FP64.
    __shared__ double sh_fin[512 * DIM];
    __shared__ double sh_ini[512 * DIM];
    const int li = sl * nb + j;
    if (t < G * nb) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        sh_fin[li * DIM + c] = X[c].value();
        sh_ini[li * DIM + c] = x0[c];
      }
    }
    __syncthreads();
    if (t < G) {
      const int64_t seed = lb * G + t;
      if (seed < a.n_groups) {
        double F[DIM * DIM];
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
          const int pp = t * nb + pair_plus(c);
          const int pm = t * nb + pair_minus(c);
          const double d0 = sh_ini[pp * DIM + c] - sh_ini[pm * DIM + c];
#pragma unroll
          for (int r = 0; r < DIM; ++r)
            F[r * DIM + c] = (sh_fin[pp * DIM + r] - sh_fin[pm * DIM + r]) / d0;
        }
        a.ftle_out[seed] = ftle_of_jacobian<DIM>(F, a.map_period);
      }
    }
  }
}

template <int DIM, int METHOD, int TMODE, bool PAIR, typename Real, bool FTLE>
static void launch_one(const AdvectArgs &a, unsigned blocks, unsigned threads, hipStream_t st) {
  advect_kernel<DIM, METHOD, TMODE, PAIR, Real, FTLE><<<blocks, threads, 0, st>>>(a);
}

template <int DIM, int METHOD, int TMODE, bool PAIR, bool FTLE>
static void dispatch_real(const AdvectArgs &a, int f64, unsigned blocks, unsigned threads,
                          hipStream_t st) {
  if (f64)
    launch_one<DIM, METHOD, TMODE, PAIR, double, FTLE>(a, blocks, threads, st);
  else
    launch_one<DIM, METHOD, TMODE, PAIR, float, FTLE>(a, blocks, threads, st);
}

template <int DIM, int METHOD, bool FTLE>
static void dispatch_time(const AdvectArgs &a, int tmode, int pair, int f64, unsigned blocks,
                          unsigned threads, hipStream_t st) {
  if (tmode == PYFTLE_TIME_LINEAR) {
    dispatch_real<DIM, METHOD, PYFTLE_TIME_LINEAR, true, FTLE>(a, f64, blocks, threads, st);
  } else if (tmode == PYFTLE_TIME_FROZEN) {
    if (pair)
      dispatch_real<DIM, METHOD, PYFTLE_TIME_FROZEN, true, FTLE>(a, f64, blocks, threads, st);
    else
      dispatch_real<DIM, METHOD, PYFTLE_TIME_FROZEN, false, FTLE>(a, f64, blocks, threads, st);
  } else {
    if (pair)
      dispatch_real<DIM, METHOD, PYFTLE_TIME_STEADY, true, FTLE>(a, f64, blocks, threads, st);
    else
      dispatch_real<DIM, METHOD, PYFTLE_TIME_STEADY, false, FTLE>(a, f64, blocks, threads, st);
  }
}

template <bool FTLE>
static void dispatch_all(const AdvectArgs &a, int dim, int method, int tmode, int pair, int f64,
                         unsigned blocks, unsigned threads, hipStream_t st) {
  if (dim == 2) {
    if (method == PYFTLE_RK4) dispatch_time<2, PYFTLE_RK4, FTLE>(a, tmode, pair, f64, blocks, threads, st);
    else if (method == PYFTLE_EULER) dispatch_time<2, PYFTLE_EULER, FTLE>(a, tmode, pair, f64, blocks, threads, st);
    else dispatch_time<2, PYFTLE_AB2, FTLE>(a, tmode, pair, f64, blocks, threads, st);
  } else {
    if (method == PYFTLE_RK4) dispatch_time<3, PYFTLE_RK4, FTLE>(a, tmode, pair, f64, blocks, threads, st);
    else if (method == PYFTLE_EULER) dispatch_time<3, PYFTLE_EULER, FTLE>(a, tmode, pair, f64, blocks, threads, st);
    else dispatch_time<3, PYFTLE_AB2, FTLE>(a, tmode, pair, f64, blocks, threads, st);
  }
}

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
    if (field->layout != PYFTLE_LAYOUT_PAIR) return fail("advect: LINEAR time needs the PAIR layout");
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
  return PYFTLE_OK;
}

}  // namespace pyftle

using namespace pyftle;

extern "C" {

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
  const int64_t blocks = (n_particles + a.G - 1) / a.G;
  if (blocks > INT32_MAX) return fail("advect: too many particles");
  dispatch_all<false>(a, field->dim, opts->method, opts->time_mode,
                      field->layout == PYFTLE_LAYOUT_PAIR, field->dtype == PYFTLE_F64,
                      (unsigned)blocks, 256, as_stream(stream));
  return check_launch("advect");
}

int pyftle_advect_ftle(const pyftle_field *field, const pyftle_advect_opts *opts,
                       const void *positions_in, int32_t pos_dtype, int64_t n_groups,
                       void *positions_out, double map_period, double *ftle_out, void *stream) {
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
  const int64_t blocks = (n_groups + G - 1) / G;
  if (blocks > INT32_MAX) return fail("advect_ftle: too many seeds");
  int threads = G * nb;
  threads = ((threads + 63) / 64) * 64;
  dispatch_all<true>(a, field->dim, opts->method, opts->time_mode,
                     field->layout == PYFTLE_LAYOUT_PAIR, field->dtype == PYFTLE_F64,
                     (unsigned)blocks, (unsigned)threads, as_stream(stream));
  return check_launch("advect_ftle");
}

}  // extern "C"
