// FTLE tensor kernels, FP64 (cauchy_green.py:15-210, ftle.py:9-233).
// One lane per seed; every stage is a pure per-seed map.
#include <string>

#include "common.hpp"

namespace pyftle {

template <typename P>
__device__ __forceinline__ double ld(const P *p, int64_t i) {
  return (double)p[i];
}

// J[:, r, c] = delta_c[:, r] / initial_delta_c[:, c], delta pairs (R-L, T-B, F-Bk)
template <int DIM, typename P>
__device__ __forceinline__ void jacobian_of(const P *pos, const P *ini, int64_t n, int64_t s,
                                            double *F) {
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    const int64_t pp = (int64_t)pair_plus(c) * n + s;
    const int64_t pm = (int64_t)pair_minus(c) * n + s;
    const double d0 = ld(ini, pp * DIM + c) - ld(ini, pm * DIM + c);
#pragma unroll
    for (int r = 0; r < DIM; ++r) F[r * DIM + c] = (ld(pos, pp * DIM + r) - ld(pos, pm * DIM + r)) / d0;
  }
}

template <int DIM, typename P>
__global__ void __launch_bounds__(256) jacobian_kernel(const P *pos, const P *ini, int64_t n,
                                                       double *J) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double F[DIM * DIM];
  jacobian_of<DIM>(pos, ini, n, s, F);
#pragma unroll
  for (int e = 0; e < DIM * DIM; ++e) J[s * DIM * DIM + e] = F[e];
}

template <int DIM, typename P>
__global__ void __launch_bounds__(256) ftle_groups_kernel(const P *pos, const P *ini, int64_t n,
                                                          double T, double *out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double F[DIM * DIM];
  jacobian_of<DIM>(pos, ini, n, s, F);
  out[s] = ftle_of_jacobian<DIM>(F, T);
}

// ftle.py:9-69 (2D, F^T F written out) and ftle.py:72-96 (3D, F F^T)
template <int DIM>
__global__ void __launch_bounds__(256) cauchy_green_kernel(const double *J, int64_t n, double *C) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double *F = J + s * DIM * DIM;
  double *o = C + s * DIM * DIM;
  if (DIM == 2) {
    o[0] = F[0] * F[0] + F[2] * F[2];
    o[1] = F[0] * F[1] + F[2] * F[3];
    o[2] = F[1] * F[0] + F[3] * F[2];
    o[3] = F[1] * F[1] + F[3] * F[3];
  } else {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        o[i * 3 + j] = F[i * 3 + 0] * F[j * 3 + 0] + F[i * 3 + 1] * F[j * 3 + 1] + F[i * 3 + 2] * F[j * 3 + 2];
  }
}

template <int DIM>
__global__ void __launch_bounds__(256) max_eig_kernel(const double *C, int64_t n, double *lam) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const double *c = C + s * DIM * DIM;
  if (DIM == 2) {
    lam[s] = max_eig_2x2(c[0], c[1], c[2], c[3]);
  } else {
    // symmetric part (C is symmetric when it comes from cauchy_green)
    lam[s] = max_eig_sym3(c[0], 0.5 * (c[1] + c[3]), 0.5 * (c[2] + c[6]), c[4], 0.5 * (c[5] + c[7]), c[8]);
  }
}

template <int DIM>
__global__ void __launch_bounds__(256) ftle_jacobian_kernel(const double *J, int64_t n, double T,
                                                            double *out) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  double F[DIM * DIM];
#pragma unroll
  for (int e = 0; e < DIM * DIM; ++e) F[e] = J[s * DIM * DIM + e];
  out[s] = ftle_of_jacobian<DIM>(F, T);
}

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

static int check_common(int dim, int64_t n, const char *who) {
  if (dim != 2 && dim != 3) return fail(std::string(who) + ": dim must be 2 or 3");
  if (n < 0) return fail(std::string(who) + ": negative count");
  if (n > (int64_t)INT32_MAX * 256) return fail(std::string(who) + ": too many seeds");
  return PYFTLE_OK;
}

}  // namespace pyftle

using namespace pyftle;

extern "C" {

int pyftle_flow_map_jacobian(const void *positions, const void *initial_positions,
                             int32_t pos_dtype, int32_t dim, int64_t n, double *J_out,
                             void *stream) {
  int rc = check_common(dim, n, "flow_map_jacobian");
  if (rc) return rc;
  if (n == 0) return PYFTLE_OK;
  if (!positions || !initial_positions || !J_out) return fail("flow_map_jacobian: NULL pointer");
  hipStream_t st = as_stream(stream);
  if (pos_dtype == PYFTLE_F64) {
    if (dim == 2) jacobian_kernel<2, double><<<nblk(n), 256, 0, st>>>((const double *)positions, (const double *)initial_positions, n, J_out);
    else jacobian_kernel<3, double><<<nblk(n), 256, 0, st>>>((const double *)positions, (const double *)initial_positions, n, J_out);
  } else if (pos_dtype == PYFTLE_F32) {
    if (dim == 2) jacobian_kernel<2, float><<<nblk(n), 256, 0, st>>>((const float *)positions, (const float *)initial_positions, n, J_out);
    else jacobian_kernel<3, float><<<nblk(n), 256, 0, st>>>((const float *)positions, (const float *)initial_positions, n, J_out);
  } else {
    return fail("flow_map_jacobian: bad dtype");
  }
  return check_launch("flow_map_jacobian");
}

int pyftle_cauchy_green(const double *J, int32_t dim, int64_t n, double *C_out, void *stream) {
  int rc = check_common(dim, n, "cauchy_green");
  if (rc) return rc;
  if (n == 0) return PYFTLE_OK;
  if (!J || !C_out) return fail("cauchy_green: NULL pointer");
  hipStream_t st = as_stream(stream);
  if (dim == 2) cauchy_green_kernel<2><<<nblk(n), 256, 0, st>>>(J, n, C_out);
  else cauchy_green_kernel<3><<<nblk(n), 256, 0, st>>>(J, n, C_out);
  return check_launch("cauchy_green");
}

int pyftle_max_eigenvalue(const double *C, int32_t dim, int64_t n, double *lam_out, void *stream) {
  int rc = check_common(dim, n, "max_eigenvalue");
  if (rc) return rc;
  if (n == 0) return PYFTLE_OK;
  if (!C || !lam_out) return fail("max_eigenvalue: NULL pointer");
  hipStream_t st = as_stream(stream);
  if (dim == 2) max_eig_kernel<2><<<nblk(n), 256, 0, st>>>(C, n, lam_out);
  else max_eig_kernel<3><<<nblk(n), 256, 0, st>>>(C, n, lam_out);
  return check_launch("max_eigenvalue");
}

int pyftle_ftle_from_jacobian(const double *J, int32_t dim, int64_t n, double map_period,
                              double *ftle_out, void *stream) {
  int rc = check_common(dim, n, "ftle_from_jacobian");
  if (rc) return rc;
  if (n == 0) return PYFTLE_OK;
  if (!J || !ftle_out) return fail("ftle_from_jacobian: NULL pointer");
  hipStream_t st = as_stream(stream);
  if (dim == 2) ftle_jacobian_kernel<2><<<nblk(n), 256, 0, st>>>(J, n, map_period, ftle_out);
  else ftle_jacobian_kernel<3><<<nblk(n), 256, 0, st>>>(J, n, map_period, ftle_out);
  return check_launch("ftle_from_jacobian");
}

int pyftle_ftle_from_groups(const void *positions, const void *initial_positions, int32_t pos_dtype,
                            int32_t dim, int64_t n, double map_period, double *ftle_out,
                            void *stream) {
  int rc = check_common(dim, n, "ftle_from_groups");
  if (rc) return rc;
  if (n == 0) return PYFTLE_OK;
  if (!positions || !initial_positions || !ftle_out) return fail("ftle_from_groups: NULL pointer");
  hipStream_t st = as_stream(stream);
  if (pos_dtype == PYFTLE_F64) {
    if (dim == 2) ftle_groups_kernel<2, double><<<nblk(n), 256, 0, st>>>((const double *)positions, (const double *)initial_positions, n, map_period, ftle_out);
    else ftle_groups_kernel<3, double><<<nblk(n), 256, 0, st>>>((const double *)positions, (const double *)initial_positions, n, map_period, ftle_out);
  } else if (pos_dtype == PYFTLE_F32) {
    if (dim == 2) ftle_groups_kernel<2, float><<<nblk(n), 256, 0, st>>>((const float *)positions, (const float *)initial_positions, n, map_period, ftle_out);
    else ftle_groups_kernel<3, float><<<nblk(n), 256, 0, st>>>((const float *)positions, (const float *)initial_positions, n, map_period, ftle_out);
  } else {
    return fail("ftle_from_groups: bad dtype");
  }
  return check_launch("ftle_from_groups");
}

}  // extern "C"
