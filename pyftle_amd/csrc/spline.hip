// Cubic-spline ingest for the RegularGridInterpolator(method="cubic") mode
// (interpolate.py:546-762).  SciPy fits a tensor-product not-a-knot cubic
// spline through the grid values; the same piecewise cubic is written here in
// the uniform cubic B-spline basis with one coefficient per node plus one on
// either side (c[-1..n] per axis), so evaluation is the same 4-weight stencil
// in every interval.
//
// Per grid line of n values f[0..n-1] the coefficients solve
//   (c[k-1] + 4 c[k] + c[k+1]) / 6 = f[k]                     k = 0..n-1
//   c[-1] - 4 c[0] + 6 c[1] - 4 c[2] + c[3] = 0               (no third-derivative
//   c[n-4] - 4 c[n-3] + 6 c[n-2] - 4 c[n-1] + c[n] = 0         jump at x1, x(n-2))
// Eliminating c[-1], c[0], c[3] (and the mirror image) gives
//   c[1] = (8 f[1] - f[0] - f[2]) / 6,   c[n-2] = (8 f[n-2] - f[n-3] - f[n-1]) / 6,
// a diagonally dominant tridiagonal system (4 on the diagonal) for c[2..n-3],
// solved by the Thomas algorithm, and the four outer coefficients from the
// interpolation rows 0, 1, n-2, n-1.  The tensor product applies this along x,
// then y, then z.  All sweeps run in FP64; the last result is converted into
// the slot's record dtype.
#include <string>

#include "common.hpp"

namespace pyftle {

constexpr int kBetaTable = 64;  // 1/beta_k; beta_k = 4 - 1/beta_{k-1} reaches 2+sqrt(3) to
                                // double precision within ~15 steps

// One thread per line along `axis`.  in: FP64 records (C doubles) on dims
// din[]; out: FP64 records on dims dout[] (= din with dout[axis] = din[axis]+2).
template <int C>
__global__ void __launch_bounds__(256) spline_axis_kernel(const double *__restrict__ in,
                                                          double *__restrict__ out, int axis,
                                                          int dx, int dy, int dz) {
  __shared__ double inv_beta[kBetaTable];
  if (threadIdx.x == 0) {
    double b = 4.0;
    for (int k = 0; k < kBetaTable; ++k) {
      if (k > 0) b = 4.0 - 1.0 / b;
      inv_beta[k] = 1.0 / b;
    }
  }
  __syncthreads();
  const int din[3] = {dx, dy, dz};
  int dout[3] = {dx, dy, dz};
  dout[axis] += 2;
  const int n = din[axis];
  // the two other axes enumerate the lines
  const int oa = axis == 0 ? 1 : 0;
  const int ob = axis == 2 ? 1 : 2;
  const int64_t lines = (int64_t)din[oa] * din[ob];
  const int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (line >= lines) return;
  int idx[3] = {0, 0, 0};
  idx[oa] = (int)(line % din[oa]);
  idx[ob] = (int)(line / din[oa]);
  const int64_t sin_[3] = {1, (int64_t)din[0], (int64_t)din[0] * din[1]};
  const int64_t sout[3] = {1, (int64_t)dout[0], (int64_t)dout[0] * dout[1]};
  const int64_t bin = idx[0] * sin_[0] + idx[1] * sin_[1] + idx[2] * sin_[2];
  const int64_t bout = idx[0] * sout[0] + idx[1] * sout[1] + idx[2] * sout[2];
  const int64_t si = sin_[axis], so = sout[axis];
  auto F = [&](int k, int c) { return in[(bin + k * si) * C + c]; };
  auto O = [&](int j) -> double * { return out + (bout + (int64_t)(j + 1) * so) * C; };  // c_j

  double c1[C], cn2[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    c1[c] = (8.0 * F(1, c) - F(0, c) - F(2, c)) / 6.0;
    cn2[c] = (8.0 * F(n - 2, c) - F(n - 3, c) - F(n - 1, c)) / 6.0;
  }
  const int m = n - 4;  // unknowns c_2 .. c_{n-3}
  // forward sweep: d'_k stored in place of c_{k+2}
  double dp[C];
#pragma unroll
  for (int c = 0; c < C; ++c) dp[c] = 0.0;
  for (int k = 0; k < m; ++k) {
    const double ib = inv_beta[k < kBetaTable ? k : kBetaTable - 1];
    double *o = O(k + 2);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      double rhs = 6.0 * F(k + 2, c);
      if (k == 0) rhs -= c1[c];
      if (k == m - 1) rhs -= cn2[c];
      dp[c] = (rhs - (k > 0 ? dp[c] : 0.0)) * ib;
      o[c] = dp[c];
    }
  }
  // back substitution: c_{k+2} = d'_k - c_{k+3} / beta_k
  double nxt[C];
#pragma unroll
  for (int c = 0; c < C; ++c) nxt[c] = dp[c];
  for (int k = m - 2; k >= 0; --k) {
    const double ib = inv_beta[k < kBetaTable ? k : kBetaTable - 1];
    double *o = O(k + 2);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      nxt[c] = o[c] - nxt[c] * ib;
      o[c] = nxt[c];
    }
  }
  double *o1 = O(1), *on2 = O(n - 2);
#pragma unroll
  for (int c = 0; c < C; ++c) {
    o1[c] = c1[c];
    on2[c] = cn2[c];
  }
  // outer coefficients from the interpolation rows 1, 0 and n-2, n-1
  double *o0 = O(0), *om1 = O(-1), *on1 = O(n - 1), *on = O(n);
  const double *o2 = O(2), *on3 = O(n - 3);
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const double c2 = o2[c], cn3 = on3[c];
    const double c0 = 6.0 * F(1, c) - 4.0 * c1[c] - c2;
    o0[c] = c0;
    om1[c] = 6.0 * F(0, c) - 4.0 * c0 - c1[c];
    const double cn1 = 6.0 * F(n - 2, c) - 4.0 * cn2[c] - cn3;
    on1[c] = cn1;
    on[c] = 6.0 * F(n - 1, c) - 4.0 * cn1 - cn2[c];
  }
}

template <typename Dst, int C>
__global__ void __launch_bounds__(256) spline_store_kernel(const double *__restrict__ in,
                                                           Dst *__restrict__ slot, int64_t recs) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= recs) return;
#pragma unroll
  for (int c = 0; c < C; ++c) slot[r * C + c] = (Dst)in[r * C + c];
}

// FP64 node values (records of C doubles, x fastest) from the component-planar source
template <typename Src, int C>
__global__ void __launch_bounds__(256) spline_gather_kernel(const Src *__restrict__ src, int64_t s0,
                                                            int64_t s1, int64_t s2, int64_t s3,
                                                            int dim, int nx, int ny, int nz,
                                                            double *__restrict__ nodes) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (node >= (int64_t)nx * ny * nz) return;
  const int ix = (int)(node % nx);
  const int iy = (int)((node / nx) % ny);
  const int iz = (int)(node / ((int64_t)nx * ny));
  const int64_t base = (int64_t)ix * s1 + (int64_t)iy * s2 + (int64_t)iz * s3;
#pragma unroll
  for (int c = 0; c < C; ++c) nodes[node * C + c] = (c < dim) ? (double)src[base + (int64_t)c * s0] : 0.0;
}

static inline unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }

static int64_t spline_work_elems(const pyftle_field *f) {
  const int C = f->dim == 2 ? 2 : 4;
  int64_t nodes = 1, ext = 1;
  for (int a = 0; a < f->dim; ++a) {
    nodes *= f->n[a];
    ext *= (int64_t)f->n[a] + 2;
  }
  return (int64_t)C * (nodes + 2 * ext);
}

}  // namespace pyftle

using namespace pyftle;

extern "C" {

int64_t pyftle_spline_work_bytes(const pyftle_field *field) {
  if (!field || (field->dim != 2 && field->dim != 3)) return -1;
  return spline_work_elems(field) * (int64_t)sizeof(double);
}

int pyftle_pack_spline(const pyftle_field *field, int32_t snapshot, const void *src, int32_t src_dtype,
                       const int64_t src_strides[4], void *work, void *stream) {
  int rc = validate_field(field);
  if (rc) return rc;
  if (field->interp != PYFTLE_INTERP_RGI_CUBIC) return fail("pack_spline: the field is not in the cubic mode");
  if (!src || !src_strides || !work) return fail("pack_spline: NULL pointer");
  if (src_dtype != PYFTLE_F32 && src_dtype != PYFTLE_F64) return fail("pack_spline: bad src dtype");
  if (snapshot < 0 || snapshot >= field->n_snapshots) return fail("pack_spline: snapshot index out of range");
  if ((reinterpret_cast<uintptr_t>(work) % 16) != 0) return fail("pack_spline: work must be 16-byte aligned");
  const int dim = field->dim;
  const int C = dim == 2 ? 2 : 4;
  const int nx = field->n[0], ny = field->n[1], nz = dim == 3 ? field->n[2] : 1;
  const int64_t nodes = (int64_t)nx * ny * nz;
  const int64_t ext = (int64_t)(nx + 2) * (ny + 2) * (dim == 3 ? nz + 2 : 1);
  double *w_nodes = reinterpret_cast<double *>(work);
  double *w_a = w_nodes + nodes * C;
  double *w_b = w_a + ext * C;
  hipStream_t st = as_stream(stream);
  const int64_t *s = src_strides;
  if (dim == 2) {
    if (src_dtype == PYFTLE_F64)
      spline_gather_kernel<double, 2><<<grid_of(nodes), 256, 0, st>>>((const double *)src, s[0], s[1], s[2], s[3],
                                                                        dim, nx, ny, nz, w_nodes);
    else
      spline_gather_kernel<float, 2><<<grid_of(nodes), 256, 0, st>>>((const float *)src, s[0], s[1], s[2], s[3],
                                                                       dim, nx, ny, nz, w_nodes);
  } else {
    if (src_dtype == PYFTLE_F64)
      spline_gather_kernel<double, 4><<<grid_of(nodes), 256, 0, st>>>((const double *)src, s[0], s[1], s[2], s[3],
                                                                        dim, nx, ny, nz, w_nodes);
    else
      spline_gather_kernel<float, 4><<<grid_of(nodes), 256, 0, st>>>((const float *)src, s[0], s[1], s[2], s[3],
                                                                       dim, nx, ny, nz, w_nodes);
  }
  // sweeps: x (nodes -> A), y (A -> B), z (B -> A); the last buffer goes to the slot
  int d[3] = {nx, ny, nz};
  const double *in = w_nodes;
  double *bufs[2] = {w_a, w_b};
  int which = 0;
  for (int axis = 0; axis < dim; ++axis) {
    const int oa = axis == 0 ? 1 : 0, ob = axis == 2 ? 1 : 2;
    const int64_t lines = (int64_t)d[oa] * d[ob];
    double *out = bufs[which];
    if (C == 2)
      spline_axis_kernel<2><<<grid_of(lines), 256, 0, st>>>(in, out, axis, d[0], d[1], d[2]);
    else
      spline_axis_kernel<4><<<grid_of(lines), 256, 0, st>>>(in, out, axis, d[0], d[1], d[2]);
    d[axis] += 2;
    in = out;
    which ^= 1;
  }
  int sl = snapshot % field->ring_slots;
  char *slot = reinterpret_cast<char *>(field->data) + (int64_t)sl * field->slot_stride;
  if (field->dtype == PYFTLE_F64) {
    if (C == 2) spline_store_kernel<double, 2><<<grid_of(ext), 256, 0, st>>>(in, (double *)slot, ext);
    else spline_store_kernel<double, 4><<<grid_of(ext), 256, 0, st>>>(in, (double *)slot, ext);
  } else {
    if (C == 2) spline_store_kernel<float, 2><<<grid_of(ext), 256, 0, st>>>(in, (float *)slot, ext);
    else spline_store_kernel<float, 4><<<grid_of(ext), 256, 0, st>>>(in, (float *)slot, ext);
  }
  return check_launch("pack_spline");
}

}  // extern "C"
