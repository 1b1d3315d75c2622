// Gather-only interpolation (the reference's ginterp module), snapshot ingest
// into the device ring, and physical-space field sampling.
#include <string>

#include "common.hpp"
#include "rgi.hpp"
#include "scatter.hpp"

namespace pyftle {

// GridInterp.cpp:93-142, one lane per query point.
__global__ void __launch_bounds__(256) interp2d_vec_kernel(const double *__restrict__ v, int nx,
                                                           int ny, const double *__restrict__ pts,
                                                           int64_t n, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double px = pts[i * 2 + 0];
  const double py = pts[i * 2 + 1];
  const double fx = floor(px), fy = floor(py);
  // static_cast<int> of a value outside int range (or NaN) yields INT_MIN on
  // x86, which fails the bounds test below; reproduce that.
  const bool rx = (fx > -2147483649.0) && (fx < 2147483648.0);
  const bool ry = (fy > -2147483649.0) && (fy < 2147483648.0);
  const int x0 = rx ? (int)fx : INT32_MIN;
  const int y0 = ry ? (int)fy : INT32_MIN;
  const double xd = px - x0;
  const double yd = py - y0;
  double c = 0.0;
  if (x0 >= 0 && x0 < nx - 1 && y0 >= 0 && y0 < ny - 1) {
    const int x1 = x0 + 1, y1 = y0 + 1;
    const double v00 = v[(int64_t)x0 * ny + y0];
    const double v10 = v[(int64_t)x1 * ny + y0];
    const double v01 = v[(int64_t)x0 * ny + y1];
    const double v11 = v[(int64_t)x1 * ny + y1];
    c = v00 * (1.0 - xd) * (1.0 - yd) + v10 * xd * (1.0 - yd) + v01 * (1.0 - xd) * yd +
        v11 * xd * yd;
  }
  out[i] = c;
}

// GridInterp.cpp:15-91
__global__ void __launch_bounds__(256) interp3d_vec_kernel(const double *__restrict__ v, int X,
                                                           int Y, int Z,
                                                           const double *__restrict__ pts,
                                                           int64_t n, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double px = pts[i * 3 + 0], py = pts[i * 3 + 1], pz = pts[i * 3 + 2];
  const double fx = floor(px), fy = floor(py), fz = floor(pz);
  const int x0 = ((fx > -2147483649.0) && (fx < 2147483648.0)) ? (int)fx : INT32_MIN;
  const int y0 = ((fy > -2147483649.0) && (fy < 2147483648.0)) ? (int)fy : INT32_MIN;
  const int z0 = ((fz > -2147483649.0) && (fz < 2147483648.0)) ? (int)fz : INT32_MIN;
  const double xd = px - x0, yd = py - y0, zd = pz - z0;
  double c = 0.0;
  if (x0 >= 0 && x0 < X - 1 && y0 >= 0 && y0 < Y - 1 && z0 >= 0 && z0 < Z - 1) {
    const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    const int64_t YZ = (int64_t)Y * Z;
    auto idx = [&](int a, int b, int d) -> int64_t { return (int64_t)a * YZ + (int64_t)b * Z + d; };
    const double c00 = v[idx(x0, y0, z0)] * (1.0 - xd) + v[idx(x1, y0, z0)] * xd;
    const double c01 = v[idx(x0, y0, z1)] * (1.0 - xd) + v[idx(x1, y0, z1)] * xd;
    const double c10 = v[idx(x0, y1, z0)] * (1.0 - xd) + v[idx(x1, y1, z0)] * xd;
    const double c11 = v[idx(x0, y1, z1)] * (1.0 - xd) + v[idx(x1, y1, z1)] * xd;
    const double c0 = c00 * (1.0 - yd) + c10 * yd;
    const double c1 = c01 * (1.0 - yd) + c11 * yd;
    c = c0 * (1.0 - zd) + c1 * zd;
  }
  out[i] = c;
}

// Snapshot ingest: one lane per (padded) grid node, reads the component-planar
// source through element strides, writes one time level of the interleaved
// record.  A periodic axis carries one ghost node at index n holding node 0,
// so an upper cell corner is always lower index + 1.
template <typename Dst, typename Src>
__global__ void __launch_bounds__(256) pack_kernel(char *__restrict__ slot_a, int level_a,
                                                   char *__restrict__ slot_b, int level_b,
                                                   const Src *__restrict__ src, int64_t s0,
                                                   int64_t s1, int64_t s2, int64_t s3, int dim,
                                                   int nx, int ny, int nz, int gx, int gy, int gz,
                                                   int per_level, int comps) {
  const int64_t node = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nodes = (int64_t)gx * gy * gz;
  if (node >= nodes) return;
  int ix = (int)(node % gx);
  int iy = (int)((node / gx) % gy);
  int iz = (int)(node / ((int64_t)gx * gy));
  ix = ix >= nx ? ix - nx : ix;
  iy = iy >= ny ? iy - ny : iy;
  iz = iz >= nz ? iz - nz : iz;
  const int64_t base = (int64_t)ix * s1 + (int64_t)iy * s2 + (int64_t)iz * s3;
  Dst vals[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) vals[c] = (c < dim) ? (Dst)src[base + (int64_t)c * s0] : (Dst)0;
  Dst *ra = reinterpret_cast<Dst *>(slot_a) + node * comps + level_a * per_level;
  for (int c = 0; c < per_level; ++c) ra[c] = vals[c];
  if (slot_b) {
    Dst *rb = reinterpret_cast<Dst *>(slot_b) + node * comps + level_b * per_level;
    for (int c = 0; c < per_level; ++c) rb[c] = vals[c];
  }
}

// Snapshot ingest into per-cell records (PYFTLE_LAYOUT_CELL, CELL_PAIR): one lane per cell
// writes the dim components of the cell's 2^dim corner nodes, corner after corner, into one
// time level.  Upper corners wrap on periodic axes; on the others the last cell (never read
// as valid) repeats the last node.
template <int DIM, typename Src>
__global__ void __launch_bounds__(256) pack_cell_kernel(char *__restrict__ slot_a, int level_a,
                                                        char *__restrict__ slot_b, int level_b,
                                                        const Src *__restrict__ src, int64_t s0, int64_t s1,
                                                        int64_t s2, int64_t s3, int nx, int ny, int nz,
                                                        int px, int py, int pz, int rec_floats) {
  constexpr int NQ = 1 << DIM;
  const int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= (int64_t)nx * ny * nz) return;
  const int ix = (int)(cell % nx);
  const int iy = (int)((cell / nx) % ny);
  const int iz = (int)(cell / ((int64_t)nx * ny));
  auto up = [](int i, int n, int periodic) { return i + 1 < n ? i + 1 : (periodic ? 0 : n - 1); };
  float vals[NQ * DIM];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int jx = (q & 1) ? up(ix, nx, px) : ix;
    const int jy = (q & 2) ? up(iy, ny, py) : iy;
    const int jz = (DIM == 3 && (q & 4)) ? up(iz, nz, pz) : iz;
    const int64_t base = (int64_t)jx * s1 + (int64_t)jy * s2 + (int64_t)jz * s3;
#pragma unroll
    for (int c = 0; c < DIM; ++c) vals[q * DIM + c] = (float)src[base + (int64_t)c * s0];
  }
  float *ra = reinterpret_cast<float *>(slot_a) + cell * rec_floats + level_a * NQ * DIM;
#pragma unroll
  for (int e = 0; e < NQ * DIM; ++e) ra[e] = vals[e];
  if (slot_b) {
    float *rb = reinterpret_cast<float *>(slot_b) + cell * rec_floats + level_b * NQ * DIM;
#pragma unroll
    for (int e = 0; e < NQ * DIM; ++e) rb[e] = vals[e];
  }
}

// Physical-space sampling, one lane per point.
template <int DIM, bool PAIR, typename Real, typename P>
__global__ void __launch_bounds__(256) sample_kernel(FieldDev f, const char *slot, double w,
                                                     const P *__restrict__ pts, int64_t n,
                                                     P *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Real x[DIM], v[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) x[a] = (Real)pts[i * DIM + a];
  sample_field<DIM, PAIR>(f, slot, (Real)w, x, v);
#pragma unroll
  for (int a = 0; a < DIM; ++a) out[i * DIM + a] = (P)v[a];
}

// The same over per-cell records (FP32 fields): bit-identical to the node records.
template <int DIM, bool PAIR, typename P>
__global__ void __launch_bounds__(256) sample_cell_kernel(FieldDev f, const char *slot, double w,
                                                          const P *__restrict__ pts, int64_t n,
                                                          P *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x[DIM], v[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) x[a] = (float)pts[i * DIM + a];
  sample_field_f32<DIM, PAIR, float, float, true>(f, slot, (float)w, x, v);
#pragma unroll
  for (int a = 0; a < DIM; ++a) out[i * DIM + a] = (P)v[a];
}

// RegularGridInterpolator modes (rgi.hpp) and scattered-data modes (scatter.hpp),
// one lane per point, FP64 arithmetic.
template <int DIM, int INTERP, typename T, typename P>
__global__ void __launch_bounds__(256) sample_rgi_kernel(FieldDev f, const char *slot,
                                                         const P *__restrict__ pts, int64_t n,
                                                         P *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[DIM], v[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) x[a] = (double)pts[i * DIM + a];
  if constexpr (INTERP == PYFTLE_INTERP_SCATTER_LINEAR || INTERP == PYFTLE_INTERP_SCATTER_NEAREST ||
                INTERP == PYFTLE_INTERP_SCATTER_CUBIC)
    sample_scatter<DIM, INTERP, T>(f, slot, x, v);
  else
    sample_rgi<DIM, INTERP, T>(f, slot, x, v);
#pragma unroll
  for (int a = 0; a < DIM; ++a) out[i * DIM + a] = (P)v[a];
}

static inline unsigned blocks_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

template <int DIM, int INTERP, typename T>
static void launch_sample_rgi_t(const FieldDev &d, const char *slot, const void *pts, int pd, int64_t n,
                                void *out, hipStream_t st) {
  if (pd == PYFTLE_F64)
    sample_rgi_kernel<DIM, INTERP, T, double><<<blocks_for(n, 256), 256, 0, st>>>(
        d, slot, (const double *)pts, n, (double *)out);
  else
    sample_rgi_kernel<DIM, INTERP, T, float><<<blocks_for(n, 256), 256, 0, st>>>(
        d, slot, (const float *)pts, n, (float *)out);
}

template <int DIM, int INTERP>
static void launch_sample_rgi_i(const FieldDev &d, bool f64, const char *slot, const void *pts, int pd,
                                int64_t n, void *out, hipStream_t st) {
  if (f64) launch_sample_rgi_t<DIM, INTERP, double>(d, slot, pts, pd, n, out, st);
  else launch_sample_rgi_t<DIM, INTERP, float>(d, slot, pts, pd, n, out, st);
}

template <int DIM>
static void launch_sample_rgi(const FieldDev &d, bool f64, const char *slot, const void *pts, int pd,
                              int64_t n, void *out, hipStream_t st) {
  if (d.interp == PYFTLE_INTERP_RGI_LINEAR)
    launch_sample_rgi_i<DIM, PYFTLE_INTERP_RGI_LINEAR>(d, f64, slot, pts, pd, n, out, st);
  else if (d.interp == PYFTLE_INTERP_RGI_NEAREST)
    launch_sample_rgi_i<DIM, PYFTLE_INTERP_RGI_NEAREST>(d, f64, slot, pts, pd, n, out, st);
  else if (d.interp == PYFTLE_INTERP_RGI_CUBIC)
    launch_sample_rgi_i<DIM, PYFTLE_INTERP_RGI_CUBIC>(d, f64, slot, pts, pd, n, out, st);
  else if (d.interp == PYFTLE_INTERP_SCATTER_LINEAR)
    launch_sample_rgi_i<DIM, PYFTLE_INTERP_SCATTER_LINEAR>(d, f64, slot, pts, pd, n, out, st);
  else if (d.interp == PYFTLE_INTERP_SCATTER_NEAREST)
    launch_sample_rgi_i<DIM, PYFTLE_INTERP_SCATTER_NEAREST>(d, f64, slot, pts, pd, n, out, st);
  else if constexpr (DIM == 2)
    launch_sample_rgi_i<DIM, PYFTLE_INTERP_SCATTER_CUBIC>(d, f64, slot, pts, pd, n, out, st);
}

template <int DIM, bool PAIR, typename Real>
static void launch_sample(const FieldDev &d, const char *slot, double w, const void *pts, int pd,
                          int64_t n, void *out, hipStream_t st) {
  if (pd == PYFTLE_F64)
    sample_kernel<DIM, PAIR, Real, double><<<blocks_for(n, 256), 256, 0, st>>>(
        d, slot, w, (const double *)pts, n, (double *)out);
  else
    sample_kernel<DIM, PAIR, Real, float><<<blocks_for(n, 256), 256, 0, st>>>(
        d, slot, w, (const float *)pts, n, (float *)out);
}

template <int DIM, bool PAIR>
static void launch_sample_cell(const FieldDev &d, const char *slot, double w, const void *pts, int pd,
                               int64_t n, void *out, hipStream_t st) {
  if (pd == PYFTLE_F64)
    sample_cell_kernel<DIM, PAIR, double><<<blocks_for(n, 256), 256, 0, st>>>(
        d, slot, w, (const double *)pts, n, (double *)out);
  else
    sample_cell_kernel<DIM, PAIR, float><<<blocks_for(n, 256), 256, 0, st>>>(
        d, slot, w, (const float *)pts, n, (float *)out);
}

}  // namespace pyftle

using namespace pyftle;

extern "C" {

int pyftle_interp2d_vec(const double *v, int32_t nx, int32_t ny, const double *points, int64_t n,
                        double *out, void *stream) {
  if (!v || !points || !out) return fail("interp2d_vec: NULL pointer");
  if (nx < 0 || ny < 0 || n < 0) return fail("interp2d_vec: negative size");
  if (n == 0) return PYFTLE_OK;
  interp2d_vec_kernel<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(v, nx, ny, points, n, out);
  return check_launch("interp2d_vec");
}

int pyftle_interp3d_vec(const double *v, int32_t X, int32_t Y, int32_t Z, const double *points,
                        int64_t n, double *out, void *stream) {
  if (!v || !points || !out) return fail("interp3d_vec: NULL pointer");
  if (X < 0 || Y < 0 || Z < 0 || n < 0) return fail("interp3d_vec: negative size");
  if (n == 0) return PYFTLE_OK;
  interp3d_vec_kernel<<<blocks_for(n, 256), 256, 0, as_stream(stream)>>>(v, X, Y, Z, points, n, out);
  return check_launch("interp3d_vec");
}

int pyftle_pack_snapshot(const pyftle_field *field, int32_t snapshot, const void *src,
                         int32_t src_dtype, const int64_t src_strides[4], void *stream) {
  int rc = validate_field(field);
  if (rc) return rc;
  if (!src || !src_strides) return fail("pack_snapshot: NULL source");
  if (src_dtype != PYFTLE_F32 && src_dtype != PYFTLE_F64) return fail("pack_snapshot: bad src dtype");
  if (snapshot < 0 || snapshot >= field->n_snapshots)
    return fail("pack_snapshot: snapshot index out of range");
  if (field->interp == PYFTLE_INTERP_RGI_CUBIC)
    return fail("pack_snapshot: cubic-mode slots hold spline coefficients (use pyftle_pack_spline)");
  if (field->interp == PYFTLE_INTERP_SCATTER_CUBIC)
    return fail("pack_snapshot: scattered cubic slots hold values and gradients (write the records directly)");
  const int dim = field->dim;
  const int nz = dim == 3 ? field->n[2] : 1;
  const int per_level = dim == 2 ? 2 : 4;
  const int pair = layout_pair(field->layout);
  const int comps = per_level * (pair ? 2 : 1);
  char *base = reinterpret_cast<char *>(field->data);
  auto slot = [&](int k) {
    int s = k % field->ring_slots;
    return base + (int64_t)s * field->slot_stride;
  };
  char *a = nullptr, *b = nullptr;
  int la = 0, lb = 0;
  if (!pair) {
    a = slot(snapshot);
  } else {
    // level 0 of pair k; level 1 of pair k-1.  The last snapshot also fills
    // level 1 of its own pair so a clamped lookup at the end stays exact.
    a = slot(snapshot);
    la = 0;
    if (snapshot > 0) {
      b = slot(snapshot - 1);
      lb = 1;
    }
    if (snapshot == field->n_snapshots - 1) {
      if (b == nullptr) {
        b = slot(snapshot);
        lb = 1;
      } else {
        // write level 1 of its own slot in a second launch
      }
    }
  }
  const int gx = field->n[0] + (field->periodic[0] ? 1 : 0);
  const int gy = field->n[1] + (field->periodic[1] ? 1 : 0);
  const int gz = nz + (dim == 3 && field->periodic[2] ? 1 : 0);
  const int64_t nodes = (int64_t)gx * gy * gz;
  hipStream_t st = as_stream(stream);
  const int cells = field->n[0] * field->n[1] * nz;  // < 2^30 (validate_field)
  const int rec_floats = (1 << dim) * dim * (pair ? 2 : 1);
  auto launch = [&](char *sa, int lva, char *sb, int lvb) {
    if (layout_cell(field->layout)) {
      const int px = field->periodic[0], py = field->periodic[1], pz = dim == 3 ? field->periodic[2] : 0;
#define PYFTLE_PACK_CELL(D, S)                                                                        \
  pack_cell_kernel<D, S><<<blocks_for(cells, 256), 256, 0, st>>>(                                    \
      sa, lva, sb, lvb, (const S *)src, src_strides[0], src_strides[1], src_strides[2], src_strides[3], \
      field->n[0], field->n[1], nz, px, py, pz, rec_floats)
      if (dim == 2 && src_dtype == PYFTLE_F64) PYFTLE_PACK_CELL(2, double);
      else if (dim == 2) PYFTLE_PACK_CELL(2, float);
      else if (src_dtype == PYFTLE_F64) PYFTLE_PACK_CELL(3, double);
      else PYFTLE_PACK_CELL(3, float);
#undef PYFTLE_PACK_CELL
    } else if (field->dtype == PYFTLE_F32) {
      if (src_dtype == PYFTLE_F64)
        pack_kernel<float, double><<<blocks_for(nodes, 256), 256, 0, st>>>(
            sa, lva, sb, lvb, (const double *)src, src_strides[0], src_strides[1], src_strides[2],
            src_strides[3], dim, field->n[0], field->n[1], nz, gx, gy, gz, per_level, comps);
      else
        pack_kernel<float, float><<<blocks_for(nodes, 256), 256, 0, st>>>(
            sa, lva, sb, lvb, (const float *)src, src_strides[0], src_strides[1], src_strides[2],
            src_strides[3], dim, field->n[0], field->n[1], nz, gx, gy, gz, per_level, comps);
    } else {
      if (src_dtype == PYFTLE_F64)
        pack_kernel<double, double><<<blocks_for(nodes, 256), 256, 0, st>>>(
            sa, lva, sb, lvb, (const double *)src, src_strides[0], src_strides[1], src_strides[2],
            src_strides[3], dim, field->n[0], field->n[1], nz, gx, gy, gz, per_level, comps);
      else
        pack_kernel<double, float><<<blocks_for(nodes, 256), 256, 0, st>>>(
            sa, lva, sb, lvb, (const float *)src, src_strides[0], src_strides[1], src_strides[2],
            src_strides[3], dim, field->n[0], field->n[1], nz, gx, gy, gz, per_level, comps);
    }
  };
  launch(a, la, b, lb);
  if (pair && snapshot == field->n_snapshots - 1 && snapshot > 0) launch(slot(snapshot), 1, nullptr, 0);
  return check_launch("pack_snapshot");
}

int pyftle_field_sample(const pyftle_field *field, int32_t time_mode, int32_t snapshot, double t,
                        const void *points, int32_t pts_dtype, int64_t n, void *out, void *stream) {
  FieldDev d;
  int rc = make_field_dev(field, &d);
  if (rc) return rc;
  if (!points || !out) return fail("field_sample: NULL pointer");
  if (pts_dtype != PYFTLE_F32 && pts_dtype != PYFTLE_F64) return fail("field_sample: bad dtype");
  if (n < 0) return fail("field_sample: negative n");
  if (n == 0) return PYFTLE_OK;
  int k = snapshot;
  double w = 0.0;
  const bool pair = layout_pair(field->layout);
  if (time_mode == PYFTLE_TIME_LINEAR) {
    if (!pair) return fail("field_sample: LINEAR time needs the PAIR layout");
    const double s = (t - field->t0) / field->dt_snap;
    double fl = floor(s);
    int kk = (fl > -1e9 && fl < 1e9) ? (int)fl : 0;
    const int kmax = field->n_snapshots >= 2 ? field->n_snapshots - 2 : 0;
    kk = kk < 0 ? 0 : (kk > kmax ? kmax : kk);
    double ww = s - kk;
    ww = ww < 0.0 ? 0.0 : (ww > 1.0 ? 1.0 : ww);
    if (field->n_snapshots < 2) ww = 0.0;
    k = kk;
    w = ww;
  } else if (time_mode != PYFTLE_TIME_STEADY && time_mode != PYFTLE_TIME_FROZEN) {
    return fail("field_sample: bad time mode");
  }
  int s = k % field->ring_slots;
  if (s < 0) s += field->ring_slots;
  const char *slot = reinterpret_cast<const char *>(field->data) + (int64_t)s * field->slot_stride;
  hipStream_t st = as_stream(stream);
  const bool f64 = field->dtype == PYFTLE_F64;
  if (layout_cell(field->layout)) {  // FP32 GridInterp fields only (checked when created)
    if (field->dim == 2) {
      if (pair) launch_sample_cell<2, true>(d, slot, w, points, pts_dtype, n, out, st);
      else launch_sample_cell<2, false>(d, slot, w, points, pts_dtype, n, out, st);
    } else {
      if (pair) launch_sample_cell<3, true>(d, slot, w, points, pts_dtype, n, out, st);
      else launch_sample_cell<3, false>(d, slot, w, points, pts_dtype, n, out, st);
    }
    return check_launch("field_sample");
  }
  if (field->interp != PYFTLE_INTERP_GINTERP) {
    if (field->dim == 2) launch_sample_rgi<2>(d, f64, slot, points, pts_dtype, n, out, st);
    else launch_sample_rgi<3>(d, f64, slot, points, pts_dtype, n, out, st);
    return check_launch("field_sample");
  }
  if (field->dim == 2) {
    if (pair) {
      if (f64) launch_sample<2, true, double>(d, slot, w, points, pts_dtype, n, out, st);
      else launch_sample<2, true, float>(d, slot, w, points, pts_dtype, n, out, st);
    } else {
      if (f64) launch_sample<2, false, double>(d, slot, w, points, pts_dtype, n, out, st);
      else launch_sample<2, false, float>(d, slot, w, points, pts_dtype, n, out, st);
    }
  } else {
    if (pair) {
      if (f64) launch_sample<3, true, double>(d, slot, w, points, pts_dtype, n, out, st);
      else launch_sample<3, true, float>(d, slot, w, points, pts_dtype, n, out, st);
    } else {
      if (f64) launch_sample<3, false, double>(d, slot, w, points, pts_dtype, n, out, st);
      else launch_sample<3, false, float>(d, slot, w, points, pts_dtype, n, out, st);
    }
  }
  return check_launch("field_sample");
}

}  // extern "C"
