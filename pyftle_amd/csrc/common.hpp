// Shared device code of libpyftle_hip: field sampling (bi/trilinear in space,
// linear in time), FP64 FTLE math and host-side error plumbing.
//
// Built with -ffp-contract=off: the FP64 paths follow the reference's
// evaluation order exactly (GridInterp.cpp:73-84,136-137; integrate.py:81-109),
// while the FP32 paths request fused multiply-adds explicitly through fmaf.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "pyftle_hip.h"

namespace pyftle {

// ---------------------------------------------------------------------------
// host-side error state
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
int fail(const std::string &msg, int code = PYFTLE_EINVAL);
int check_launch(const char *what);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// compute units of the current device (cached per device; 256 if the query fails)
inline int device_cu_count() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

// ---------------------------------------------------------------------------
// device view of a pyftle_field (passed by value as a kernel argument)
// ---------------------------------------------------------------------------
struct FieldDev {
  const char *data;
  int64_t slot_stride;
  int ring;
  int nsnap;
  int n[3];
  int periodic[3];
  int px, pxy;     // record pitch of a row / an xy-slice (ghost node on periodic axes)
  int per_n[3];    // n on periodic axes, 0 otherwise
  int lim[3];      // valid lower-index bound: n (periodic) or n - 1 (zero outside)
  double origin[3];
  double spacing[3];
  double inv_spacing[3];
  double n_d[3];
  double inv_n_d[3];
  float origin_f[3];
  float inv_spacing_f[3];
  float n_f[3];
  float inv_n_f[3];
  double t0, dts;
  int interp;       // PYFTLE_INTERP_*
  double upper[3];  // last grid node per axis (RGI modes)
  // scattered-data modes (pyftle_scatter)
  const double *points;
  const int *simplices;
  const double *transform;
  const int *hull;
  const int *s_off, *s_ids;
  const int *p_off, *p_ids;
  const double *ct_g;
  int64_t npts;
  int s_dims[3], p_dims[3];
  double s_org[3], s_inv[3], p_org[3], p_size[3], p_inv[3];
};

inline bool is_scatter(int interp) {
  return interp == PYFTLE_INTERP_SCATTER_LINEAR || interp == PYFTLE_INTERP_SCATTER_NEAREST ||
         interp == PYFTLE_INTERP_SCATTER_CUBIC;
}

int make_field_dev(const pyftle_field *f, FieldDev *out);
int validate_field(const pyftle_field *f);

// two time levels per record (PAIR, CELL_PAIR); per-cell records (CELL, CELL_PAIR)
inline bool layout_pair(int layout) { return (layout & 1) != 0; }
inline bool layout_cell(int layout) { return (layout & 2) != 0; }

// records per axis of one slot: n, plus the ghost node of a periodic axis, or
// n + 2 spline coefficients (-1..n) in the cubic mode, or n cells
inline int axis_records(const pyftle_field *f, int a) {
  if (f->interp == PYFTLE_INTERP_RGI_CUBIC) return f->n[a] + 2;
  if (layout_cell(f->layout)) return f->n[a];
  return f->n[a] + (f->periodic[a] ? 1 : 0);
}

// records of one slot
inline int64_t padded_nodes(const pyftle_field *f) {
  if (is_scatter(f->interp)) return f->scatter ? f->scatter->n_points : 0;
  int64_t g = 1;
  for (int a = 0; a < f->dim; ++a) g *= (int64_t)axis_records(f, a);
  return g;
}

// record size in bytes for (dim, layout, dtype); the scattered cubic mode stores 8
// components per point (value and gradient of u and v)
inline int record_bytes(int dim, int layout, int dtype, int interp = PYFTLE_INTERP_GINTERP) {
  if (interp == PYFTLE_INTERP_SCATTER_CUBIC) return 8 * (dtype == PYFTLE_F64 ? 8 : 4);
  const int levels = layout_pair(layout) ? 2 : 1;
  const int comps = layout_cell(layout) ? (1 << dim) * dim * levels : (dim == 2 ? 2 : 4) * levels;
  return comps * (dtype == PYFTLE_F64 ? 8 : 4);
}

// ---------------------------------------------------------------------------
// vector record types
// ---------------------------------------------------------------------------
template <typename Real, int N>
struct alignas(sizeof(Real) * N) RecN {
  Real v[N];
};

// number of Real components in a record
template <int DIM, bool PAIR>
struct RecInfo {
  static constexpr int per_level = (DIM == 2 ? 2 : 4);
  static constexpr int comps = per_level * (PAIR ? 2 : 1);
};

// ---------------------------------------------------------------------------
// Index computation for one axis.
//   FP64: (p - origin) / spacing, floor, reference bounds test
//         (grid_interp.py:78-79, GridInterp.cpp:44-58,116-126).
//   FP32: (p - origin) * (1/spacing).
// periodic axes wrap the lower index modulo n (extension of this engine).
// ---------------------------------------------------------------------------
template <typename Real>
struct AxisIdx {
  int i0, i1;
  Real d;
  bool ok;
};

// FP64, reference semantics: division, integer modulo on periodic axes.
__device__ __forceinline__ AxisIdx<double> axis_index_exact(const FieldDev &f, int a, double p) {
  AxisIdx<double> r;
  const double fi = (p - f.origin[a]) / f.spacing[a];
  const double fl = floor(fi);
  const int n = f.n[a];
  // A value outside int range (or NaN) is out of bounds, as the x86
  // static_cast<int> result INT_MIN is in the reference.
  const bool finite_in_range = (fl > -2147483648.0) && (fl < 2147483647.0);
  int i0 = finite_in_range ? (int)fl : INT32_MIN;
  r.d = fi - (double)i0;
  if (f.periodic[a]) {
    r.ok = finite_in_range;
    int w = i0 % n;
    if (w < 0) w += n;
    i0 = finite_in_range ? w : 0;
    r.i0 = i0;
    r.i1 = (i0 + 1 == n) ? 0 : i0 + 1;
  } else {
    r.ok = finite_in_range && (i0 >= 0) && (i0 + 1 < n);
    r.i0 = r.ok ? i0 : 0;
    r.i1 = r.ok ? i0 + 1 : 0;
  }
  return r;
}

// FP64 positions of an FP32 field: multiply by the reciprocal spacing and wrap
// periodic axes in floating point (fl - n*floor(fl/n) is exact for integral fl).
__device__ __forceinline__ AxisIdx<double> axis_index(const FieldDev &f, int a, double p) {
  AxisIdx<double> r;
  const double fi = (p - f.origin[a]) * f.inv_spacing[a];
  const double fl = floor(fi);
  r.d = fi - fl;
  const bool finite_in_range = (fl > -1.0e15) && (fl < 1.0e15);
  const int n = f.n[a];
  if (f.periodic[a]) {
    const double q = floor(fl * f.inv_n_d[a]);
    int w = finite_in_range ? (int)fma(-q, f.n_d[a], fl) : 0;
    w = (w >= n) ? w - n : w;
    w = (w < 0) ? w + n : w;
    w = w < 0 ? 0 : (w >= n ? n - 1 : w);
    r.ok = finite_in_range;
    r.i0 = w;
    r.i1 = (w + 1 == n) ? 0 : w + 1;
  } else {
    const double cl = fl < -1.0 ? -1.0 : (fl > 2147483000.0 ? 2147483000.0 : fl);
    const int i0 = finite_in_range ? (int)cl : -1;
    r.ok = finite_in_range && (i0 >= 0) && (i0 + 1 < n);
    r.i0 = r.ok ? i0 : 0;
    r.i1 = r.ok ? i0 + 1 : 0;
  }
  return r;
}

__device__ __forceinline__ AxisIdx<float> axis_index(const FieldDev &f, int a, float p) {
  AxisIdx<float> r;
  const float fi = (p - f.origin_f[a]) * f.inv_spacing_f[a];
  const float fl = floorf(fi);
  r.d = fi - fl;
  const bool finite_in_range = (fl > -2147483520.0f) && (fl < 2147483520.0f);
  const int n = f.n[a];
  if (f.periodic[a]) {
    // fl - n*floor(fl/n) is exact for |fl| < 2^24; one correction step fixes
    // the rounding of fl/n at multiples of n.
    const float q = floorf(fl * f.inv_n_f[a]);
    int w = (int)fmaf(-q, f.n_f[a], fl);
    w = (w >= n) ? w - n : w;
    w = (w < 0) ? w + n : w;
    // beyond 2^24 the reduction is inexact: keep the index in range anyway
    w = w < 0 ? 0 : (w >= n ? n - 1 : w);
    r.ok = finite_in_range;
    r.i0 = finite_in_range ? w : 0;
    r.i1 = (r.i0 + 1 == n) ? 0 : r.i0 + 1;
  } else {
    const int i0 = finite_in_range ? (int)fl : -1;
    r.ok = finite_in_range && (i0 >= 0) && (i0 + 1 < n);
    r.i0 = r.ok ? i0 : 0;
    r.i1 = r.ok ? i0 + 1 : 0;
  }
  return r;
}

template <typename Real>
__device__ __forceinline__ Real lerp_f(Real a, Real b, Real t);
template <>
__device__ __forceinline__ float lerp_f<float>(float a, float b, float t) {
  return fmaf(t, b - a, a);
}
template <>
__device__ __forceinline__ double lerp_f<double>(double a, double b, double t) {
  return fma(t, b - a, a);
}

// ---------------------------------------------------------------------------
// Sample the velocity at x from one slot.  w = time weight (PAIR layout).
// FP64: per-level reference formula, then (1-w)*a + w*b.
// FP32: lerps with fma, level blend per corner-interpolated value.
// ---------------------------------------------------------------------------
template <int DIM, bool PAIR>
__device__ __forceinline__ void sample_field(const FieldDev &f, const char *slot, double w,
                                             const double *x, double *vel) {
  constexpr int C = RecInfo<DIM, PAIR>::comps;
  constexpr int L = RecInfo<DIM, PAIR>::per_level;
  using Rec = RecN<double, (C > 4 ? 4 : C)>;  // loads in chunks of <= 32 bytes
  const Rec *rec = reinterpret_cast<const Rec *>(slot);
  constexpr int chunks = (C > 4) ? 2 : 1;
  AxisIdx<double> ax[DIM];
  bool ok = true;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    ax[a] = axis_index_exact(f, a, x[a]);
    ok = ok && ax[a].ok;
  }
  if constexpr (DIM == 2) {
    const int64_t r00 = (int64_t)ax[1].i0 * f.px + ax[0].i0;
    const int64_t r10 = (int64_t)ax[1].i0 * f.px + ax[0].i1;
    const int64_t r01 = (int64_t)ax[1].i1 * f.px + ax[0].i0;
    const int64_t r11 = (int64_t)ax[1].i1 * f.px + ax[0].i1;
    double c[4][C];
    const int64_t idx[4] = {r00, r10, r01, r11};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int h = 0; h < chunks; ++h) {
        Rec r = rec[idx[q] * chunks + h];
#pragma unroll
        for (int e = 0; e < (C > 4 ? 4 : C); ++e) c[q][h * 4 + e] = r.v[e];
      }
    const double xd = ax[0].d, yd = ax[1].d;
    double lev[2][2];
#pragma unroll
    for (int l = 0; l < (PAIR ? 2 : 1); ++l)
#pragma unroll
      for (int comp = 0; comp < 2; ++comp) {
        const int e = l * L + comp;
        // GridInterp.cpp:136-137
        lev[l][comp] = c[0][e] * (1.0 - xd) * (1.0 - yd) + c[1][e] * xd * (1.0 - yd) +
                       c[2][e] * (1.0 - xd) * yd + c[3][e] * xd * yd;
      }
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
      double v = lev[0][comp];
      if (PAIR) v = (1.0 - w) * lev[0][comp] + w * lev[1][comp];
      vel[comp] = ok ? v : 0.0;
    }
  } else {
    const int64_t nx = f.px, nxy = f.pxy;
    double lev[2][3];
    const double xd = ax[0].d, yd = ax[1].d, zd = ax[2].d;
    // corners: (x0/x1, y0/y1, z0/z1)
    double c[8][C];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int ix = (q & 1) ? ax[0].i1 : ax[0].i0;
      const int iy = (q & 2) ? ax[1].i1 : ax[1].i0;
      const int iz = (q & 4) ? ax[2].i1 : ax[2].i0;
      const int64_t node = (int64_t)iz * nxy + (int64_t)iy * nx + ix;
#pragma unroll
      for (int h = 0; h < chunks; ++h) {
        Rec r = rec[node * chunks + h];
#pragma unroll
        for (int e = 0; e < (C > 4 ? 4 : C); ++e) c[q][h * 4 + e] = r.v[e];
      }
    }
#pragma unroll
    for (int l = 0; l < (PAIR ? 2 : 1); ++l)
#pragma unroll
      for (int comp = 0; comp < 3; ++comp) {
        const int e = l * L + comp;
        // GridInterp.cpp:73-84 (x, then y, then z)
        const double c00 = c[0][e] * (1.0 - xd) + c[1][e] * xd;
        const double c01 = c[4][e] * (1.0 - xd) + c[5][e] * xd;
        const double c10 = c[2][e] * (1.0 - xd) + c[3][e] * xd;
        const double c11 = c[6][e] * (1.0 - xd) + c[7][e] * xd;
        const double c0 = c00 * (1.0 - yd) + c10 * yd;
        const double c1 = c01 * (1.0 - yd) + c11 * yd;
        lev[l][comp] = c0 * (1.0 - zd) + c1 * zd;
      }
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
      double v = lev[0][comp];
      if (PAIR) v = (1.0 - w) * lev[0][comp] + w * lev[1][comp];
      vel[comp] = ok ? v : 0.0;
    }
  }
}

// FP32-stored field, any position / blend precision:
//   Pos   = precision of the position, index and fractional weights;
//   Blend = precision of the corner blend (lerps with fused multiply-add).
// CELL: per-cell records (PYFTLE_LAYOUT_CELL, CELL_PAIR), the cell's corners
// [level][corner][component] in one record: the same corner values as the node
// records (upper corner wrapped on periodic axes), so the same result.
template <int DIM, bool PAIR, typename Pos, typename Blend, bool CELL = false>
__device__ __forceinline__ void sample_field_f32(const FieldDev &f, const char *slot, Blend w,
                                                 const Pos *x, Blend *vel) {
  constexpr int C = RecInfo<DIM, PAIR>::comps;
  constexpr int L = RecInfo<DIM, PAIR>::per_level;
  using Rec = RecN<float, (C > 4 ? 4 : C)>;  // loads of <= 16 bytes
  constexpr int chunks = (C > 4) ? 2 : 1;
  const Rec *rec = reinterpret_cast<const Rec *>(slot);
  AxisIdx<Pos> ax[DIM];
  bool ok = true;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    ax[a] = axis_index(f, a, x[a]);
    ok = ok && ax[a].ok;
  }
  constexpr int NQ = (DIM == 2) ? 4 : 8;
  float c[NQ][C];
  const int nx = f.px;
  const int nxy = f.pxy;
  if constexpr (CELL) {
    constexpr int LV = PAIR ? 2 : 1;
    constexpr int RF = NQ * DIM * LV;  // 8, 16, 24 or 48 floats
    int cell = ax[1].i0 * nx + ax[0].i0;
    if constexpr (DIM == 3) cell += ax[2 % DIM].i0 * nxy;
    const float4 *rp = reinterpret_cast<const float4 *>(slot) + (int64_t)cell * (RF / 4);
    float r[RF];
#pragma unroll
    for (int h = 0; h < RF / 4; ++h) {
      const float4 v = rp[h];
      r[4 * h] = v.x;
      r[4 * h + 1] = v.y;
      r[4 * h + 2] = v.z;
      r[4 * h + 3] = v.w;
    }
#pragma unroll
    for (int l = 0; l < LV; ++l)
#pragma unroll
      for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int comp = 0; comp < DIM; ++comp) c[q][l * L + comp] = r[(l * NQ + q) * DIM + comp];
  } else {
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int ix = (q & 1) ? ax[0].i1 : ax[0].i0;
    const int iy = (q & 2) ? ax[1].i1 : ax[1].i0;
    int node = iy * nx + ix;
    if constexpr (DIM == 3) node += ((q & 4) ? ax[2 % DIM].i1 : ax[2 % DIM].i0) * nxy;
#pragma unroll
    for (int h = 0; h < chunks; ++h) {
      Rec r = rec[node * chunks + h];
#pragma unroll
      for (int e = 0; e < (C > 4 ? 4 : C); ++e) c[q][h * 4 + e] = r.v[e];
    }
  }
  }
  const Blend xd = (Blend)ax[0].d, yd = (Blend)ax[1].d;
  const Blend zd = (Blend)ax[2 % DIM].d;
  Blend val[L];
#pragma unroll
  for (int comp = 0; comp < DIM; ++comp) {
    Blend lev[2];
#pragma unroll
    for (int l = 0; l < (PAIR ? 2 : 1); ++l) {
      const int e = l * L + comp;
      const Blend c0 = lerp_f<Blend>(c[0][e], c[1][e], xd);
      const Blend c1 = lerp_f<Blend>(c[2][e], c[3][e], xd);
      Blend r = lerp_f<Blend>(c0, c1, yd);
      if constexpr (DIM == 3) {
        const Blend c2 = lerp_f<Blend>(c[4 % NQ][e], c[5 % NQ][e], xd);
        const Blend c3 = lerp_f<Blend>(c[6 % NQ][e], c[7 % NQ][e], xd);
        const Blend r1 = lerp_f<Blend>(c2, c3, yd);
        r = lerp_f<Blend>(r, r1, zd);
      }
      lev[l] = r;
    }
    Blend v = lev[0];
    if (PAIR) v = lerp_f<Blend>(lev[0], lev[1 % (PAIR ? 2 : 1)], w);
    val[comp] = v;
  }
#pragma unroll
  for (int comp = 0; comp < DIM; ++comp) vel[comp] = ok ? val[comp] : (Blend)0;
}

template <int DIM, bool PAIR>
__device__ __forceinline__ void sample_field(const FieldDev &f, const char *slot, float w,
                                             const float *x, float *vel) {
  sample_field_f32<DIM, PAIR, float, float>(f, slot, w, x, vel);
}

// ---------------------------------------------------------------------------
// Time lookup (uniform per launch stage): slot pointer and blend weight.
// ---------------------------------------------------------------------------
__device__ __forceinline__ const char *slot_ptr(const FieldDev &f, int k) {
  int s = k % f.ring;
  if (s < 0) s += f.ring;
  return f.data + (int64_t)s * f.slot_stride;
}

__device__ __forceinline__ void time_linear_lookup(const FieldDev &f, double t, int &k,
                                                   double &w) {
  const double s = (t - f.t0) / f.dts;
  double fl = floor(s);
  int kk = (fl > -1e9 && fl < 1e9) ? (int)fl : 0;
  const int kmax = f.nsnap >= 2 ? f.nsnap - 2 : 0;
  kk = kk < 0 ? 0 : (kk > kmax ? kmax : kk);
  double ww = s - (double)kk;
  ww = ww < 0.0 ? 0.0 : (ww > 1.0 ? 1.0 : ww);
  if (f.nsnap < 2) ww = 0.0;
  k = kk;
  w = ww;
}

// ---------------------------------------------------------------------------
// FP64 FTLE math (ftle.py:9-233), symmetric closed forms.
// ---------------------------------------------------------------------------

// 2x2 largest eigenvalue of a general matrix [[A,B],[Cc,D]]:
// 0.5*(tr + sqrt((A-D)^2 + 4 B Cc)).  Equals the reference's
// 0.5*(tr + sqrt(tr^2 - 4 det)) (ftle.py:155-161) algebraically; for a
// symmetric C the radicand is a sum of squares and never rounds below zero.
__device__ __forceinline__ double max_eig_2x2(double A, double B, double Cc, double D) {
  const double tr = A + D;
  const double dm = A - D;
  const double disc = dm * dm + 4.0 * (B * Cc);
  return 0.5 * (tr + sqrt(disc));
}

// Largest eigenvalue of a symmetric 3x3 matrix by the trigonometric
// (Smith 1961) method.
__device__ __forceinline__ double max_eig_sym3(double a00, double a01, double a02, double a11,
                                               double a12, double a22) {
  const double p1 = a01 * a01 + a02 * a02 + a12 * a12;
  const double q = (a00 + a11 + a22) / 3.0;
  const double b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
  const double p2 = b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * p1;
  if (!(p2 > 0.0)) {
    // already a multiple of the identity (or NaN input)
    return fmax(fmax(a00, a11), a22);
  }
  const double p = sqrt(p2 / 6.0);
  const double ip = 1.0 / p;
  const double c00 = b00 * ip, c11 = b11 * ip, c22 = b22 * ip;
  const double c01 = a01 * ip, c02 = a02 * ip, c12 = a12 * ip;
  const double detb = c00 * (c11 * c22 - c12 * c12) - c01 * (c01 * c22 - c12 * c02) +
                      c02 * (c01 * c12 - c11 * c02);
  double r = 0.5 * detb;
  r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
  const double phi = acos(r) / 3.0;
  return q + 2.0 * p * cos(phi);
}

// FTLE of one seed from the Jacobian F (row-major dim x dim):
// 2D C = F^T F (ftle.py:62-67), 3D C = F F^T (ftle.py:96), then
// (1/T) * log(sqrt(lambda_max)) (ftle.py:198,233).
template <int DIM>
__device__ __forceinline__ double ftle_of_jacobian(const double *F, double T) {
  double lam;
  if constexpr (DIM == 2) {
    const double A = F[0] * F[0] + F[2] * F[2];
    const double B = F[0] * F[1] + F[2] * F[3];
    const double D = F[1] * F[1] + F[3] * F[3];
    lam = max_eig_2x2(A, B, B, D);
  } else {
    double C[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        C[i * 3 + j] = F[i * 3 + 0] * F[j * 3 + 0] + F[i * 3 + 1] * F[j * 3 + 1] +
                       F[i * 3 + 2] * F[j * 3 + 2];
    lam = max_eig_sym3(C[0], C[1], C[2], C[4], C[5], C[8]);
  }
  return (1.0 / T) * log(sqrt(lam));
}

// Jacobian column c of a seed from its block-ordered neighbours
// (cauchy_green.py:70-74,140-151): pairs (1,0), (2,3), (4,5).
__host__ __device__ constexpr int pair_plus(int c) { return c == 0 ? 1 : 2 * c; }
__host__ __device__ constexpr int pair_minus(int c) { return c == 0 ? 0 : 2 * c + 1; }

// ---------------------------------------------------------------------------
// XCD-aware block remap: consecutive logical blocks land on one XCD
// (hardware dispatches block b to XCD b % 8).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t logical_block(int remap) {
  const int64_t b = blockIdx.x;
  if (!remap) return b;
  const int64_t nb = gridDim.x;
  const int64_t q = nb / 8, r = nb % 8;
  const int64_t xcd = b % 8, i = b / 8;
  return (xcd < r) ? xcd * (q + 1) + i : r * (q + 1) + (xcd - r) * q + i;
}

}  // namespace pyftle
