// SciPy RegularGridInterpolator semantics on the device (interpolate.py:546-762:
// the reference's GridInterpolator builds one RegularGridInterpolator per
// velocity component with bounds_error=False, fill_value=0 on the axes
// numpy.linspace(min, max, n)).
//
//   * axis search: interval l in [0, n-2] with node(l) <= x < node(l+1), the
//     last node belonging to the last interval, and the normalised distance
//     (x - node(l)) / (node(l+1) - node(l));
//   * x < node(0) or x > node(n-1) on any axis -> 0 (fill_value); NaN -> NaN;
//   * linear: multilinear blend of the 2^dim corners;
//   * nearest: node l if the distance is <= 0.5, else l + 1, per axis;
//   * cubic: the tensor-product not-a-knot cubic spline, held as uniform cubic
//     B-spline coefficients c[-1..n] per axis (pyftle_pack_spline), evaluated
//     with the 4 B-spline weights per axis.
// Arithmetic is FP64 for every record dtype.
#pragma once

#include "common.hpp"

namespace pyftle {

// numpy.linspace(origin, upper, n)[k]: k*step + start, the last element is `upper`
__device__ __forceinline__ double rgi_node(const FieldDev &f, int a, int k) {
  return (k == f.n[a] - 1) ? f.upper[a] : (double)k * f.spacing[a] + f.origin[a];
}

struct RgiAxis {
  int i;     // interval
  double d;  // normalised distance in the interval
};

__device__ __forceinline__ RgiAxis rgi_axis(const FieldDev &f, int a, double x) {
  const int n = f.n[a];
  const double q = (x - f.origin[a]) * f.inv_spacing[a];
  // the estimate is within one interval of the exact search; NaN lands on 0
  int l = (q > 0.0) ? (q < (double)(n - 2) ? (int)q : n - 2) : 0;
  if (l > 0 && x < rgi_node(f, a, l)) --l;
  if (l > 0 && x < rgi_node(f, a, l)) --l;
  if (l < n - 2 && x >= rgi_node(f, a, l + 1)) ++l;
  if (l < n - 2 && x >= rgi_node(f, a, l + 1)) ++l;
  const double x0 = rgi_node(f, a, l);
  const double x1 = rgi_node(f, a, l + 1);
  RgiAxis r;
  r.i = l;
  r.d = (x - x0) / (x1 - x0);
  return r;
}

template <typename T>
struct RgiRec;
template <>
struct RgiRec<float> {
  template <int DIM>
  static __device__ __forceinline__ void load(const char *slot, int64_t rec, double *v) {
    if constexpr (DIM == 2) {
      const float2 r = reinterpret_cast<const float2 *>(slot)[rec];
      v[0] = r.x;
      v[1] = r.y;
    } else {
      const float4 r = reinterpret_cast<const float4 *>(slot)[rec];
      v[0] = r.x;
      v[1] = r.y;
      v[2] = r.z;
    }
  }
};
template <>
struct RgiRec<double> {
  template <int DIM>
  static __device__ __forceinline__ void load(const char *slot, int64_t rec, double *v) {
    const double2 *p = reinterpret_cast<const double2 *>(slot) + rec * (DIM == 2 ? 1 : 2);
    const double2 a = p[0];
    v[0] = a.x;
    v[1] = a.y;
    if constexpr (DIM == 3) v[2] = p[1].x;
  }
};

__device__ __forceinline__ void bspline3_weights(double t, double *w) {
  const double t2 = t * t, t3 = t2 * t, s = 1.0 - t;
  w[0] = s * s * s / 6.0;
  w[1] = (3.0 * t3 - 6.0 * t2 + 4.0) / 6.0;
  w[2] = (-3.0 * t3 + 3.0 * t2 + 3.0 * t + 1.0) / 6.0;
  w[3] = t3 / 6.0;
}

// Velocity at x (physical, FP64) from one slot of an RGI-mode field.
template <int DIM, int INTERP, typename T>
__device__ __forceinline__ void sample_rgi(const FieldDev &f, const char *slot, const double *x,
                                           double *vel) {
  bool inside = true, nan = false;
  RgiAxis ax[DIM];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    nan = nan || (x[a] != x[a]);
    inside = inside && !(x[a] < f.origin[a]) && !(x[a] > f.upper[a]);
    ax[a] = rgi_axis(f, a, x[a]);
  }
  double out[DIM];
  if constexpr (INTERP == PYFTLE_INTERP_RGI_NEAREST) {
    int64_t rec = 0;
    int64_t pitch = 1;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      const int k = ax[a].d <= 0.5 ? ax[a].i : ax[a].i + 1;
      rec += (int64_t)k * pitch;
      pitch *= (a == 0) ? f.px : (int64_t)f.pxy / f.px;
    }
    RgiRec<T>::template load<DIM>(slot, rec, out);
  } else if constexpr (INTERP == PYFTLE_INTERP_RGI_LINEAR) {
    if constexpr (DIM == 2) {
      const int i0 = ax[0].i, i1 = ax[1].i;
      const double y0 = ax[0].d, y1 = ax[1].d;
      double c00[2], c01[2], c10[2], c11[2];
      const int64_t px = f.px;
      RgiRec<T>::template load<2>(slot, (int64_t)i1 * px + i0, c00);
      RgiRec<T>::template load<2>(slot, (int64_t)(i1 + 1) * px + i0, c01);
      RgiRec<T>::template load<2>(slot, (int64_t)i1 * px + i0 + 1, c10);
      RgiRec<T>::template load<2>(slot, (int64_t)(i1 + 1) * px + i0 + 1, c11);
#pragma unroll
      for (int c = 0; c < 2; ++c)
        out[c] = c00[c] * (1.0 - y0) * (1.0 - y1) + c01[c] * (1.0 - y0) * y1 +
                 c10[c] * y0 * (1.0 - y1) + c11[c] * y0 * y1;
    } else {
      // corners in itertools.product order (x slowest), weight ((1*wx)*wy)*wz
#pragma unroll
      for (int c = 0; c < DIM; ++c) out[c] = 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int bx = (q >> 2) & 1, by = (q >> 1) & 1, bz = q & 1;
        const double wx = bx ? ax[0].d : 1.0 - ax[0].d;
        const double wy = by ? ax[1].d : 1.0 - ax[1].d;
        const double wz = bz ? ax[2 % DIM].d : 1.0 - ax[2 % DIM].d;
        const double w = ((1.0 * wx) * wy) * wz;
        const int64_t rec = (int64_t)(ax[2 % DIM].i + bz) * f.pxy + (int64_t)(ax[1].i + by) * f.px +
                            ax[0].i + bx;
        double v[3];
        RgiRec<T>::template load<3>(slot, rec, v);
#pragma unroll
        for (int c = 0; c < DIM; ++c) out[c] = out[c] + v[c] * w;
      }
    }
  } else {
    double wx[4], wy[4], wz[4];
    bspline3_weights(ax[0].d, wx);
    bspline3_weights(ax[1].d, wy);
    if constexpr (DIM == 3) bspline3_weights(ax[2 % DIM].d, wz);
#pragma unroll
    for (int c = 0; c < DIM; ++c) out[c] = 0.0;
    // coefficient c_{i-1+p} of an axis sits at record index i + p
    if constexpr (DIM == 2) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double row[2] = {0.0, 0.0};
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          double v[2];
          RgiRec<T>::template load<2>(slot, (int64_t)(ax[1].i + q) * f.px + ax[0].i + p, v);
          row[0] += wx[p] * v[0];
          row[1] += wx[p] * v[1];
        }
        out[0] += wy[q] * row[0];
        out[1] += wy[q] * row[1];
      }
    } else {
#pragma unroll 1
      for (int r = 0; r < 4; ++r) {
        double plane[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double row[3] = {0.0, 0.0, 0.0};
          const int64_t base = (int64_t)(ax[2 % DIM].i + r) * f.pxy + (int64_t)(ax[1].i + q) * f.px + ax[0].i;
#pragma unroll
          for (int p = 0; p < 4; ++p) {
            double v[3];
            RgiRec<T>::template load<3>(slot, base + p, v);
#pragma unroll
            for (int c = 0; c < 3; ++c) row[c] += wx[p] * v[c];
          }
#pragma unroll
          for (int c = 0; c < 3; ++c) plane[c] += wy[q] * row[c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c % DIM] += wz[r] * plane[c];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < DIM; ++c) vel[c] = nan ? __longlong_as_double(0x7ff8000000000000LL) : (inside ? out[c] : 0.0);
}

}  // namespace pyftle
