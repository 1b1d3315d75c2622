// Scattered-data strategies on the device (interpolate.py:188-294): SciPy's
// LinearNDInterpolator on a Delaunay triangulation and NearestNDInterpolator.
//
// LINEAR: the simplex holding x is found among those listed in x's bucket
// (every simplex is listed in each bucket its bounding box touches); a simplex
// holds x when all barycentric coordinates lie in [-eps, 1 + eps] with
// eps = 100 DBL_EPSILON, computed as SciPy does from Delaunay.transform:
//   c_i = sum_j T[i][j] (x_j - r_j),  c_dim = 1 - sum_i c_i.
// Failing that, a simplex with a facet on the convex hull is accepted within
// eps_broad = sqrt(DBL_EPSILON), as SciPy accepts points just outside the hull.
// The value is 0 + sum_k c_k v[vertex_k] (SciPy's accumulation order); no
// simplex -> NaN (LinearNDInterpolator's default fill value).
//
// NEAREST: rings of point buckets around x's bucket are searched until the
// best squared distance is no larger than the squared distance from x to any
// bucket not yet searched; ties go to the lower point index.
#pragma once

#include "common.hpp"

namespace pyftle {

constexpr double kScatterEps = 100.0 * 2.220446049250313e-16;
constexpr double kScatterEpsBroad = 1.4901161193847656e-08;  // sqrt(DBL_EPSILON)

template <typename T>
struct ScatterRec;
template <>
struct ScatterRec<float> {
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
struct ScatterRec<double> {
  template <int DIM>
  static __device__ __forceinline__ void load(const char *slot, int64_t rec, double *v) {
    const double2 *p = reinterpret_cast<const double2 *>(slot) + rec * (DIM == 2 ? 1 : 2);
    const double2 a = p[0];
    v[0] = a.x;
    v[1] = a.y;
    if constexpr (DIM == 3) v[2] = p[1].x;
  }
};

template <int DIM>
__device__ __forceinline__ bool barycentric_inside(const double *__restrict__ T, const double *x, double eps,
                                                   double *c) {
  c[DIM] = 1.0;
  bool in = true;
#pragma unroll
  for (int i = 0; i < DIM; ++i) {
    double ci = 0.0;
#pragma unroll
    for (int j = 0; j < DIM; ++j) ci += T[DIM * i + j] * (x[j] - T[DIM * DIM + j]);
    c[i] = ci;
    c[DIM] -= ci;
    in = in && (ci >= -eps) && (ci <= 1.0 + eps);
  }
  return in && (c[DIM] >= -eps) && (c[DIM] <= 1.0 + eps);
}

template <int DIM, typename T>
__device__ __forceinline__ void sample_scatter_linear(const FieldDev &f, const char *slot, const double *x,
                                                      double *out) {
  const double nan = __longlong_as_double(0x7ff8000000000000LL);
  int64_t b = 0, pitch = 1;
  bool in = true;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    const double q = (x[a] - f.s_org[a]) * f.s_inv[a];
    in = in && (q >= 0.0) && (q < (double)f.s_dims[a]);
    int k = in ? (int)q : 0;
    k = k >= f.s_dims[a] ? f.s_dims[a] - 1 : k;
    b += (int64_t)k * pitch;
    pitch *= f.s_dims[a];
  }
  int found = -1;
  double c[DIM + 1];
  if (in) {
    const int e0 = f.s_off[b], e1 = f.s_off[b + 1];
    for (int e = e0; e < e1 && found < 0; ++e) {
      const int s = f.s_ids[e];
      if (barycentric_inside<DIM>(f.transform + (int64_t)s * (DIM + 1) * DIM, x, kScatterEps, c)) found = s;
    }
    for (int e = e0; e < e1 && found < 0; ++e) {
      const int s = f.s_ids[e];
      if (f.hull[s] &&
          barycentric_inside<DIM>(f.transform + (int64_t)s * (DIM + 1) * DIM, x, kScatterEpsBroad, c))
        found = s;
    }
  }
  if (found < 0) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = nan;
    return;
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) out[d] = 0.0;
#pragma unroll
  for (int k = 0; k <= DIM; ++k) {
    const int vtx = f.simplices[(int64_t)found * (DIM + 1) + k];
    double v[3];
    ScatterRec<T>::template load<DIM>(slot, vtx, v);
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = out[d] + c[k] * v[d];
  }
}

// squared distance from t to the interval [lo, hi]
__device__ __forceinline__ double gap2(double t, double lo, double hi) {
  const double g = t < lo ? lo - t : (t > hi ? t - hi : 0.0);
  return g * g;
}

template <int DIM, typename T>
__device__ __forceinline__ void sample_scatter_nearest(const FieldDev &f, const char *slot, const double *x,
                                                       double *out) {
  bool nanx = false;
  int bc[3] = {0, 0, 0};
  double glo[3], ghi[3];
  int rmax = 0;
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    nanx = nanx || (x[a] != x[a]);
    const double q = (x[a] - f.p_org[a]) * f.p_inv[a];
    const int n = f.p_dims[a];
    int k = (q > 0.0) ? (q < (double)(n - 1) ? (int)q : n - 1) : 0;
    bc[a] = k;
    glo[a] = f.p_org[a];
    ghi[a] = f.p_org[a] + n * f.p_size[a];
    rmax = max(rmax, max(k, n - 1 - k));
  }
  if (nanx) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = __longlong_as_double(0x7ff8000000000000LL);
    return;
  }
  double base2 = 0.0;
#pragma unroll
  for (int a = 0; a < DIM; ++a) base2 += gap2(x[a], glo[a], ghi[a]);
  double best = 1.0e300;
  int64_t best_i = -1;
  for (int r = 0; r <= rmax; ++r) {
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      lo[a] = max(bc[a] - r, 0);
      hi[a] = min(bc[a] + r, f.p_dims[a] - 1);
    }
    const int z0 = DIM == 3 ? lo[2] : 0, z1 = DIM == 3 ? hi[2] : 0;
    auto visit = [&](int bx, int by, int bz) {
      const int64_t b = bx + (int64_t)f.p_dims[0] * (by + (int64_t)f.p_dims[1] * bz);
      const int e1 = f.p_off[b + 1];
      for (int e = f.p_off[b]; e < e1; ++e) {
        const int pi = f.p_ids[e];
        double d2 = 0.0;
#pragma unroll
        for (int a = 0; a < DIM; ++a) {
          const double t = x[a] - f.points[(int64_t)pi * DIM + a];
          d2 += t * t;
        }
        if (d2 < best || (d2 == best && pi < best_i)) {
          best = d2;
          best_i = pi;
        }
      }
    };
    // the shell of Chebyshev radius r: whole rows where y or z is on the shell,
    // only the two x ends elsewhere
    for (int bz = z0; bz <= z1; ++bz) {
      const bool zin = DIM == 3 ? abs(bz - bc[2 % DIM]) < r : true;
      for (int by = lo[1]; by <= hi[1]; ++by) {
        if (zin && abs(by - bc[1]) < r) {
          if (bc[0] - r >= 0) visit(bc[0] - r, by, bz);
          if (bc[0] + r <= f.p_dims[0] - 1) visit(bc[0] + r, by, bz);
        } else {
          for (int bx = lo[0]; bx <= hi[0]; ++bx) visit(bx, by, bz);
        }
      }
    }
    // lower bound on the squared distance to any bucket outside the searched box
    double lb = 1.0e300;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      const double others = base2 - gap2(x[a], glo[a], ghi[a]);
      if (lo[a] > 0) {
        const double edge = f.p_org[a] + lo[a] * f.p_size[a];
        lb = fmin(lb, others + gap2(x[a], glo[a], edge));
      }
      if (hi[a] < f.p_dims[a] - 1) {
        const double edge = f.p_org[a] + (hi[a] + 1) * f.p_size[a];
        lb = fmin(lb, others + gap2(x[a], edge, ghi[a]));
      }
    }
    // (the bound is shaded by 1e-9 relative against rounding of the host bucket assignment)
    if (best_i >= 0 && best <= lb * (1.0 - 1e-9)) break;
  }
  ScatterRec<T>::template load<DIM>(slot, best_i < 0 ? 0 : best_i, out);
}

template <int DIM, int INTERP, typename T>
__device__ __forceinline__ void sample_scatter(const FieldDev &f, const char *slot, const double *x, double *vel) {
  if constexpr (INTERP == PYFTLE_INTERP_SCATTER_LINEAR) sample_scatter_linear<DIM, T>(f, slot, x, vel);
  else sample_scatter_nearest<DIM, T>(f, slot, x, vel);
}

}  // namespace pyftle
