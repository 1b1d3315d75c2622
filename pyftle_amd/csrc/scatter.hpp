// Scattered-data strategies on the device (interpolate.py:114-294): SciPy's
// LinearNDInterpolator on a Delaunay triangulation, NearestNDInterpolator and, in
// 2D, CloughTocher2DInterpolator.
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

#include <type_traits>

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

// Index of the simplex holding x (and its barycentric coordinates), or -1.
template <int DIM>
__device__ __forceinline__ int scatter_locate(const FieldDev &f, const double *x, double *c) {
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
  return found;
}

template <int DIM, typename T>
__device__ __forceinline__ void sample_scatter_linear(const FieldDev &f, const char *slot, const double *x,
                                                      double *out) {
  double c[DIM + 1];
  const int found = scatter_locate<DIM>(f, x, c);
  if (found < 0) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = __longlong_as_double(0x7ff8000000000000LL);
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

// (u, v, du/dx, du/dy, dv/dx, dv/dy) of one point of a scattered cubic slot
template <typename T>
__device__ __forceinline__ void ct_record(const char *slot, int64_t rec, double *r) {
  if constexpr (std::is_same<T, float>::value) {
    const float4 *p = reinterpret_cast<const float4 *>(slot) + rec * 2;
    const float4 a = p[0], b = p[1];
    r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w; r[4] = b.x; r[5] = b.y;
  } else {
    const double2 *p = reinterpret_cast<const double2 *>(slot) + rec * 4;
    const double2 a = p[0], b = p[1], d = p[2];
    r[0] = a.x; r[1] = a.y; r[2] = b.x; r[3] = b.y; r[4] = d.x; r[5] = d.y;
  }
}

// Clough-Tocher patch (Farin 1986, "Triangular Bernstein-Bezier patches"): the
// triangle is split at its centroid into three cubic Bezier pieces; vertex values
// and gradients fix the ordinates at and next to the vertices, the cross-edge
// derivative along direction g (toward the neighbouring triangle's centroid,
// g = -1/2 on the hull) is made linear along each edge, and C1 continuity inside
// fixes the ordinates next to and at the centroid.  b: barycentric coordinates.
// The ordinates (c3000 ... c0003) and the cross-edge terms are formed in the order of
// SciPy's CloughTocher2DInterpolator evaluation (scipy/interpolate/interpnd.pyx,
// BSD licence), so that results agree with it to rounding.
__device__ __forceinline__ double ct_patch(const double *P, const double *fv, const double (*gr)[2],
                                           const double *g, const double *b) {
  const double e12x = P[2] - P[0], e12y = P[3] - P[1];
  const double e13x = P[4] - P[0], e13y = P[5] - P[1];
  const double e23x = P[4] - P[2], e23y = P[5] - P[3];
  const double c3000 = fv[0], c0300 = fv[1], c0030 = fv[2];
  const double c2100 = (gr[0][0] * e12x + gr[0][1] * e12y + 3.0 * c3000) / 3.0;
  const double c2010 = (gr[0][0] * e13x + gr[0][1] * e13y + 3.0 * c3000) / 3.0;
  const double c1200 = (-(gr[1][0] * e12x + gr[1][1] * e12y) + 3.0 * c0300) / 3.0;
  const double c0210 = (gr[1][0] * e23x + gr[1][1] * e23y + 3.0 * c0300) / 3.0;
  const double c1020 = (-(gr[2][0] * e13x + gr[2][1] * e13y) + 3.0 * c0030) / 3.0;
  const double c0120 = (-(gr[2][0] * e23x + gr[2][1] * e23y) + 3.0 * c0030) / 3.0;
  const double c2001 = (c2100 + c2010 + c3000) / 3.0;
  const double c0201 = (c1200 + c0300 + c0210) / 3.0;
  const double c0021 = (c1020 + c0120 + c0030) / 3.0;
  const double c0111 = (g[0] * (-c0300 + 3.0 * c0210 - 3.0 * c0120 + c0030) +
                        (-c0300 + 2.0 * c0210 - c0120 + c0021 + c0201)) / 2.0;
  const double c1011 = (g[1] * (-c0030 + 3.0 * c1020 - 3.0 * c2010 + c3000) +
                        (-c0030 + 2.0 * c1020 - c2010 + c2001 + c0021)) / 2.0;
  const double c1101 = (g[2] * (-c3000 + 3.0 * c2100 - 3.0 * c1200 + c0300) +
                        (-c3000 + 2.0 * c2100 - c1200 + c2001 + c0201)) / 2.0;
  const double c1002 = (c1101 + c1011 + c2001) / 3.0;
  const double c0102 = (c1101 + c0111 + c0201) / 3.0;
  const double c0012 = (c1011 + c0111 + c0021) / 3.0;
  const double c0003 = (c1002 + c0102 + c0012) / 3.0;
  const double mn = fmin(b[0], fmin(b[1], b[2]));
  const double b1 = b[0] - mn, b2 = b[1] - mn, b3 = b[2] - mn, b4 = 3.0 * mn;
  return b1 * b1 * b1 * c3000 + 3.0 * b1 * b1 * b2 * c2100 + 3.0 * b1 * b1 * b3 * c2010 +
         3.0 * b1 * b1 * b4 * c2001 + 3.0 * b1 * b2 * b2 * c1200 + 6.0 * b1 * b2 * b4 * c1101 +
         3.0 * b1 * b3 * b3 * c1020 + 6.0 * b1 * b3 * b4 * c1011 + 3.0 * b1 * b4 * b4 * c1002 +
         b2 * b2 * b2 * c0300 + 3.0 * b2 * b2 * b3 * c0210 + 3.0 * b2 * b2 * b4 * c0201 +
         3.0 * b2 * b3 * b3 * c0120 + 6.0 * b2 * b3 * b4 * c0111 + 3.0 * b2 * b4 * b4 * c0102 +
         b3 * b3 * b3 * c0030 + 3.0 * b3 * b3 * b4 * c0021 + 3.0 * b3 * b4 * b4 * c0012 +
         b4 * b4 * b4 * c0003;
}

template <typename T>
__device__ __forceinline__ void sample_scatter_cubic(const FieldDev &f, const char *slot, const double *x,
                                                     double *out) {
  double c[3];
  const int found = scatter_locate<2>(f, x, c);
  if (found < 0) {
    out[0] = out[1] = __longlong_as_double(0x7ff8000000000000LL);
    return;
  }
  double P[6], fu[3], fw[3], gu[3][2], gw[3][2];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int vtx = f.simplices[(int64_t)found * 3 + k];
    P[2 * k] = f.points[(int64_t)vtx * 2];
    P[2 * k + 1] = f.points[(int64_t)vtx * 2 + 1];
    double r[6];
    ct_record<T>(slot, vtx, r);
    fu[k] = r[0];
    fw[k] = r[1];
    gu[k][0] = r[2];
    gu[k][1] = r[3];
    gw[k][0] = r[4];
    gw[k][1] = r[5];
  }
  const double g[3] = {f.ct_g[(int64_t)found * 3], f.ct_g[(int64_t)found * 3 + 1], f.ct_g[(int64_t)found * 3 + 2]};
  out[0] = ct_patch(P, fu, gu, g, c);
  out[1] = ct_patch(P, fw, gw, g, c);
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
  if constexpr (INTERP == PYFTLE_INTERP_SCATTER_LINEAR) {
    sample_scatter_linear<DIM, T>(f, slot, x, vel);
  } else if constexpr (INTERP == PYFTLE_INTERP_SCATTER_CUBIC) {
    static_assert(DIM == 2, "Clough-Tocher interpolation is 2D");
    sample_scatter_cubic<T>(f, slot, x, vel);
  } else {
    sample_scatter_nearest<DIM, T>(f, slot, x, vel);
  }
}

}  // namespace pyftle
