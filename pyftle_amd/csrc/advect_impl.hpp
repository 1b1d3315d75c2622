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
// Precision modes (template PREC):
//   NATIVE: FP32 field -> the cell-local FP32 kernel (advect_kernel_cell);
//           FP64 field -> FP64 throughout in the reference's operation order
//           (stage x + (0.5*h)*k, update x += (h/6)*(k1+2k2+2k3+k4)).
//   MIXED:  FP32 field, FP64 positions / cell index / weights, FP32 blend.
//   F64C:   FP32 field, FP64 positions and FP64 blend.
#pragma once

#include <type_traits>

#include "common.hpp"
#include "rgi.hpp"
#include "scatter.hpp"

namespace pyftle {

enum { PREC_NATIVE = 0, PREC_MIXED = 1, PREC_F64C = 2 };

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
  const int *order;  // processing order of the seed groups (NULL: identity)
  unsigned long long *load_count;  // counting launches: wave-level record loads issued (else NULL)
};

// index of the k-th processed seed group (an entry outside [0, n_groups) maps to k)
__device__ __forceinline__ int64_t seed_at(const AdvectArgs &a, int64_t k) {
  if (!a.order) return k;
  const int64_t q = a.order[k];
  return (q >= 0 && q < a.n_groups) ? q : k;
}

template <int PREC, bool STORE64>
struct PrecTraits;
template <>
struct PrecTraits<PREC_NATIVE, false> {
  using Pos = float;
  using Blend = float;
};
template <>
struct PrecTraits<PREC_NATIVE, true> {
  using Pos = double;
  using Blend = double;
};
template <>
struct PrecTraits<PREC_MIXED, false> {
  using Pos = double;
  using Blend = float;
};
template <>
struct PrecTraits<PREC_F64C, false> {
  using Pos = double;
  using Blend = double;
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

template <int DIM, bool PAIR, bool STORE64, typename Pos, typename Blend>
__device__ __forceinline__ void sample(const FieldDev &f, const char *slot, double w, const Pos *x,
                                       Blend *v) {
  if constexpr (STORE64) {
    sample_field<DIM, PAIR>(f, slot, w, x, v);  // reference operation order
  } else {
    sample_field_f32<DIM, PAIR, Pos, Blend>(f, slot, (Blend)w, x, v);
  }
}

// GridInterp (INTERP 0), a RegularGridInterpolator mode or a scattered-data mode
// (the last two with FP64 positions and blend).
template <int DIM, bool PAIR, bool STORE64, int INTERP, typename Pos, typename Blend>
__device__ __forceinline__ void sample_any(const FieldDev &f, const char *slot, double w, const Pos *x,
                                           Blend *v) {
  if constexpr (INTERP == PYFTLE_INTERP_GINTERP) {
    sample<DIM, PAIR, STORE64>(f, slot, w, x, v);
  } else {
    static_assert(std::is_same<Pos, double>::value && std::is_same<Blend, double>::value,
                  "RegularGridInterpolator and scattered-data modes run in FP64");
    using T = typename std::conditional<STORE64, double, float>::type;
    if constexpr (INTERP == PYFTLE_INTERP_SCATTER_LINEAR || INTERP == PYFTLE_INTERP_SCATTER_NEAREST ||
                  INTERP == PYFTLE_INTERP_SCATTER_CUBIC)
      sample_scatter<DIM, INTERP, T>(f, slot, x, v);
    else
      sample_rgi<DIM, INTERP, T>(f, slot, x, v);
  }
}

// FTLE epilogue: LDS exchange of the final and initial positions of every
// particle of the work-group, then lane sl < G forms the Jacobian of seed sl
// (cauchy_green.py:70-74,140-151) and its FTLE in FP64.
template <int DIM>
__device__ __forceinline__ void ftle_epilogue(const AdvectArgs &a, int64_t lb, int t, int sl, int j,
                                              const double *fin, const double *x0) {
  __shared__ double sh_fin[512 * DIM];
  __shared__ double sh_ini[512 * DIM];
  const int nb = a.nb, G = a.G;
  const int li = sl * nb + j;
  if (t < G * nb) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      sh_fin[li * DIM + c] = fin[c];
      sh_ini[li * DIM + c] = x0[c];
    }
  }
  __syncthreads();
  if (t < G) {
    const int64_t k = lb * G + t;
    if (k < a.n_groups) {
      const int64_t seed = seed_at(a, k);
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

template <int DIM, int METHOD, int TMODE, bool PAIR, bool STORE64, int PREC, bool FTLE,
          int INTERP = PYFTLE_INTERP_GINTERP>
__global__ void __launch_bounds__(512) advect_kernel(AdvectArgs a) {
  using Pos = typename PrecTraits<PREC, STORE64>::Pos;
  using Blend = typename PrecTraits<PREC, STORE64>::Blend;
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
  const int64_t pidx = (int64_t)j * a.n_groups + (active ? seed_at(a, s) : s);

  double x0[DIM];
  PosState<Pos> X[DIM];
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
  const Pos h = (Pos)a.h;
  Pos vprev[DIM];
  bool have_prev = (METHOD == PYFTLE_AB2) && a.have_prev;
  if (METHOD == PYFTLE_AB2) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      Pos v = (Pos)0;
      if (active && have_prev && a.vprev)
        v = a.pos_f64 ? (Pos)((const double *)a.vprev)[pidx * DIM + c]
                      : (Pos)((const float *)a.vprev)[pidx * DIM + c];
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
    Pos xs[DIM];
    Blend k1[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) xs[c] = X[c].hi();
    sample_any<DIM, PAIR, STORE64, INTERP>(f, s0, w0, xs, k1);
    if (METHOD == PYFTLE_RK4) {
      Blend k2[DIM], k3[DIM], k4[DIM];
      const Pos hh = (Pos)(0.5 * a.h);
#pragma unroll
      for (int c = 0; c < DIM; ++c) xs[c] = stage_pos<Pos>(X[c].hi(), hh, (Pos)k1[c]);
      sample_any<DIM, PAIR, STORE64, INTERP>(f, s1, w1, xs, k2);
#pragma unroll
      for (int c = 0; c < DIM; ++c) xs[c] = stage_pos<Pos>(X[c].hi(), hh, (Pos)k2[c]);
      sample_any<DIM, PAIR, STORE64, INTERP>(f, s1, w1, xs, k3);
#pragma unroll
      for (int c = 0; c < DIM; ++c) xs[c] = stage_pos<Pos>(X[c].hi(), h, (Pos)k3[c]);
      sample_any<DIM, PAIR, STORE64, INTERP>(f, s2, w2, xs, k4);
      const Pos h6 = (Pos)(a.h / 6.0);
#pragma unroll
      for (int c = 0; c < DIM; ++c)
        X[c].add(rk4_increment<Pos>(h6, (Pos)k1[c], (Pos)k2[c], (Pos)k3[c], (Pos)k4[c]));
    } else if (METHOD == PYFTLE_EULER) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) X[c].add(h * (Pos)k1[c]);  // integrate.py:44
    } else {
      if (have_prev) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) X[c].add(ab2_increment<Pos>(h, (Pos)k1[c], vprev[c]));
      } else {
#pragma unroll
        for (int c = 0; c < DIM; ++c) X[c].add(h * (Pos)k1[c]);  // first AB2 step is Euler
      }
#pragma unroll
      for (int c = 0; c < DIM; ++c) vprev[c] = (Pos)k1[c];
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
    double fin[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) fin[c] = X[c].value();
    ftle_epilogue<DIM>(a, lb, t, sl, j, fin, x0);
  }
}

// ---------------------------------------------------------------------------
// FP32 cell-local kernel (the FP32 fast path).
//
// Each coordinate is held in index space as an integer cell plus an FP32
// fraction in [0, 1) with a Kahan compensation term (index = cell + frac -
// comp), so every stage position, cell lookup and weight is plain FP32/int
// arithmetic yet carries ~1e-8 cell of resolution anywhere in the grid.
// Periodic axes keep the cell in [0, n) and count wraps separately so the
// final (unwrapped) position and the Jacobian stay exact.  Snapshot records
// are fetched with buffer loads (32-bit offsets, hardware range check).
// ---------------------------------------------------------------------------
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 raw buffer descriptor word 3

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const char *slot, int bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(slot), (short)0, bytes, kRsrcWord3);
}

// Buffer loads of one record chunk.  The loaded vector is copied out whole:
// extracting its elements one by one made this compiler emit a single-dword
// load (the remaining elements were left undefined).
template <int NF>
struct BufVec;
template <>
struct BufVec<2> {
  static __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t r, int off, float *o) {
    auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    __builtin_memcpy(o, &v, 8);
  }
};
template <>
struct BufVec<4> {
  static __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t r, int off, float *o) {
    auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    __builtin_memcpy(o, &v, 16);
  }
};

// Lower cell index of one axis for a stage offset t (index units relative to
// the particle's cell).  Branch-free: per_n is n on a periodic axis (0
// otherwise) and lim the first invalid lower index (n, or n - 1 when the field
// is zero outside).  The upper corner is always i + 1 (ghost node on periodic
// axes).  Offsets of invalid points may be garbage: the loads are range
// checked by the buffer descriptor and the result is masked by `ok`.
__device__ __forceinline__ int cell_axis(const FieldDev &f, int a, int cell, float t, float &d,
                                         bool &ok) {
  const float fl = floorf(t);
  d = t - fl;
  int i = cell + (int)fl;
  const int pn = f.per_n[a];
  i += (i < 0) ? pn : 0;
  i -= (i >= pn) ? pn : 0;
  ok = ok && ((unsigned)i < (unsigned)f.lim[a]);
  return i;
}

// Per-lane cache of the last per-cell record a lane read: its record index,
// its time slot and its floats.  A sample of the same record in the same slot
// reads no memory; only the lanes whose sample moved to another cell or slot
// take part in the loads, and a wave in which none did skips them.  Within a
// launch a record's contents depend on (slot, index) alone, so the results are
// the same bit for bit.  Steady fields pass slot 0 throughout (a constant the
// compiler folds away); the others start from slot -1, which no sample has.
// COUNT adds the wave-level load instructions the sample issues to *nload
// (uniform over the wave: a wave issues a record's loads when any lane misses).
template <int DIM, bool PAIR, bool CELL, bool CACHE = false, bool SLOTS = false, bool COUNT = false>
__device__ __forceinline__ void sample_cell(const FieldDev &f, __amdgpu_buffer_rsrc_t rs, float w,
                                            const int *cell, const float *t, float *vel,
                                            int *ckey = nullptr, int *cslot = nullptr,
                                            float *crec = nullptr, int slot = 0, int *nload = nullptr) {
  constexpr int C = RecInfo<DIM, PAIR>::comps;
  constexpr int L = RecInfo<DIM, PAIR>::per_level;
  constexpr int NF = C > 4 ? 4 : C;
  constexpr int chunks = C > 4 ? 2 : 1;
  constexpr int RB = C * 4;  // record bytes
  int i0[3];
  float d[3];
  bool ok = true;
#pragma unroll
  for (int a = 0; a < DIM; ++a) i0[a] = cell_axis(f, a, cell[a], t[a], d[a], ok);
  // node offsets in records; rows and slices hold < 2^24 records, so 24-bit
  // multiplies (full rate) are exact
  int base = i0[0] + (int)__umul24((unsigned)i0[1], (unsigned)f.px);
  if constexpr (DIM == 3) base += (int)__umul24((unsigned)i0[2], (unsigned)f.pxy);
  constexpr int NQ = (DIM == 2) ? 4 : 8;
  float c[NQ][C];
  if constexpr (CELL) {
    // one record per cell: level l, corner q, component comp at float (l NQ + q) DIM + comp
    constexpr int LV = PAIR ? 2 : 1;
    constexpr int RF = NQ * DIM * LV;  // 8, 16, 24 or 48 floats
    float local[RF];
    float *rec = CACHE ? crec : local;
    const unsigned roff = (unsigned)base * (unsigned)(RF * 4);  // out-of-range offsets read 0
    const bool miss = !CACHE || base != *ckey || (SLOTS && slot != *cslot);
    if constexpr (COUNT) *nload += __builtin_amdgcn_ballot_w64(miss) ? RF / 4 : 0;
    if (miss) {
#pragma unroll
      for (int h = 0; h < RF / 4; ++h) BufVec<4>::load(rs, (int)(roff + 16u * h), &rec[h * 4]);
      if constexpr (CACHE) *ckey = base;
    }
    if constexpr (CACHE && SLOTS) *cslot = slot;
#pragma unroll
    for (int l = 0; l < LV; ++l)
#pragma unroll
      for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int comp = 0; comp < DIM; ++comp) c[q][l * L + comp] = rec[(l * NQ + q) * DIM + comp];
  } else {
    if constexpr (COUNT) *nload += NQ * chunks;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int node = base + ((q & 1) ? 1 : 0) + ((q & 2) ? f.px : 0) + ((q & 4) ? f.pxy : 0);
#pragma unroll
      for (int h = 0; h < chunks; ++h) BufVec<NF>::load(rs, node * RB + h * 16, &c[q][h * 4]);
    }
  }
  const float xd = d[0], yd = d[1], zd = d[DIM - 1];
#pragma unroll
  for (int comp = 0; comp < DIM; ++comp) {
    float lev[2];
#pragma unroll
    for (int l = 0; l < (PAIR ? 2 : 1); ++l) {
      const int e = l * L + comp;
      const float c0 = lerp_f<float>(c[0][e], c[1][e], xd);
      const float c1 = lerp_f<float>(c[2][e], c[3][e], xd);
      float r = lerp_f<float>(c0, c1, yd);
      if constexpr (DIM == 3) {
        const float c2 = lerp_f<float>(c[4 % NQ][e], c[5 % NQ][e], xd);
        const float c3 = lerp_f<float>(c[6 % NQ][e], c[7 % NQ][e], xd);
        r = lerp_f<float>(r, lerp_f<float>(c2, c3, yd), zd);
      }
      lev[l] = r;
    }
    float v = lev[0];
    if (PAIR) v = lerp_f<float>(lev[0], lev[1 % (PAIR ? 2 : 1)], w);
    vel[comp] = ok ? v : 0.0f;
  }
}

// wrap an integer cell into [0, n) and return the number of wraps
__device__ __forceinline__ int wrap_cell(int &cell, int n, float inv_n) {
  int q = (int)floorf((float)cell * inv_n);
  int r = cell - q * n;
  if (r < 0) {
    r += n;
    q -= 1;
  }
  if (r >= n) {
    r -= n;
    q += 1;
  }
  cell = r;
  return q;
}

// One work-group's share of a window: logical block lb of the seeds of `a`.
template <int DIM, int METHOD, int TMODE, bool PAIR, bool FTLE, bool CELL, bool COUNT = false>
__device__ __forceinline__ void advect_block_cell(const AdvectArgs &a, int64_t lb) {
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
  const int64_t pidx = (int64_t)j * a.n_groups + (active ? seed_at(a, s) : s);
  const FieldDev &f = a.f;

  int cell[DIM], wraps[DIM];
  float frac[DIM], comp[DIM];
  float sh[DIM], sf[DIM];
  double sfd[DIM], s6d[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    double v = 0.0;
    if (active)
      v = a.pos_f64 ? ((const double *)a.pos_in)[pidx * DIM + c]
                    : (double)((const float *)a.pos_in)[pidx * DIM + c];
    const double fi = (v - f.origin[c]) * f.inv_spacing[c];
    double fl = floor(fi);
    fl = fl < -1.0e9 ? -1.0e9 : (fl > 1.0e9 ? 1.0e9 : fl);
    fl = (fl == fl) ? fl : -1.0e9;  // NaN -> far outside
    int ci = (int)fl;
    const double r = fi - fl;
    float fr = (float)r;
    float cp = (float)((double)fr - r);
    if (fr >= 1.0f) {  // r rounded up to 1
      fr -= 1.0f;
      ci += 1;
    }
    wraps[c] = f.periodic[c] ? wrap_cell(ci, f.n[c], f.inv_n_f[c]) : 0;
    cell[c] = ci;
    frac[c] = (r == r) ? fr : 0.0f;
    comp[c] = (r == r) ? cp : 0.0f;
    sh[c] = (float)(0.5 * a.h * f.inv_spacing[c]);
    sf[c] = (float)(a.h * f.inv_spacing[c]);
    sfd[c] = a.h * f.inv_spacing[c];
    s6d[c] = a.h / 6.0 * f.inv_spacing[c];
  }

  float vprev[DIM];
  bool have_prev = (METHOD == PYFTLE_AB2) && a.have_prev;
  if (METHOD == PYFTLE_AB2) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      float v = 0.0f;
      if (active && have_prev && a.vprev)
        v = a.pos_f64 ? (float)((const double *)a.vprev)[pidx * DIM + c]
                      : ((const float *)a.vprev)[pidx * DIM + c];
      vprev[c] = v;
    }
  }

  const int bytes = (int)f.slot_stride;
  const __amdgpu_buffer_rsrc_t steady = make_rsrc(slot_ptr(f, a.first), bytes);
  // per-lane record cache for per-cell records (key -1: byte offset 2^32 minus
  // the record size, past the end of any slot, whose load reads zeros)
  constexpr bool CACHE = CELL;
  constexpr int CRF = CACHE ? (1 << DIM) * DIM * (PAIR ? 2 : 1) : 1;
  int ckey = -1, cslot = -1;
  int nload = 0;  // COUNT: wave-level record loads issued by this wave
  float crec[CRF];
#pragma unroll
  for (int i = 0; i < CRF; ++i) crec[i] = 0.0f;
  for (int step = 0; step < a.n_steps; ++step) {
    __amdgpu_buffer_rsrc_t r0 = steady, r1 = steady, r2 = steady;
    double w0 = 0.0, w1 = 0.0, w2 = 0.0;
    int q0 = 0, q1 = 0, q2 = 0;  // time slots of the three sample times (cache keys)
    if (TMODE == PYFTLE_TIME_FROZEN) {
      q0 = q1 = q2 = a.first + step;
      r0 = r1 = r2 = make_rsrc(slot_ptr(f, q0), bytes);
    } else if (TMODE == PYFTLE_TIME_LINEAR) {
      const double tt = a.t_start + step * a.h;
      time_linear_lookup(f, tt, q0, w0);
      r0 = make_rsrc(slot_ptr(f, q0), bytes);
      if (METHOD == PYFTLE_RK4) {
        time_linear_lookup(f, tt + 0.5 * a.h, q1, w1);
        r1 = make_rsrc(slot_ptr(f, q1), bytes);
        time_linear_lookup(f, tt + a.h, q2, w2);
        r2 = make_rsrc(slot_ptr(f, q2), bytes);
      }
    }
    float tp[DIM], k1[DIM];
    double inc[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) tp[c] = frac[c];
    sample_cell<DIM, PAIR, CELL, CACHE, TMODE != PYFTLE_TIME_STEADY, COUNT>(f, r0, (float)w0, cell, tp, k1, &ckey, &cslot, crec, q0, &nload);
    if (METHOD == PYFTLE_RK4) {
      float k2[DIM], k3[DIM], k4[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) tp[c] = fmaf(sh[c], k1[c], frac[c]);
      sample_cell<DIM, PAIR, CELL, CACHE, TMODE != PYFTLE_TIME_STEADY, COUNT>(f, r1, (float)w1, cell, tp, k2, &ckey, &cslot, crec, q1, &nload);
#pragma unroll
      for (int c = 0; c < DIM; ++c) tp[c] = fmaf(sh[c], k2[c], frac[c]);
      sample_cell<DIM, PAIR, CELL, CACHE, TMODE != PYFTLE_TIME_STEADY, COUNT>(f, r1, (float)w1, cell, tp, k3, &ckey, &cslot, crec, q1, &nload);
#pragma unroll
      for (int c = 0; c < DIM; ++c) tp[c] = fmaf(sf[c], k3[c], frac[c]);
      sample_cell<DIM, PAIR, CELL, CACHE, TMODE != PYFTLE_TIME_STEADY, COUNT>(f, r2, (float)w2, cell, tp, k4, &ckey, &cslot, crec, q2, &nload);
#pragma unroll
      for (int c = 0; c < DIM; ++c)
        inc[c] = s6d[c] * ((double)k1[c] + 2.0 * ((double)k2[c] + (double)k3[c]) + (double)k4[c]);
    } else if (METHOD == PYFTLE_EULER) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) inc[c] = sfd[c] * (double)k1[c];
    } else {
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        inc[c] = have_prev ? sfd[c] * (1.5 * (double)k1[c] - 0.5 * (double)vprev[c])
                           : sfd[c] * (double)k1[c];
        vprev[c] = k1[c];
      }
      have_prev = true;
    }
    // Once per step: fold the increment into (cell, frac, comp) in FP64 so
    // the accumulated position carries no FP32 rounding of the increment.
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const double fr = ((double)frac[c] - (double)comp[c]) + inc[c];
      const double fl = floor(fr);
      const double rem = fr - fl;
      float nf = (float)rem;
      float nc = (float)((double)nf - rem);
      int dc = (int)fl;
      if (nf >= 1.0f) {
        nf -= 1.0f;
        dc += 1;
      }
      frac[c] = nf;
      comp[c] = nc;
      cell[c] += dc;
      if (f.periodic[c] && (cell[c] < 0 || cell[c] >= f.n[c]))
        wraps[c] += wrap_cell(cell[c], f.n[c], f.inv_n_f[c]);
    }
  }

  double fin[DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c) {
    const double idx = (double)cell[c] + (double)wraps[c] * (double)f.n[c] + (double)frac[c] -
                       (double)comp[c];
    fin[c] = f.origin[c] + idx * f.spacing[c];
  }
  // The seed position is read again for the epilogue instead of being held
  // through the step loop (six registers of FP64).  The empty asm hides that
  // the index equals the first read's, so that read is not reused; pos_out,
  // which may be pos_in, is written after this.
  double x0[DIM];
  if constexpr (FTLE) {
    int64_t q = pidx;
    asm volatile("" : "+v"(q));
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      double v = 0.0;
      if (active)
        v = a.pos_f64 ? ((const double *)a.pos_in)[q * DIM + c]
                      : (double)((const float *)a.pos_in)[q * DIM + c];
      x0[c] = v;
    }
  }
  if (active && a.pos_out) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      if (a.pos_f64)
        ((double *)a.pos_out)[pidx * DIM + c] = fin[c];
      else
        ((float *)a.pos_out)[pidx * DIM + c] = (float)fin[c];
    }
  }
  if (METHOD == PYFTLE_AB2 && active && a.vprev && a.n_steps > 0) {
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      if (a.pos_f64)
        ((double *)a.vprev)[pidx * DIM + c] = (double)vprev[c];
      else
        ((float *)a.vprev)[pidx * DIM + c] = vprev[c];
    }
  }
  if constexpr (FTLE) ftle_epilogue<DIM>(a, lb, t, sl, j, fin, x0);
  if constexpr (COUNT) {
    if ((threadIdx.x & 63) == 0 && a.load_count) atomicAdd(a.load_count, (unsigned long long)nload);
  }
}

template <int DIM, int METHOD, int TMODE, bool PAIR, bool FTLE, bool CELL, bool COUNT = false>
__global__ void __launch_bounds__(512) advect_kernel_cell(AdvectArgs a) {
  advect_block_cell<DIM, METHOD, TMODE, PAIR, FTLE, CELL, COUNT>(a, logical_block(a.remap));
}

// ---------------------------------------------------------------------------
// host-side dispatch over the template parameters (one translation unit per
// dimension instantiates these)
// ---------------------------------------------------------------------------
struct LaunchSpec {
  int method, tmode, pair, store64, prec, ftle;
  unsigned blocks, threads;
  hipStream_t stream;
  int cell;   // per-cell records (FP32, PREC_NATIVE only)
  int count;  // counting launch (RK4 FTLE windows on per-cell records; AdvectArgs.load_count)
};

template <int DIM, int METHOD, int TMODE, bool PAIR, bool STORE64, int PREC>
void launch_leaf(const AdvectArgs &a, const LaunchSpec &sp) {
  if constexpr (!STORE64 && PREC == PREC_NATIVE) {
    if constexpr (METHOD == PYFTLE_RK4) {
      if (sp.count) {  // the caller checked: per-cell records, FTLE window
        advect_kernel_cell<DIM, METHOD, TMODE, PAIR, true, true, true><<<sp.blocks, sp.threads, 0, sp.stream>>>(a);
        return;
      }
    }
    if (sp.cell) {
      if (sp.ftle)
        advect_kernel_cell<DIM, METHOD, TMODE, PAIR, true, true><<<sp.blocks, sp.threads, 0, sp.stream>>>(a);
      else
        advect_kernel_cell<DIM, METHOD, TMODE, PAIR, false, true><<<sp.blocks, sp.threads, 0, sp.stream>>>(a);
    } else if (sp.ftle) {
      advect_kernel_cell<DIM, METHOD, TMODE, PAIR, true, false><<<sp.blocks, sp.threads, 0, sp.stream>>>(a);
    } else {
      advect_kernel_cell<DIM, METHOD, TMODE, PAIR, false, false><<<sp.blocks, sp.threads, 0, sp.stream>>>(a);
    }
  } else {
    if (sp.ftle)
      advect_kernel<DIM, METHOD, TMODE, PAIR, STORE64, PREC, true><<<sp.blocks, sp.threads, 0, sp.stream>>>(a);
    else
      advect_kernel<DIM, METHOD, TMODE, PAIR, STORE64, PREC, false><<<sp.blocks, sp.threads, 0, sp.stream>>>(a);
  }
}

template <int DIM, int METHOD, int TMODE, bool PAIR>
void launch_prec(const AdvectArgs &a, const LaunchSpec &sp) {
  if (sp.store64) {
    launch_leaf<DIM, METHOD, TMODE, PAIR, true, PREC_NATIVE>(a, sp);
  } else if (sp.prec == PREC_MIXED) {
    launch_leaf<DIM, METHOD, TMODE, PAIR, false, PREC_MIXED>(a, sp);
  } else if (sp.prec == PREC_F64C) {
    launch_leaf<DIM, METHOD, TMODE, PAIR, false, PREC_F64C>(a, sp);
  } else {
    launch_leaf<DIM, METHOD, TMODE, PAIR, false, PREC_NATIVE>(a, sp);
  }
}

template <int DIM, int METHOD>
void launch_time(const AdvectArgs &a, const LaunchSpec &sp) {
  if (sp.tmode == PYFTLE_TIME_LINEAR) {
    launch_prec<DIM, METHOD, PYFTLE_TIME_LINEAR, true>(a, sp);
  } else if (sp.tmode == PYFTLE_TIME_FROZEN) {
    if (sp.pair) launch_prec<DIM, METHOD, PYFTLE_TIME_FROZEN, true>(a, sp);
    else launch_prec<DIM, METHOD, PYFTLE_TIME_FROZEN, false>(a, sp);
  } else {
    if (sp.pair) launch_prec<DIM, METHOD, PYFTLE_TIME_STEADY, true>(a, sp);
    else launch_prec<DIM, METHOD, PYFTLE_TIME_STEADY, false>(a, sp);
  }
}

template <int DIM>
void launch_advect(const AdvectArgs &a, const LaunchSpec &sp) {
  if (sp.method == PYFTLE_RK4) launch_time<DIM, PYFTLE_RK4>(a, sp);
  else if (sp.method == PYFTLE_EULER) launch_time<DIM, PYFTLE_EULER>(a, sp);
  else launch_time<DIM, PYFTLE_AB2>(a, sp);
}

// defined in advect_2d.hip / advect_3d.hip / advect_rgi.hip
void launch_advect_2d(const AdvectArgs &a, const LaunchSpec &sp);
void launch_advect_3d(const AdvectArgs &a, const LaunchSpec &sp);
void launch_advect_rgi(int dim, int interp, const AdvectArgs &a, const LaunchSpec &sp);

}  // namespace pyftle
