/*
 * C restatement of pyFTLE's CPU hot path (test infrastructure + CPU baseline).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library.  It follows the reference structure step by step so that its
 * timing is an honest stand-in for the reference CPU path:
 *
 *  - interp2d / interp3d: the scalar bilinear / trilinear loop of
 *    GridInterp.cpp:15-142 (index space, zero outside [0, n-2] cells,
 *    x -> y -> z lerp order), FP64;
 *  - the physical -> index conversion is redone for every component call, as
 *    Interp2D/Interp3D.__call__ do (grid_interp.py:78-79,164-166);
 *  - RK4 keeps full-array stage buffers k1..k4 and tmp (integrate.py:291-342)
 *    and combines them as integrate.py:81-109;
 *  - mode 0 ("reference"): step i uses snapshot i frozen for all stages
 *    (ftle_solver.py:99-102); mode 1 ("time-linear", this build's extension):
 *    stage time t selects snapshots k, k+1 and blends them linearly.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* GridInterp.cpp:93-142 */
void cpu_interp2d(const double *v, int nx, int ny, const double *pts, int64_t n, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double px = pts[i * 2 + 0];
    const double py = pts[i * 2 + 1];
    const int x0 = (int)floor(px);
    const int y0 = (int)floor(py);
    const int x1 = x0 + 1;
    const int y1 = y0 + 1;
    const double xd = px - x0;
    const double yd = py - y0;
    double c = 0.0;
    if (x0 >= 0 && x1 < nx && y0 >= 0 && y1 < ny) {
      const double v00 = v[(int64_t)x0 * ny + y0];
      const double v10 = v[(int64_t)x1 * ny + y0];
      const double v01 = v[(int64_t)x0 * ny + y1];
      const double v11 = v[(int64_t)x1 * ny + y1];
      c = v00 * (1.0 - xd) * (1.0 - yd) + v10 * xd * (1.0 - yd) + v01 * (1.0 - xd) * yd +
          v11 * xd * yd;
    }
    out[i] = c;
  }
}

/* GridInterp.cpp:15-91 */
void cpu_interp3d(const double *v, int X, int Y, int Z, const double *pts, int64_t n,
                  double *out) {
  const int64_t YZ = (int64_t)Y * Z;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double px = pts[i * 3 + 0];
    const double py = pts[i * 3 + 1];
    const double pz = pts[i * 3 + 2];
    const int x0 = (int)floor(px);
    const int y0 = (int)floor(py);
    const int z0 = (int)floor(pz);
    const int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
    const double xd = px - x0, yd = py - y0, zd = pz - z0;
    double c = 0.0;
    if (x0 >= 0 && x1 < X && y0 >= 0 && y1 < Y && z0 >= 0 && z1 < Z) {
#define IDX(a, b, d) ((int64_t)(a) * YZ + (int64_t)(b) * Z + (int64_t)(d))
      const double c00 = v[IDX(x0, y0, z0)] * (1.0 - xd) + v[IDX(x1, y0, z0)] * xd;
      const double c01 = v[IDX(x0, y0, z1)] * (1.0 - xd) + v[IDX(x1, y0, z1)] * xd;
      const double c10 = v[IDX(x0, y1, z0)] * (1.0 - xd) + v[IDX(x1, y1, z0)] * xd;
      const double c11 = v[IDX(x0, y1, z1)] * (1.0 - xd) + v[IDX(x1, y1, z1)] * xd;
#undef IDX
      const double c0 = c00 * (1.0 - yd) + c10 * yd;
      const double c1 = c01 * (1.0 - yd) + c11 * yd;
      c = c0 * (1.0 - zd) + c1 * zd;
    }
    out[i] = c;
  }
}

typedef struct {
  int dim;
  int n[3];           /* grid points per axis (C++ kernel axis order) */
  double mins[3];
  double deltas[3];
  const double *snaps; /* [n_snap][dim][Ngrid], each component C-ordered */
  int n_snap;
  double snap_t0, snap_dt;
  int mode;            /* 0 frozen snapshot per step, 1 time-linear */
  double *idx;         /* (P, dim) index buffer (grid_interp.py:72-79) */
  double *comp;        /* (P,) component buffer (interpolate.py:505) */
  double *comp2;
} field_t;

static int64_t grid_count(const field_t *f) {
  int64_t g = 1;
  for (int a = 0; a < f->dim; ++a) g *= f->n[a];
  return g;
}

/* One component call: index conversion then the scalar loop
 * (grid_interp.py:40-85 / 123-172). */
static void component_call(const field_t *f, const double *v, const double *x, int64_t P,
                           double *out) {
  const int dim = f->dim;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < P; ++i)
    for (int a = 0; a < dim; ++a) f->idx[i * dim + a] = (x[i * dim + a] - f->mins[a]) / f->deltas[a];
  if (dim == 2)
    cpu_interp2d(v, f->n[0], f->n[1], f->idx, P, out);
  else
    cpu_interp3d(v, f->n[0], f->n[1], f->n[2], f->idx, P, out);
}

/* interpolate.py:475-543 for snapshot k (or the k/k+1 blend in mode 1). */
static void velocity(const field_t *f, int step, double t, const double *x, int64_t P,
                     double *out) {
  const int dim = f->dim;
  const int64_t G = grid_count(f);
  int k = step;
  double w = 0.0;
  if (f->mode == 1) {
    double s = (t - f->snap_t0) / f->snap_dt;
    k = (int)floor(s);
    if (k < 0) k = 0;
    if (k > f->n_snap - 2) k = f->n_snap - 2;
    if (k < 0) k = 0;
    w = s - k;
    if (w < 0.0) w = 0.0;
    if (w > 1.0) w = 1.0;
  }
  for (int c = 0; c < dim; ++c) {
    component_call(f, f->snaps + ((int64_t)k * dim + c) * G, x, P, f->comp);
    if (f->mode == 1 && w != 0.0 && f->n_snap > 1) {
      component_call(f, f->snaps + ((int64_t)(k + 1) * dim + c) * G, x, P, f->comp2);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < P; ++i) out[i * dim + c] = (1.0 - w) * f->comp[i] + w * f->comp2[i];
    } else {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < P; ++i) out[i * dim + c] = f->comp[i];
    }
  }
}

/*
 * Advect P particles (positions (P, dim), in place) for n_steps.
 * method: 0 euler, 1 ab2, 2 rk4.  Returns 0 on success.
 */
int cpu_advect(int dim, const int *grid_n, const double *mins, const double *deltas,
               const double *snaps, int n_snap, double snap_t0, double snap_dt, int mode,
               double *x, int64_t P, double t0, double h, int n_steps, int method,
               int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  field_t f;
  memset(&f, 0, sizeof f);
  f.dim = dim;
  for (int a = 0; a < dim; ++a) {
    f.n[a] = grid_n[a];
    f.mins[a] = mins[a];
    f.deltas[a] = deltas[a];
  }
  f.snaps = snaps;
  f.n_snap = n_snap;
  f.snap_t0 = snap_t0;
  f.snap_dt = snap_dt;
  f.mode = mode;
  const int64_t M = P * dim;
  f.idx = (double *)malloc(sizeof(double) * M);
  f.comp = (double *)malloc(sizeof(double) * P);
  f.comp2 = (double *)malloc(sizeof(double) * P);
  double *k1 = (double *)malloc(sizeof(double) * M);
  double *k2 = (double *)malloc(sizeof(double) * M);
  double *k3 = (double *)malloc(sizeof(double) * M);
  double *k4 = (double *)malloc(sizeof(double) * M);
  double *tmp = (double *)malloc(sizeof(double) * M);
  if (!f.idx || !f.comp || !f.comp2 || !k1 || !k2 || !k3 || !k4 || !tmp) return 1;
  int have_prev = 0;
  for (int s = 0; s < n_steps; ++s) {
    const double t = t0 + s * h;
    if (method == 2) {
      velocity(&f, s, t, x, P, k1);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < M; ++i) tmp[i] = x[i] + 0.5 * h * k1[i];
      velocity(&f, s, t + 0.5 * h, tmp, P, k2);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < M; ++i) tmp[i] = x[i] + 0.5 * h * k2[i];
      velocity(&f, s, t + 0.5 * h, tmp, P, k3);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < M; ++i) tmp[i] = x[i] + h * k3[i];
      velocity(&f, s, t + h, tmp, P, k4);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < M; ++i)
        x[i] += (h / 6.0) * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]);
    } else {
      velocity(&f, s, t, x, P, k1);
      if (method == 0 || !have_prev) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < M; ++i) x[i] += h * k1[i];
      } else {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < M; ++i) x[i] += h * (1.5 * k1[i] - 0.5 * k2[i]);
      }
      if (method == 1) {
        memcpy(k2, k1, sizeof(double) * M);
        have_prev = 1;
      }
    }
  }
  free(f.idx);
  free(f.comp);
  free(f.comp2);
  free(k1);
  free(k2);
  free(k3);
  free(k4);
  free(tmp);
  return 0;
}

int cpu_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
