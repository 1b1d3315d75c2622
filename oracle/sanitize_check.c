/*
 * Driver for running the oracle's C restatement (cpu_path.c) under the address and
 * undefined-behaviour sanitizers (test infrastructure; built and run by
 * tests/test_oracle.py, never by the product).
 *
 * Usage: sanitize_check IN OUT
 *   IN:  int64 header [dim, nx, ny, nz, n_snap, P, n_steps, npts], then doubles
 *        mins[3], deltas[3], snap_t0, snap_dt, t0, h, the snapshots
 *        (n_snap, dim, nx*ny*nz), the particles (P, dim) and the index-space
 *        interpolation points (npts, dim).
 *   OUT: doubles: interp of snapshot 0's first component at the points (npts), then
 *        the advected particles (P, dim) for method euler, ab2, rk4 (outer) and
 *        mode frozen, time-linear (inner).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void cpu_interp2d(const double *v, int nx, int ny, const double *pts, int64_t n, double *out);
void cpu_interp3d(const double *v, int X, int Y, int Z, const double *pts, int64_t n, double *out);
int cpu_advect(int dim, const int *grid_n, const double *mins, const double *deltas,
               const double *snaps, int n_snap, double snap_t0, double snap_dt, int mode,
               double *x, int64_t P, double t0, double h, int n_steps, int method, int threads);

static double *read_doubles(FILE *f, int64_t n) {
  double *p = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  if (!p || fread(p, sizeof(double), (size_t)n, f) != (size_t)n) {
    fprintf(stderr, "short input\n");
    exit(2);
  }
  return p;
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
    return 2;
  }
  FILE *fi = fopen(argv[1], "rb");
  if (!fi) return 2;
  int64_t hd[8];
  if (fread(hd, sizeof(int64_t), 8, fi) != 8) return 2;
  const int dim = (int)hd[0], nx = (int)hd[1], ny = (int)hd[2], nz = (int)hd[3];
  const int n_snap = (int)hd[4], n_steps = (int)hd[6];
  const int64_t P = hd[5], npts = hd[7];
  const int64_t grid = (int64_t)nx * ny * (dim == 3 ? nz : 1);
  double *sc = read_doubles(fi, 10);
  double *snaps = read_doubles(fi, (int64_t)n_snap * dim * grid);
  double *x0 = read_doubles(fi, P * dim);
  double *pts = read_doubles(fi, npts * dim);
  fclose(fi);

  FILE *fo = fopen(argv[2], "wb");
  if (!fo) return 2;
  double *out = (double *)malloc((size_t)(npts > 0 ? npts : 1) * sizeof(double));
  if (dim == 2)
    cpu_interp2d(snaps, nx, ny, pts, npts, out);
  else
    cpu_interp3d(snaps, nx, ny, nz, pts, npts, out);
  fwrite(out, sizeof(double), (size_t)npts, fo);
  const int n[3] = {nx, ny, dim == 3 ? nz : 1};
  double *x = (double *)malloc((size_t)(P * dim > 0 ? P * dim : 1) * sizeof(double));
  for (int method = 0; method < 3; ++method) {
    for (int mode = 0; mode < 2; ++mode) {
      memcpy(x, x0, (size_t)(P * dim) * sizeof(double));
      if (cpu_advect(dim, n, sc, sc + 3, snaps, n_snap, sc[6], sc[7], mode, x, P, sc[8], sc[9], n_steps,
                     method, 1) != 0)
        return 3;
      fwrite(x, sizeof(double), (size_t)(P * dim), fo);
    }
  }
  fclose(fo);
  free(out);
  free(x);
  free(sc);
  free(snaps);
  free(x0);
  free(pts);
  return 0;
}
