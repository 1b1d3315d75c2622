// Library-level C ABI: version, thread-local error text, field validation.
#include <math.h>

#include <cstdio>
#include <string>

#include "common.hpp"

namespace pyftle {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(const std::string &msg, int code) {
  set_error(msg);
  return code;
}

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    return fail(std::string(what) + ": " + hipGetErrorString(e), PYFTLE_EHIP);
  }
  return PYFTLE_OK;
}

static int validate_scatter(const pyftle_field *f) {
  const pyftle_scatter *sc = f->scatter;
  if (!sc) return fail("scattered-data modes need field.scatter");
  if (f->layout != PYFTLE_LAYOUT_SINGLE) return fail("scattered-data modes need the SINGLE layout");
  if (sc->n_points < 1 || sc->n_points >= (int64_t)1 << 31) return fail("scatter.n_points out of range");
  if (f->n[0] != sc->n_points || f->n[1] != 1 || (f->dim == 3 && f->n[2] != 1))
    return fail("scattered-data fields have n = (n_points, 1, 1)");
  const bool simplex_grid = f->interp != PYFTLE_INTERP_SCATTER_NEAREST;
  const int32_t *dims = simplex_grid ? sc->simplex_dims : sc->point_dims;
  const double *size = simplex_grid ? sc->simplex_size : sc->point_size;
  int64_t buckets = 1;
  for (int a = 0; a < f->dim; ++a) {
    if (dims[a] < 1) return fail("scatter bucket grid dims must be >= 1");
    if (!(size[a] > 0.0) || !isfinite(size[a])) return fail("scatter bucket sizes must be finite and > 0");
    buckets *= dims[a];
  }
  if (buckets >= (int64_t)1 << 31) return fail("scatter bucket grid too large");
  if (f->interp == PYFTLE_INTERP_SCATTER_LINEAR || f->interp == PYFTLE_INTERP_SCATTER_CUBIC) {
    if (sc->n_simplices < 1 || !sc->simplices || !sc->transform || !sc->hull || !sc->simplex_offsets ||
        !sc->simplex_ids)
      return fail("LINEAR/CUBIC scatter needs simplices, transform, hull flags and the simplex buckets");
    if (f->interp == PYFTLE_INTERP_SCATTER_CUBIC) {
      if (f->dim != 2) return fail("cubic interpolator is only valid for 2D cases");
      if (!sc->points || !sc->ct_g) return fail("CUBIC scatter needs the points and the cross-edge factors");
    }
  } else {
    if (!sc->points || !sc->point_offsets || !sc->point_ids)
      return fail("NEAREST scatter needs the points and the point buckets");
  }
  return PYFTLE_OK;
}

int validate_field(const pyftle_field *f) {
  if (!f) return fail("field is NULL");
  if (f->dim != 2 && f->dim != 3) return fail("field.dim must be 2 or 3");
  const bool scatter = is_scatter(f->interp);
  if (scatter) {
    int rc = validate_scatter(f);
    if (rc) return rc;
  }
  for (int a = 0; a < f->dim && !scatter; ++a) {
    if (f->n[a] < 2) return fail("field.n must be >= 2 on every axis");
    if (!(f->spacing[a] != 0.0) || !isfinite(f->spacing[a]))
      return fail("field.spacing must be finite and non-zero");
    if (!isfinite(f->origin[a])) return fail("field.origin must be finite");
  }
  if (f->dtype != PYFTLE_F32 && f->dtype != PYFTLE_F64) return fail("field.dtype must be F32 or F64");
  if (f->layout < PYFTLE_LAYOUT_SINGLE || f->layout > PYFTLE_LAYOUT_CELL_PAIR)
    return fail("field.layout must be SINGLE, PAIR, CELL or CELL_PAIR");
  if (layout_cell(f->layout) && (f->dtype != PYFTLE_F32 || f->interp != PYFTLE_INTERP_GINTERP))
    return fail("the CELL layouts hold FP32 GridInterp-mode fields");
  if (f->interp < PYFTLE_INTERP_GINTERP || f->interp > PYFTLE_INTERP_SCATTER_CUBIC)
    return fail("field.interp must be one of PYFTLE_INTERP_*");
  if (f->interp != PYFTLE_INTERP_GINTERP && !scatter) {
    if (f->layout != PYFTLE_LAYOUT_SINGLE) return fail("RegularGridInterpolator modes need the SINGLE layout");
    for (int a = 0; a < f->dim; ++a) {
      if (f->periodic[a]) return fail("RegularGridInterpolator modes have no periodic axes");
      if (!isfinite(f->upper[a]) || !(f->spacing[a] > 0.0) || !(f->upper[a] > f->origin[a]))
        return fail("RegularGridInterpolator modes need ascending axes: origin < upper, spacing > 0");
      if (f->interp == PYFTLE_INTERP_RGI_CUBIC && f->n[a] < 4)
        return fail("method cubic requires at least 4 points per dimension");
    }
  }
  if (f->ring_slots < 1) return fail("field.ring_slots must be >= 1");
  if (f->n_snapshots < 1) return fail("field.n_snapshots must be >= 1");
  if (!f->data) return fail("field.data is NULL");
  int64_t nodes = padded_nodes(f);
  if (nodes > (int64_t)INT32_MAX / 2) return fail("field grid too large (> 2^30 nodes)");
  if (!scatter && (axis_records(f, 0) >= (1 << 24) || (int64_t)axis_records(f, 0) * axis_records(f, 1) >= (1 << 24)))
    return fail("field rows / slices must hold fewer than 2^24 nodes");
  int64_t need = nodes * record_bytes(f->dim, f->layout, f->dtype, f->interp);
  if (f->slot_stride < need) return fail("field.slot_stride smaller than one slot of records");
  if (f->slot_stride % 16 != 0) return fail("field.slot_stride must be a multiple of 16 bytes");
  if ((reinterpret_cast<uintptr_t>(f->data) % 32) != 0) return fail("field.data must be 32-byte aligned");
  if (f->dtype == PYFTLE_F32 && f->slot_stride > (int64_t)INT32_MAX)
    return fail("FP32 field slots are limited to 2 GiB (32-bit buffer offsets)");
  return PYFTLE_OK;
}

int make_field_dev(const pyftle_field *f, FieldDev *d) {
  int rc = validate_field(f);
  if (rc) return rc;
  d->data = reinterpret_cast<const char *>(f->data);
  d->slot_stride = f->slot_stride;
  d->ring = f->ring_slots;
  d->nsnap = f->n_snapshots;
  for (int a = 0; a < 3; ++a) {
    const bool used = a < f->dim;
    d->n[a] = used ? f->n[a] : 1;
    d->periodic[a] = used ? (f->periodic[a] ? 1 : 0) : 0;
    d->origin[a] = used ? f->origin[a] : 0.0;
    d->spacing[a] = used ? f->spacing[a] : 1.0;
    d->inv_spacing[a] = 1.0 / d->spacing[a];
    d->n_d[a] = (double)d->n[a];
    d->inv_n_d[a] = 1.0 / d->n_d[a];
    d->origin_f[a] = (float)d->origin[a];
    d->inv_spacing_f[a] = (float)(1.0 / d->spacing[a]);
    d->n_f[a] = (float)d->n[a];
    d->inv_n_f[a] = (float)(1.0 / d->n[a]);
  }
  for (int a = 0; a < 3; ++a) {
    d->per_n[a] = d->periodic[a] ? d->n[a] : 0;
    d->lim[a] = d->periodic[a] ? d->n[a] : d->n[a] - 1;
  }
  d->px = axis_records(f, 0);
  d->pxy = d->px * (f->dim >= 2 ? axis_records(f, 1) : 1);
  d->t0 = f->t0;
  d->dts = f->dt_snap;
  d->interp = f->interp;
  for (int a = 0; a < 3; ++a) d->upper[a] = a < f->dim ? f->upper[a] : 0.0;
  d->points = nullptr;
  d->simplices = nullptr;
  d->transform = nullptr;
  d->hull = nullptr;
  d->s_off = d->s_ids = d->p_off = d->p_ids = nullptr;
  d->npts = 0;
  for (int a = 0; a < 3; ++a) {
    d->s_dims[a] = d->p_dims[a] = 1;
    d->s_org[a] = d->p_org[a] = 0.0;
    d->s_inv[a] = d->p_inv[a] = 1.0;
    d->p_size[a] = 1.0;
  }
  d->ct_g = nullptr;
  if (f->scatter && is_scatter(f->interp)) {
    const pyftle_scatter *sc = f->scatter;
    d->points = sc->points;
    d->simplices = sc->simplices;
    d->transform = sc->transform;
    d->hull = sc->hull;
    d->s_off = sc->simplex_offsets;
    d->s_ids = sc->simplex_ids;
    d->p_off = sc->point_offsets;
    d->p_ids = sc->point_ids;
    d->ct_g = sc->ct_g;
    d->npts = sc->n_points;
    for (int a = 0; a < f->dim; ++a) {
      d->s_dims[a] = sc->simplex_dims[a] > 0 ? sc->simplex_dims[a] : 1;
      d->s_org[a] = sc->simplex_origin[a];
      d->s_inv[a] = sc->simplex_size[a] > 0.0 ? 1.0 / sc->simplex_size[a] : 1.0;
      d->p_dims[a] = sc->point_dims[a] > 0 ? sc->point_dims[a] : 1;
      d->p_org[a] = sc->point_origin[a];
      d->p_size[a] = sc->point_size[a] > 0.0 ? sc->point_size[a] : 1.0;
      d->p_inv[a] = 1.0 / d->p_size[a];
    }
  }
  return PYFTLE_OK;
}

}  // namespace pyftle

extern "C" {

int pyftle_version(void) { return 10000 * 0 + 100 * 1 + 0; }

const char *pyftle_last_error(void) { return pyftle::g_last_error.c_str(); }

int pyftle_device_sync(void) {
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return pyftle::fail(std::string("hipDeviceSynchronize: ") + hipGetErrorString(e), PYFTLE_EHIP);
  return PYFTLE_OK;
}

}  // extern "C"
