// Stage arithmetic of the integrators (integrate.py:16-109) for velocities the
// caller evaluates itself: the analytical strategy calls the user's
// velocity_fn(t, points) on device tensors at every stage (interpolate.py:
// 765-812), and these element-wise kernels form the stage positions and the
// update in the reference's operation order (FP64 bit-exact with it for the
// same velocities; contraction is off for the whole library).
#include <algorithm>
#include <string>

#include "common.hpp"

namespace pyftle {

template <typename T, int OP>
__global__ void __launch_bounds__(256) combine_kernel(int64_t n, T c, T *x, const T *k1, const T *k2,
                                                      const T *k3, const T *k4, T *out) {
  // no __restrict__: AB2 passes the previous velocity as both k2 and out
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (OP == PYFTLE_COMBINE_STAGE) {
      out[i] = x[i] + c * k1[i];  // tmp = x; tmp += (c*h) * k   (integrate.py:327-338)
    } else if (OP == PYFTLE_COMBINE_RK4) {
      // positions += (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)   (integrate.py:109)
      x[i] = x[i] + c * (((k1[i] + (T)2 * k2[i]) + (T)2 * k3[i]) + k4[i]);
    } else if (OP == PYFTLE_COMBINE_EULER) {
      const T v = k1[i];
      x[i] = x[i] + c * v;  // positions += h * velocity   (integrate.py:44)
      if (out) out[i] = v;  // AB2's previous velocity after its Euler start (integrate.py:252-257)
    } else {                // PYFTLE_COMBINE_AB2
      const T v = k1[i];
      x[i] = x[i] + c * ((T)1.5 * v - (T)0.5 * k2[i]);  // integrate.py:78
      if (out) out[i] = v;
    }
  }
}

template <typename T>
static int launch(int op, int64_t n, double c, void *x, const void *k1, const void *k2, const void *k3,
                  const void *k4, void *out, hipStream_t st) {
  const int threads = 256;
  const int64_t want = (n + threads - 1) / threads;
  const int blocks = (int)std::min<int64_t>(want, (int64_t)device_cu_count() * 16);
  T *xx = static_cast<T *>(x);
  const T *a = static_cast<const T *>(k1), *b = static_cast<const T *>(k2);
  const T *cc = static_cast<const T *>(k3), *d = static_cast<const T *>(k4);
  T *o = static_cast<T *>(out);
  const T cv = (T)c;
  switch (op) {
    case PYFTLE_COMBINE_STAGE:
      combine_kernel<T, PYFTLE_COMBINE_STAGE><<<blocks, threads, 0, st>>>(n, cv, xx, a, b, cc, d, o);
      break;
    case PYFTLE_COMBINE_RK4:
      combine_kernel<T, PYFTLE_COMBINE_RK4><<<blocks, threads, 0, st>>>(n, cv, xx, a, b, cc, d, o);
      break;
    case PYFTLE_COMBINE_EULER:
      combine_kernel<T, PYFTLE_COMBINE_EULER><<<blocks, threads, 0, st>>>(n, cv, xx, a, b, cc, d, o);
      break;
    default:
      combine_kernel<T, PYFTLE_COMBINE_AB2><<<blocks, threads, 0, st>>>(n, cv, xx, a, b, cc, d, o);
      break;
  }
  return check_launch("stage_combine");
}

}  // namespace pyftle

using namespace pyftle;

extern "C" int pyftle_stage_combine(int32_t op, int32_t dtype, int64_t n, double c, void *x, const void *k1,
                                    const void *k2, const void *k3, const void *k4, void *out,
                                    void *stream) {
  if (op < PYFTLE_COMBINE_STAGE || op > PYFTLE_COMBINE_AB2) return fail("stage_combine: unknown op");
  if (dtype != PYFTLE_F32 && dtype != PYFTLE_F64) return fail("stage_combine: dtype must be F32 or F64");
  if (n < 0) return fail("stage_combine: negative size");
  if (n == 0) return PYFTLE_OK;
  if (!x || !k1) return fail("stage_combine: x and k1 are required");
  if (op == PYFTLE_COMBINE_STAGE && !out) return fail("stage_combine: STAGE needs out");
  if (op == PYFTLE_COMBINE_RK4 && (!k2 || !k3 || !k4)) return fail("stage_combine: RK4 needs k1..k4");
  if (op == PYFTLE_COMBINE_AB2 && !k2) return fail("stage_combine: AB2 needs the previous velocity k2");
  hipStream_t st = as_stream(stream);
  return dtype == PYFTLE_F64 ? launch<double>(op, n, c, x, k1, k2, k3, k4, out, st)
                             : launch<float>(op, n, c, x, k1, k2, k3, k4, out, st);
}
