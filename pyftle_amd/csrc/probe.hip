// Measurement kernel (no reference counterpart): the peak rate of 16-byte-per-lane
// wave loads through a CU's vector-memory path (texture address, L1, texture data
// return), against which the gather-bound advection kernel's issued record loads are
// set.  Every lane of every wave reads 16 bytes at 16 * lane of a 1 KiB row of a
// 16 KiB window (8 whole 128-byte lines per wave load, L1 hits after the first
// pass); four independent accumulators keep four loads in flight per lane.
#include <string>

#include "advect_impl.hpp"

namespace pyftle {

__global__ void __launch_bounds__(256) load_rate_kernel(const char *buf, int iters, float *out) {
  const __amdgpu_buffer_rsrc_t r = make_rsrc(buf, 16384);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  float acc[4][4] = {};
  for (int k = 0; k < iters; k += 4) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int row = (k + u + wave) & 15;
      auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (row << 10) + lane * 16, 0, 0);
      float f[4];
      __builtin_memcpy(f, &v, 16);
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[u][e] += f[e];
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int e = 0; e < 4; ++e) s += acc[u][e];
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace pyftle

using namespace pyftle;

extern "C" int pyftle_probe_load_rate(const void *buf, int32_t blocks, int32_t iters, float *out, void *stream) {
  if (!buf || !out) return fail("probe_load_rate: NULL pointer");
  if (blocks < 1 || iters < 4 || iters % 4 != 0) return fail("probe_load_rate: blocks >= 1, iters a multiple of 4");
  load_rate_kernel<<<blocks, 256, 0, as_stream(stream)>>>(static_cast<const char *>(buf), iters, out);
  return check_launch("probe_load_rate");
}
