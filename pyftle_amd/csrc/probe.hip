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

// Issue costs of the fused MLP kernels' instruction mix on one SIMD: MFMA 16x16x32 bf16,
// FP32 multiply, exp2 and reciprocal, alone (one wave per SIMD) and with an MFMA wave and
// a VALU wave sharing the SIMD (waves w and w + 4 of an 8-wave work-group share a SIMD), or
// interleaved in one wave.  Each work-group reserves 96 KB of LDS, so one work-group runs per
// CU; a wave's s_memtime cycles over `iters` iterations of 16 slots go to cycles[2 b + k]
// (k = 0: waves 0-3, k = 1: waves 4-7).
enum { kIssueMfma = 0, kIssueMul = 1, kIssueExp = 2, kIssueRcp = 3, kIssueMfmaMul = 4 };

template <int OP>
__device__ __forceinline__ float issue_body(int iters) {
  typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
  typedef float f32x4 __attribute__((ext_vector_type(4)));
  const int tid = threadIdx.x;
  bf16x8 a, b;
  for (int j = 0; j < 8; ++j) {
    a[j] = (__bf16)(0.001f * (tid + j));
    b[j] = (__bf16)(0.002f * (tid - j));
  }
  f32x4 acc[8];
  float x[16];
  for (int j = 0; j < 8; ++j) acc[j] = f32x4{0.001f * j, 0.002f * tid, 0.0f, 1.0f * j};
  for (int j = 0; j < 16; ++j) x[j] = 0.25f + 0.001f * j + 0.0001f * (tid & 15);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (OP == kIssueMfma || OP == kIssueMfmaMul) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
#pragma unroll
      for (int v = 2 * j; v < 2 * j + 2; ++v) {
        if (OP == kIssueMul || OP == kIssueMfmaMul) x[v] *= 0.999f;
        if (OP == kIssueExp) x[v] = __builtin_amdgcn_exp2f(-x[v]);
        if (OP == kIssueRcp) x[v] = __builtin_amdgcn_rcpf(x[v]);
      }
      if (OP == kIssueMfmaMul) {  // one MFMA, then its two multiplies
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
      }
    }
  }
  float s = 0.0f;
  for (int j = 0; j < 8; ++j) s += acc[j][0] + acc[j][3];
  for (int j = 0; j < 16; ++j) s += x[j];
  return s;
}

template <int OP>
__device__ __forceinline__ void issue_timed(int iters, unsigned long long *cycles, float *out) {
  const unsigned long long c0 = __builtin_amdgcn_s_memtime();
  const float s = issue_body<OP>(iters);
  const unsigned long long c1 = __builtin_amdgcn_s_memtime();
  out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
  if ((threadIdx.x & 255) == 0) cycles[2 * blockIdx.x + (threadIdx.x >> 8)] = c1 - c0;
}

// one wave per SIMD
template <int OP>
__global__ void __launch_bounds__(256, 2) issue_one_kernel(int iters, unsigned long long *cycles, float *out) {
  extern __shared__ float pad[];  // 96 KB at launch: one work-group per CU
  if (iters < 0) pad[threadIdx.x] = 0.0f;
  issue_timed<OP>(iters, cycles, out);
}

// two waves per SIMD: waves 0-3 MFMA only, waves 4-7 run OP
template <int OP>
__global__ void __launch_bounds__(512, 1) issue_pair_kernel(int iters, unsigned long long *cycles, float *out) {
  extern __shared__ float pad[];
  if (iters < 0) pad[threadIdx.x] = 0.0f;
  if (threadIdx.x < 256) issue_timed<kIssueMfma>(iters, cycles, out);
  else issue_timed<OP>(iters, cycles, out);
}

}  // namespace pyftle

using namespace pyftle;

// mode: 0 MFMA, 1 mul, 2 exp2, 3 rcp, 4 MFMA and mul interleaved in one wave (one wave per
// SIMD); 5 MFMA wave | mul wave, 6 MFMA wave | exp2 wave, 7 MFMA wave | rcp wave (two per SIMD)
extern "C" int pyftle_probe_issue_rates(int32_t mode, int32_t blocks, int32_t iters, unsigned long long *cycles,
                                        float *out, void *stream) {
  if (!cycles || !out) return fail("probe_issue_rates: NULL pointer");
  if (blocks < 1 || iters < 1) return fail("probe_issue_rates: blocks and iters must be positive");
  hipStream_t st = as_stream(stream);
  constexpr size_t kLds = 96 * 1024;
  auto go = [&](auto kern, int threads) -> int {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)kLds) != hipSuccess)
      return fail("probe_issue_rates: cannot reserve 96 KB of LDS");
    kern<<<blocks, threads, kLds, st>>>(iters, cycles, out);
    return check_launch("probe_issue_rates");
  };
  switch (mode) {
    case 0: return go(issue_one_kernel<kIssueMfma>, 256);
    case 1: return go(issue_one_kernel<kIssueMul>, 256);
    case 2: return go(issue_one_kernel<kIssueExp>, 256);
    case 3: return go(issue_one_kernel<kIssueRcp>, 256);
    case 4: return go(issue_one_kernel<kIssueMfmaMul>, 256);
    case 5: return go(issue_pair_kernel<kIssueMul>, 512);
    case 6: return go(issue_pair_kernel<kIssueExp>, 512);
    case 7: return go(issue_pair_kernel<kIssueRcp>, 512);
    default: return fail("probe_issue_rates: mode must be in [0, 7]");
  }
}

extern "C" int pyftle_probe_load_rate(const void *buf, int32_t blocks, int32_t iters, float *out, void *stream) {
  if (!buf || !out) return fail("probe_load_rate: NULL pointer");
  if (blocks < 1 || iters < 4 || iters % 4 != 0) return fail("probe_load_rate: blocks >= 1, iters a multiple of 4");
  load_rate_kernel<<<blocks, 256, 0, as_stream(stream)>>>(static_cast<const char *>(buf), iters, out);
  return check_launch("probe_load_rate");
}
