// Shared pieces of the fused MLP kernels (inference in mlp_fused.hip, training in
// mlp_train.hip): shapes, the packed-weight fragment loads, the swizzled LDS activation
// layout, the per-layer MFMA loop over LDS activations and the SiLU epilogue.
// Formulation and register layouts are described at the top of mlp_fused.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cstdint>

#include "common.hpp"

namespace pyftle {
namespace mlpf {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kFeat = 128;  // Fourier features: 64 frequencies, sin and cos
constexpr int kNFreq = kFeat / 2;
constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kMaxDepth = 8;
constexpr int kMaxIn = 4;

// Shape of one instantiation: hidden width W, RT row tiles of 16 per work-group, WV waves
// (each owning kTilesPerWave of the layer's 16-feature tiles for all the rows).
template <int W, int RT, int WV = kWaves>
struct Shape {
  static constexpr int kFragTiles = W / 16;             // 16-feature tiles of a layer
  static constexpr int kTilesPerWave = kFragTiles / WV;
  static constexpr int kRows = 16 * RT;
  static_assert(W % (16 * WV) == 0 && W >= kFeat, "the width splits over the waves");
};

struct FusedArgs {
  const float *x;      // (M, d_in) inputs
  const float *scale;  // (d_in) input scaling before the Fourier projection
  const float *Bf;     // (64, d_in) frequencies, in revolutions per unit
  const __bf16 *W;     // packed hidden weights (pyftle_mlp_pack_fused order), layer after layer
  int w_bytes;         // their size
  const float *bias;   // (depth, 512)
  const float *Wout;   // (O, 512)
  const float *bout;   // (O)
  float *Y;            // (1 + n_tangents, M, O)
  float out_scale;
  int d_in;
  int M;
  int depth;
};

// Byte offset of layer L in the packed weights of width W.
template <int W>
__device__ __forceinline__ int layer_bytes(int L) {
  return L == 0 ? 0 : 2 * (W * kFeat + (L - 1) * W * W);
}

// This lane's 8 values of a fragment of the packed weights: buffer load at the wave-uniform
// byte offset `sofs` (layer, k-step and feature tile, in a scalar register) plus the lane's
// 16-byte piece `vofs`, so no per-load 64-bit address is formed.
__device__ __forceinline__ bf16x8_t load_frag(__amdgpu_buffer_rsrc_t w, int vofs, int sofs) {
  return __builtin_bit_cast(bf16x8_t, __builtin_amdgcn_raw_buffer_load_b128(w, vofs, sofs, 0));
}

// Byte offset of fragment (k-step ks, feature tile ft) inside a packed layer of width W.
template <int W>
constexpr int frag_bytes(int ks, int ft) { return (ks * (W / 16) + ft) * 1024; }

// LDS activations: row r holds W bf16 (W/8 units of 16 bytes); unit u of row r is stored
// at unit u ^ (r % 16), so the 16 rows read by one B fragment land in distinct banks.
template <int W>
__device__ __forceinline__ int act_off(int r, int u) { return r * W + ((u ^ (r & 15)) << 3); }

// The lane index as a value the compiler must take as new at each call: the per-lane
// offsets derived from it are then recomputed where they are used (a few VALU
// instructions) instead of being hoisted out of the tile and layer loops, where dozens of
// them would stay live next to the accumulators and spill.  The statement emits no
// instruction.
__device__ __forceinline__ int opaque_lane(int lane) {
  asm volatile("" : "+v"(lane));
  return lane;
}

// 1 / (1 + e^-z) with the hardware exp2 and reciprocal (about 1 ulp each; the results are
// rounded to bf16 or summed in FP32 afterwards).  e^-z overflows to +inf for z < -88 and the
// reciprocal then gives 0, the correct limit.
__device__ __forceinline__ float sigmoid_f(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.4426950408889634f));
}

// acc[i][t] = (output features of this wave's tile i) x (rows of row tile t), over the K
// inputs held in LDS.  The weight fragments travel through a ring PD k-steps deep: on entry
// ring[p][i] holds tile i of k-step p (p < PD) of this layer; as soon as the MFMAs of a tile
// have been issued at k-step ks its slot ks % PD is refilled with the same tile of k-step
// ks + PD (of this layer, or of the layer at `next`, whose k-steps continue the count), so
// each load has PD k-steps of the wave's MFMAs (interleaved with the other waves on the SIMD)
// to land.  Every layer's k-step count must be a multiple of PD, so the slot of a k-step is
// the same whichever layer prefetched it.
template <int W, int RT, int K, int PD, int WV = kWaves>
__device__ __forceinline__ void layer_mma_ring(__amdgpu_buffer_rsrc_t w, int lbytes, int next_lbytes,
                                               const __bf16 *act,
                                               bf16x8_t (&ring)[PD][Shape<W, RT, WV>::kTilesPerWave],
                                               f32x4_t (&acc)[Shape<W, RT, WV>::kTilesPerWave][RT], int ft0,
                                               int lane_in) {
  constexpr int kTilesPerWave = Shape<W, RT, WV>::kTilesPerWave, kRowTiles = RT;
  constexpr int KS = K / 32;
  static_assert(KS % PD == 0, "the ring depth divides the k-steps of every layer");
  const int lane = opaque_lane(lane_in);
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < kTilesPerWave; ++i)
#pragma unroll
    for (int t = 0; t < kRowTiles; ++t) acc[i][t] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  bf16x8_t bfr[2][kRowTiles];
#pragma unroll
  for (int t = 0; t < kRowTiles; ++t) bfr[0][t] = *reinterpret_cast<const bf16x8_t *>(act + act_off<W>(t * 16 + r, g));
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    if (ks + 1 < KS) {
#pragma unroll
      for (int t = 0; t < kRowTiles; ++t)
        bfr[(ks + 1) & 1][t] =
            *reinterpret_cast<const bf16x8_t *>(act + act_off<W>(t * 16 + r, (ks + 1) * 4 + g));
    }
    __builtin_amdgcn_sched_barrier(0);  // the next k-step's LDS reads go first
#pragma unroll
    for (int i = 0; i < kTilesPerWave; ++i) {
#pragma unroll
      for (int t = 0; t < kRowTiles; ++t)
        acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ring[ks % PD][i], bfr[ks & 1][t], acc[i][t], 0, 0, 0);
      const int kn = ks + PD;
      const int sofs = kn < KS ? lbytes + frag_bytes<W>(kn, ft0 + i) : next_lbytes + frag_bytes<W>(kn - KS, ft0 + i);
      ring[ks % PD][i] = load_frag(w, lane * 16, sofs);
    }
    // keep the next k-step's loads in this one: left free, the scheduler sinks each load
    // to just before its use and the MFMAs wait on L2 latency
    __builtin_amdgcn_sched_barrier(0);
  }
}

// The ring one k-step deep (the inference kernel: its accumulators leave no room for more).
template <int W, int RT, int K>
__device__ __forceinline__ void layer_mma(__amdgpu_buffer_rsrc_t w, int lbytes, int next_lbytes,
                                          const __bf16 *act, bf16x8_t (&ring)[Shape<W, RT>::kTilesPerWave],
                                          f32x4_t (&acc)[Shape<W, RT>::kTilesPerWave][RT], int ft0, int lane_in) {
  constexpr int kTilesPerWave = Shape<W, RT>::kTilesPerWave;
  layer_mma_ring<W, RT, K, 1>(w, lbytes, next_lbytes, act,
                              *reinterpret_cast<bf16x8_t(*)[1][kTilesPerWave]>(&ring), acc, ft0, lane_in);
}

// Inference stores its activations scaled by -log2(e).  The accumulators then hold
// -log2(e) (z - b), so z' = -log2(e) z costs one fma with the bias, sigmoid(z) =
// 1 / (1 + 2^z') needs no multiply, and h' = z' sigmoid(z) = -log2(e) SiLU(z) feeds the next
// layer's unscaled weights consistently (the layer-0 inputs, the Fourier features, are stored
// scaled alike).  SiLU'(z) = s + h (1 - s) is applied up to the constant c = -ln 2, as
// d / c = h' + s (-log2(e) - h') (two operations instead of four): every hidden layer thus
// scales the tangent rows by a further 1 / c, and the output undoes the product, c^depth.
// For z < -88, 2^z' overflows to +inf and the reciprocal gives s = 0, h' = 0, d = 0: the
// limits.
constexpr float kNegLog2e = -1.4426950408889634f;
constexpr float kNegLn2 = -0.6931471805599453f;

// Hidden layer: h' and, for the tangent rows, t (d / c), stored as bf16 over the layer's
// inputs in LDS (every wave has finished reading them).  Lane (r, g) holds features
// 16 ft + 4 g + e of row 16 t + r for its tiles ft.
template <int W, int RT, int NT>
__device__ __forceinline__ void hidden_epilogue(const f32x4_t (&acc)[Shape<W, RT>::kTilesPerWave][RT],
                                                const float *bias_l, __bf16 *act, int ft0, int lane_in) {
  constexpr int kTilesPerWave = Shape<W, RT>::kTilesPerWave;
  constexpr int G = RT / (1 + NT);
  const int lane = opaque_lane(lane_in);
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < kTilesPerWave; ++i) {
    const int f0 = (ft0 + i) * 16 + 4 * g;
    const f32x4_t bv = *reinterpret_cast<const f32x4_t *>(bias_l + f0);
    const int off = act_off<W>(r, f0 >> 3) + (f0 & 7);  // row tile t: + 16 t rows
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int b0 = q * (1 + NT);
      bf16x4_t h, tg[NT > 0 ? NT : 1];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float z = fmaf(bv[e], kNegLog2e, acc[i][b0][e]);
        const float s = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z));
        const float hp = z * s;
        h[e] = (__bf16)hp;
        if constexpr (NT > 0) {
          const float dd = fmaf(s, kNegLog2e - hp, hp);
#pragma unroll
          for (int a = 0; a < NT; ++a) tg[a][e] = (__bf16)(acc[i][b0 + 1 + a][e] * dd);
        }
      }
      *reinterpret_cast<bf16x4_t *>(act + off + b0 * 16 * W) = h;
      if constexpr (NT > 0) {
#pragma unroll
        for (int a = 0; a < NT; ++a) *reinterpret_cast<bf16x4_t *>(act + off + (b0 + 1 + a) * 16 * W) = tg[a];
      }
    }
  }
}

}  // namespace mlpf
}  // namespace pyftle
