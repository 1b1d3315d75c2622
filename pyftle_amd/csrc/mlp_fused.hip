// Fused surrogate inference with the activations in LDS: Fourier features -> depth x
// (W, SiLU) -> O outputs, with forward-mode tangents, in one kernel, for hidden width
// W = 256 or 512.
//
// No reference counterpart: the learned flow-map approximation of the RK4 path
// (integrate.py:265-342).  The layer-by-layer kernels move 2 W bytes per row and layer
// through HBM, below the MFMA ridge at these widths.  Here a work-group keeps the
// activations of its rows in LDS for the whole network, and only the weights move: from L2
// straight into registers.
//
// Formulation: out^T = A in^T per layer (A the layer's weights) with
// v_mfma_f32_16x16x32_bf16.  The weights are the MFMA A operand (16 output features x 32
// inputs per fragment), the activations the B operand (32 inputs x 16 rows) read from
// LDS.  Wave w of 4 owns output features
// W/4 w .. W/4 (w + 1) - 1 (W/64 tiles of 16) for all 16 RT rows of the work-group (RT
// row tiles of 16), so each weight fragment feeds RT MFMAs and each activation fragment
// W/64.  The weights are packed fragment by fragment (pyftle_mlp_pack_fused): the 64 lanes'
// 16-byte pieces of one fragment are 1 KB contiguous, one coalesced load per wave.
//
// Rows of a tile: S seeds x (1 + NT) row blocks of 16; row tile q (1 + NT) + a holds seeds
// 16 q .. 16 q + 15 of the tile, a = 0 the primal rows, a > 0 the derivative along input
// a - 1.  A lane holds the same (seed, feature) pair for the primal and its tangent rows,
// so SiLU'(z) of the tangent rows uses the primal pre-activation from registers, in FP32.
//
// Per layer: the MFMAs over K (the weight fragments of the next k-step in flight; the
// next layer's first ones are issued during the last k-step), a barrier, the epilogue
// SiLU(z + b) written back to LDS in place as bf16 (scaled by -log2(e); mlp_fused.hpp
// gives the bookkeeping) at raised wave priority, a barrier.
// Two work-groups share a CU (57-75 KB of LDS each), so one's epilogue runs under the
// other's MFMAs.  The output layer (O <= 3) is FP32 on the last hidden layer's registers,
// reduced over lanes and waves.

#include "mlp_fused.hpp"

namespace pyftle {
namespace {

using namespace mlpf;

// Last hidden layer: this lane's partial sums of the output layer, in FP32, with c Wout on
// the scaled activations (the primal rows come out exact; the tangent rows still carry
// the factor c^-depth).
template <int W, int RT, int NT, int O>
__device__ __forceinline__ void output_partials(const f32x4_t (&acc)[Shape<W, RT>::kTilesPerWave][RT],
                                                const float *bias_l, const float *wout, int ft0, int lane_in,
                                                float (&yp)[RT][O]) {
  constexpr int kTilesPerWave = Shape<W, RT>::kTilesPerWave, kRowTiles = RT;
  constexpr int G = RT / (1 + NT);
  const int g = opaque_lane(lane_in) >> 4;
#pragma unroll
  for (int t = 0; t < kRowTiles; ++t)
#pragma unroll
    for (int o = 0; o < O; ++o) yp[t][o] = 0.0f;
#pragma unroll
  for (int i = 0; i < kTilesPerWave; ++i) {
    const int f0 = (ft0 + i) * 16 + 4 * g;
    const f32x4_t bv = *reinterpret_cast<const f32x4_t *>(bias_l + f0);
    f32x4_t wv[O];
#pragma unroll
    for (int o = 0; o < O; ++o) wv[o] = *reinterpret_cast<const f32x4_t *>(wout + o * W + f0) * kNegLn2;
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int b0 = q * (1 + NT);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float z = fmaf(bv[e], kNegLog2e, acc[i][b0][e]);
        const float s = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z));
        const float hp = z * s;
#pragma unroll
        for (int o = 0; o < O; ++o) yp[b0][o] = fmaf(hp, wv[o][e], yp[b0][o]);
        if constexpr (NT > 0) {
          const float dd = fmaf(s, kNegLog2e - hp, hp);
#pragma unroll
          for (int a = 1; a <= NT; ++a) {
            const float tv = acc[i][b0 + a][e] * dd;
#pragma unroll
            for (int o = 0; o < O; ++o) yp[b0 + a][o] = fmaf(tv, wv[o][e], yp[b0 + a][o]);
          }
        }
      }
    }
  }
}

// Fourier features of the tile's seeds (and their input derivatives), scaled by -log2(e),
// into LDS rows: features [sin | cos] of the 64 frequencies, one task per (seed, block of 8
// frequencies); dfreq_s already holds 2 pi B scale (-log2(e)).
template <int W, int RT, int NT>
__device__ __forceinline__ void fourier_rows(const FusedArgs &a, int64_t tile, const float *freq_s,
                                             const float *dfreq_s, __bf16 *act, int tid) {
  constexpr int S = Shape<W, RT>::kRows / (1 + NT);
  for (int task = tid; task < S * 8; task += kThreads) {
    const int j = task >> 3, fb = task & 7;
    const int64_t seed = tile * S + j;
    float xin[kMaxIn];
#pragma unroll
    for (int c = 0; c < kMaxIn; ++c) xin[c] = (seed < a.M && c < a.d_in) ? a.x[seed * a.d_in + c] : 0.0f;
    bf16x8_t sn, cs, tsn[NT > 0 ? NT : 1], tcs[NT > 0 ? NT : 1];
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const int f = fb * 8 + jj;
      const f32x4_t fr = *reinterpret_cast<const f32x4_t *>(freq_s + f * kMaxIn);
      float ph = 0.0f;  // revolutions
#pragma unroll
      for (int c = 0; c < kMaxIn; ++c) ph += fr[c] * xin[c];
      ph -= rintf(ph);
      const float s = __builtin_amdgcn_sinf(ph), c = __builtin_amdgcn_cosf(ph);
      sn[jj] = (__bf16)(s * kNegLog2e);
      cs[jj] = (__bf16)(c * kNegLog2e);
      if constexpr (NT > 0) {
        const f32x4_t dq = *reinterpret_cast<const f32x4_t *>(dfreq_s + f * kMaxIn);
#pragma unroll
        for (int a2 = 0; a2 < NT; ++a2) {
          tsn[a2][jj] = (__bf16)(c * dq[a2]);
          tcs[a2][jj] = (__bf16)(-s * dq[a2]);
        }
      }
    }
    const int row0 = (j >> 4) * (1 + NT) * 16 + (j & 15);
    *reinterpret_cast<bf16x8_t *>(act + act_off<W>(row0, fb)) = sn;
    *reinterpret_cast<bf16x8_t *>(act + act_off<W>(row0, kNFreq / 8 + fb)) = cs;
    if constexpr (NT > 0) {
#pragma unroll
      for (int a2 = 0; a2 < NT; ++a2) {
        const int row = row0 + (1 + a2) * 16;
        *reinterpret_cast<bf16x8_t *>(act + act_off<W>(row, fb)) = tsn[a2];
        *reinterpret_cast<bf16x8_t *>(act + act_off<W>(row, kNFreq / 8 + fb)) = tcs[a2];
      }
    }
  }
}

// Persistent: work-group b handles tiles b, b + gridDim.x, ...
template <int W, int RT, int NT, int O>
__global__ void __launch_bounds__(kThreads, 2) fused_mlp_kernel(FusedArgs a) {
  using Sh = Shape<W, RT>;
  constexpr int kTilesPerWave = Sh::kTilesPerWave, kRows = Sh::kRows;
  constexpr int S = kRows / (1 + NT);  // seeds per tile
  static_assert(RT % (1 + NT) == 0, "row tiles hold whole seed groups");
  __shared__ __attribute__((aligned(16))) __bf16 act[kRows * W];
  __shared__ __attribute__((aligned(16))) float freq_s[kNFreq * kMaxIn];   // B * scale, zero-padded
  __shared__ __attribute__((aligned(16))) float dfreq_s[kNFreq * kMaxIn];  // 2 pi B * scale * -log2(e)
  __shared__ float red[kWaves][kRows][4];                                  // output partials per wave

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ft0 = wave * kTilesPerWave;
  const int64_t ntiles = ((int64_t)a.M + S - 1) / S;
  {
    const int q = tid >> 2, c = tid & 3;  // kThreads == kNFreq * kMaxIn
    float fv = 0.0f, sc = 0.0f;
    if (c < a.d_in) {
      fv = a.Bf[q * a.d_in + c];
      sc = a.scale[c];
    }
    freq_s[tid] = fv * sc;
    dfreq_s[tid] = 6.283185307179586f * fv * sc * kNegLog2e;
  }
  float tscale = a.out_scale;  // c^depth undoes the tangents' scale
  for (int L = 0; L < a.depth; ++L) tscale *= kNegLn2;
  const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc((void *)a.W, 0, a.w_bytes, 0x00020000);
  bf16x8_t ring[kTilesPerWave];
#pragma unroll
  for (int i = 0; i < kTilesPerWave; ++i) ring[i] = load_frag(wr, lane * 16, frag_bytes<W>(0, ft0 + i));
  __syncthreads();
  // odd work-groups start about 7 us late, so the two work-groups of a CU tend to
  // alternate MFMA and epilogue phases instead of contending in step (measured: primal-only
  // width 512 runs at 1.19-1.21 PFLOP/s staggered, 1.07-1.19 in step)
  if (blockIdx.x & 1) {
    __builtin_amdgcn_s_sleep(127);
    __builtin_amdgcn_s_sleep(127);
  }

  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    fourier_rows<W, RT, NT>(a, tile, freq_s, dfreq_s, act, tid);
    __syncthreads();
    f32x4_t acc[kTilesPerWave][RT];
    for (int L = 0; L < a.depth; ++L) {
      const int lb = layer_bytes<W>(L), nlb = layer_bytes<W>(L + 1 < a.depth ? L + 1 : 0);
      if (L == 0) layer_mma<W, RT, kFeat>(wr, lb, nlb, act, ring, acc, ft0, lane);
      else layer_mma<W, RT, W>(wr, lb, nlb, act, ring, acc, ft0, lane);
      const float *bl = a.bias + (int64_t)L * W;
      if (L + 1 < a.depth) {
        __syncthreads();  // every wave has read this layer's inputs
        // the epilogue's VALU work issues ahead of the MFMAs of the CU's other work-group
        // (primal-only width 256: 0.85 -> 1.0 PFLOP/s)
        __builtin_amdgcn_s_setprio(1);
        hidden_epilogue<W, RT, NT>(acc, bl, act, ft0, lane);
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
      } else {
        float yp[RT][O];
        output_partials<W, RT, NT, O>(acc, bl, a.Wout, ft0, lane, yp);
        // the 4 lanes of a row (lane % 16 equal) hold disjoint feature subsets
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
          for (int o = 0; o < O; ++o) {
            float v = yp[t][o];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            yp[t][o] = v;
          }
        if (lane < 16) {
#pragma unroll
          for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int o = 0; o < O; ++o) red[wave][t * 16 + lane][o] = yp[t][o];
        }
        // also orders this tile's reads of `act` before the next tile's Fourier rows
        __syncthreads();
        for (int idx = tid; idx < kRows * 4; idx += kThreads) {
          const int n = idx >> 2, o = idx & 3;
          const int t = n >> 4, q = t / (1 + NT), a2 = t % (1 + NT);
          const int64_t seed = tile * S + q * 16 + (n & 15);
          if (o < O && seed < a.M) {
            float v = red[0][n][o];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) v += red[w][n][o];
            a.Y[((int64_t)a2 * a.M + seed) * O + o] = a2 == 0 ? (v + a.bout[o]) * a.out_scale : v * tscale;
          }
        }
      }
    }
  }
}

static_assert(kThreads == kNFreq * kMaxIn, "one thread per staged frequency entry");

__global__ void pack_fused_kernel(const float *W, int rows, int K, __bf16 *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * K) return;
  const int j = (int)(i & 7), lane = (int)((i >> 3) & 63);
  const int64_t frag = i >> 9;
  const int tiles = rows / 16;
  const int ks = (int)(frag / tiles), ft = (int)(frag % tiles);
  const int m = ft * 16 + (lane & 15), k = ks * 32 + 8 * (lane >> 4) + j;
  out[i] = (__bf16)W[(int64_t)m * K + k];
}

template <int W, int RT, int NT, int O>
void launch_fused(const FusedArgs &a, hipStream_t st) {
  // two work-groups per CU (LDS and registers admit two); each loops over tiles
  constexpr int64_t S = Shape<W, RT>::kRows / (1 + NT);
  const int64_t tiles = ((int64_t)a.M + S - 1) / S;
  const unsigned grid = (unsigned)std::min<int64_t>(tiles, 2 * (int64_t)device_cu_count());
  fused_mlp_kernel<W, RT, NT, O><<<grid, kThreads, 0, st>>>(a);
}

}  // namespace
}  // namespace pyftle

using namespace pyftle;

extern "C" {

int pyftle_mlp_pack_fused(const float *W, int32_t rows, int32_t K, void *out, void *stream) {
  if (!W || !out) return fail("mlp_pack_fused: NULL pointer");
  if (rows < 16 || rows % 16 || K < 32 || K % 32)
    return fail("mlp_pack_fused: rows must be a positive multiple of 16 and K of 32");
  const int64_t n = (int64_t)rows * K;
  pack_fused_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(W, rows, K, (__bf16 *)out);
  return check_launch("mlp_pack_fused");
}

int pyftle_mlp_fused_forward(const float *x, int32_t d_in, int32_t M, const float *scale, const float *Bf,
                            int32_t n_freq, const void *W_packed, const float *bias, int32_t depth,
                            int32_t width, const float *Wout, const float *bout, int32_t O, float out_scale,
                            int32_t n_tangents, float *Y, void *stream) {
  if (!x || !scale || !Bf || !W_packed || !bias || !Wout || !bout || !Y)
    return fail("mlp_fused_forward: NULL pointer");
  if (n_freq != kNFreq || (width != 256 && width != 512))
    return fail("mlp_fused_forward: needs 64 Fourier frequencies and hidden width 256 or 512");
  if (depth < 2 || depth > kMaxDepth) return fail("mlp_fused_forward: depth must be in [2, 8]");
  if (d_in < 1 || d_in > kMaxIn || M < 0 || n_tangents < 0 || n_tangents > d_in)
    return fail("mlp_fused_forward: bad input sizes");
  if ((uintptr_t)W_packed % 16 || (uintptr_t)bias % 16 || (uintptr_t)Wout % 16)
    return fail("mlp_fused_forward: packed weights, bias and Wout must be 16-byte aligned");
  if (M == 0) return PYFTLE_OK;
  const int w_bytes = 2 * (width * kFeat + (depth - 1) * width * width);
  FusedArgs a{x, scale, Bf, (const __bf16 *)W_packed, w_bytes, bias, Wout, bout, Y, out_scale, d_in, M, depth};
  hipStream_t st = as_stream(stream);
  // row tiles per work-group: 64 rows at width 512 (the accumulators fill the registers);
  // 128 at width 256, or 96 (32 seeds x 3 row blocks) with 2 tangents
  if (width == 512 && O == 2 && n_tangents == 0) launch_fused<512, 4, 0, 2>(a, st);
  else if (width == 512 && O == 3 && n_tangents == 0) launch_fused<512, 4, 0, 3>(a, st);
  else if (width == 512 && O == 3 && n_tangents == 3) launch_fused<512, 4, 3, 3>(a, st);
  else if (width == 256 && O == 2 && n_tangents == 0) launch_fused<256, 8, 0, 2>(a, st);
  else if (width == 256 && O == 2 && n_tangents == 2) launch_fused<256, 6, 2, 2>(a, st);
  else if (width == 256 && O == 3 && n_tangents == 0) launch_fused<256, 8, 0, 3>(a, st);
  else
    return fail("mlp_fused_forward: supported (width, outputs, tangents): (512, 2, 0), (512, 3, 0), "
                "(512, 3, 3), (256, 2, 0), (256, 2, 2), (256, 3, 0)");
  return check_launch("mlp_fused_forward");
}

}  // extern "C"
