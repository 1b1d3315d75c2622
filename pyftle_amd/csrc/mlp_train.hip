// Fused surrogate training step: one minibatch of the Fourier-feature MLP in three
// stages instead of ~45 layer-by-layer launches.
//
// No reference counterpart: the learned flow-map approximation of the RK4 path
// (integrate.py:265-342) is this engine's addition.
//
// 1. train_kernel (one work-group per tile of 16*RT rows, activations in LDS as in the
//    fused inference kernel; 64 rows, on eight waves at width 256 and four at width 512,
//    each wave owning a share of the layer's 16-feature tiles for all the rows, so a weight
//    fragment feeds RT MFMAs): Fourier features, every hidden layer forward, the output
//    layer and the MSE gradient, then the backward pass through the hidden layers with the
//    transposed weights (dH_{l-1} = dZ_l W_l, the same MFMA loop with W_l^T packed), each
//    dZ_l = dH_l * SiLU'(z_l).  It writes what the weight gradients need, transposed so that
//    the batch rows are contiguous (the K dimension of the next stage): the features and
//    hidden activations (HT) and every dZ (DZT); per-tile partial sums of the bias, output
//    weight and output bias gradients and of the loss; and SiLU'(z) of the forward pass,
//    in its own register order, for its own backward pass.
// 2. Weight gradients dW_l = DZT_l HT_{l-1}^T: the BF16 MFMA GEMM of mlp.hip with both
//    operands K-contiguous, split over K (FP32 partials).
// 3. train_reduce_kernel: every parameter's gradient summed over splits or tiles in a fixed
//    order (deterministic), the loss accumulated, the batch offset advanced; in
//    pyftle_mlp_train_step it also applies Adam to each summed gradient and rewrites the
//    forward and transposed fragment-packed bf16 weights the next step reads
//    (pyftle_mlp_train_grads leaves that to train_adam_kernel, pyftle_mlp_train_adam).
//
// Transposed copies of the hidden activations and of every dZ leave the registers they are
// computed in: the four lanes of a quad hold a 4 x 4 block (rows x features) of a tile, which
// two DPP exchanges inside the quad turn into 4 rows of one feature per lane; one more
// exchange between neighbouring quads pairs two row tiles into 16-byte stores of 8 rows.
// The features' copy leaves LDS as 16-byte stores: a lane takes a feature pair and 8 rows,
// reading the 8 rows in an order rotated by its row octet so that the 64 lanes of a read
// hit distinct banks, and rotates the words back in registers.

#include "mlp_fused.hpp"
#include "mlp_gemm.hpp"

namespace pyftle {
namespace {

using namespace mlpf;

struct TrainArgs {
  const float *x;          // (rows, d_in)
  const float *t;          // (rows, O) targets of the network output (already / out_scale)
  const int64_t *base_row; // first row of this step
  int M;                   // rows per step
  int d_in, depth;
  const float *scale, *Bf;
  const __bf16 *Wp;        // forward layers, then transposed layers 1..depth-1 (packed)
  int wp_bytes;
  const float *params;     // flat FP32 parameters
  float loss_scale;        // 1 / (M * O)
  __bf16 *FT, *HT, *DZT, *DACT;
  float *bias_part, *wout_part, *bout_part, *loss_part;
  int *step;               // step index, += 1 here (nothing reads it before the reduction)
};

// flat parameter layout: per layer W_l (W x fan_l) then b_l (W); then Wout (O x W), bout (O)
template <int W>
__host__ __device__ __forceinline__ int64_t w_off(int l) {
  return l == 0 ? 0 : (int64_t)W * kFeat + W + (int64_t)(l - 1) * ((int64_t)W * W + W);
}
template <int W>
__host__ __device__ __forceinline__ int64_t b_off(int l) {
  return w_off<W>(l) + (int64_t)W * (l == 0 ? kFeat : W);
}
// byte offset of transposed layer l (1 <= l < depth) in the packed buffer
template <int W>
__host__ __device__ __forceinline__ int tlayer_bytes(int l, int depth) {
  return 2 * (W * kFeat + (depth - 1) * W * W) + 2 * (l - 1) * W * W;
}

// LDS rows [0, 16 RT) x features [0, nfeat) -> dst[f * ld + col0 + row] (bf16), NTH threads
template <int W, int RT, int NTH>
__device__ __forceinline__ void transpose_out(const __bf16 *act, int nfeat, __bf16 *dst, int64_t ld,
                                              int64_t col0, int tid) {
  constexpr int Q = 2 * RT;  // row octets
  const int tasks = (nfeat / 2) * Q;
  for (int task = tid; task < tasks; task += NTH) {
    const int q = task % Q, p = task / Q;
    const int rot = (q >> 1) & 7;
    const int f = 2 * p;
    uint32_t v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int r = 8 * q + ((j + rot) & 7);
      v[j] = *reinterpret_cast<const uint32_t *>(act + act_off<W>(r, f >> 3) + (f & 7));
    }
    // v[j] holds row 8 q + ((j + rot) & 7); rotate so that v[k] holds row 8 q + k
#pragma unroll
    for (int s = 1; s < 8; s <<= 1) {
      uint32_t w[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) w[k] = (rot & s) ? v[(k - s) & 7] : v[k];
#pragma unroll
      for (int k = 0; k < 8; ++k) v[k] = w[k];
    }
    uint4 lo, hi;  // feature f (low halves), feature f + 1 (high halves), rows 8 q .. 8 q + 7
    lo.x = (v[0] & 0xffffu) | (v[1] << 16);
    lo.y = (v[2] & 0xffffu) | (v[3] << 16);
    lo.z = (v[4] & 0xffffu) | (v[5] << 16);
    lo.w = (v[6] & 0xffffu) | (v[7] << 16);
    hi.x = (v[0] >> 16) | (v[1] & 0xffff0000u);
    hi.y = (v[2] >> 16) | (v[3] & 0xffff0000u);
    hi.z = (v[4] >> 16) | (v[5] & 0xffff0000u);
    hi.w = (v[6] >> 16) | (v[7] & 0xffff0000u);
    *reinterpret_cast<uint4 *>(dst + (int64_t)f * ld + col0 + 8 * q) = lo;
    *reinterpret_cast<uint4 *>(dst + (int64_t)(f + 1) * ld + col0 + 8 * q) = hi;
  }
}

// DPP move of x (lanes outside the bank mask keep `old`).  Every lane of the wave must be active.
template <int CTRL, int BANKS = 0xF>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t old, uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)x, CTRL, 0xF, BANKS, false);
}
constexpr int kDppXor1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kDppXor2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kDppShl4 = 0x104;  // row_shl:4, lane l reads l + 4 of its row of 16
constexpr int kDppShr4 = 0x114;  // row_shr:4, lane l reads l - 4
constexpr int kDppRor8 = 0x128;  // row_ror:8, lane l reads (l + 8) mod 16 of its row: l ^ 8

// The word of lane l ^ 4: banks 0 and 2 of each row of 16 lanes read 4 lanes up, banks 1 and
// 3 four lanes down.
__device__ __forceinline__ uint32_t swap_xor4(uint32_t x) {
  return dpp_u32<kDppShr4, 0xA>(dpp_u32<kDppShl4, 0x5>(0u, x), x);
}

// v summed over the 16 lanes of its row (lane / 16), in the order of an xor butterfly over
// lane bits 0, 1, 2, 3 (every lane gets the same bits); DPP moves instead of LDS permutes.
__device__ __forceinline__ float row16_sum(float v) {
  v += __builtin_bit_cast(float, dpp_u32<kDppXor1>(0u, __builtin_bit_cast(uint32_t, v)));
  v += __builtin_bit_cast(float, dpp_u32<kDppXor2>(0u, __builtin_bit_cast(uint32_t, v)));
  v += __builtin_bit_cast(float, swap_xor4(__builtin_bit_cast(uint32_t, v)));
  v += __builtin_bit_cast(float, dpp_u32<kDppRor8>(0u, __builtin_bit_cast(uint32_t, v)));
  return v;
}

// The 4 x 4 bf16 block of a quad of lanes transposed inside the quad.  On entry lane q
// (q = lane % 4) holds row q of the block as two words, features (0, 1) in `lo` and (2, 3) in
// `hi` (the first of each pair in the low half); on return it holds feature q, rows (0, 1) in
// .x and rows (2, 3) in .y.  First the lanes q and q ^ 2 swap word pairs (DPP quad_perm
// [2,3,0,1]): lane q then holds features (q & 2, +1) of rows q & 1 (u) and (q & 1) + 2 (v).
// Then the lanes q and q ^ 1 swap halves (quad_perm [1,0,3,2]); v_perm_b32 picks the low
// halves (even lanes: own, partner's) or the high halves (odd lanes: partner's, own) of the
// combined bytes {partner word: 4..7, own word: 0..3}.  Every lane of the wave must be active.
__device__ __forceinline__ uint2 quad_transpose(uint32_t lo, uint32_t hi, int q) {
  const bool upper = (q & 2) != 0;
  const uint32_t x = upper ? lo : hi;
  const uint32_t y = dpp_u32<kDppXor2>(0u, x);
  const uint32_t u = upper ? y : lo;
  const uint32_t v = upper ? hi : y;
  const uint32_t up = dpp_u32<kDppXor1>(0u, u), vp = dpp_u32<kDppXor1>(0u, v);
  const uint32_t sel = (q & 1) ? 0x03020706u : 0x05040100u;
  return uint2{__builtin_amdgcn_perm(up, u, sel), __builtin_amdgcn_perm(vp, v, sel)};
}

// Row tiles t and t + 1 of lane (r, g) -- rows 16 t + r and 16 (t + 1) + r, features
// f0 .. f0 + 3 (f0 = 16 ft + 4 g), as bf16x4 h0, h1 -- stored transposed, dst[f * ld + col0 +
// row], as one 16-byte store per lane: after the quad transposes lane (a = r / 4, q = r % 4)
// holds rows 4 a .. 4 a + 3 of feature f0 + q in both tiles, and lanes a and a ^ 1 swap one
// tile's words, so that an even a stores rows 4 a .. 4 a + 7 of tile t and an odd a rows
// 4 (a - 1) .. 4 (a - 1) + 7 of tile t + 1.  A wave's store covers 32 rows of 16 features.
__device__ __forceinline__ void store_transposed(bf16x4_t h0, bf16x4_t h1, __bf16 *dst, int64_t ld, int64_t col0,
                                                 int f0, int t, int r) {
  const uint2 w0 = __builtin_bit_cast(uint2, h0), w1 = __builtin_bit_cast(uint2, h1);
  const uint2 o0 = quad_transpose(w0.x, w0.y, r & 3);
  const uint2 o1 = quad_transpose(w1.x, w1.y, r & 3);
  const bool odd = (r & 4) != 0;
  const uint32_t y0 = swap_xor4(odd ? o0.x : o1.x), y1 = swap_xor4(odd ? o0.y : o1.y);
  const uint4 out = odd ? uint4{y0, y1, o1.x, o1.y} : uint4{o0.x, o0.y, y0, y1};
  const int row = odd ? 16 * (t + 1) + 4 * ((r >> 2) - 1) : 16 * t + 4 * (r >> 2);
  *reinterpret_cast<uint4 *>(dst + (int64_t)(f0 + (r & 3)) * ld + col0 + row) = out;
}

// Per-feature sum over this lane's rows (its RT row tiles), then over the 16 lanes that
// share the lane's features (lane % 16 = row within a tile): lanes with lane % 16 == 0 write.
template <int W, int RT, int WV>
__device__ __forceinline__ void feature_sums(const f32x4_t (&acc)[Shape<W, RT, WV>::kTilesPerWave][RT], float *dst,
                                             int ft0, int lane_in) {
  const int lane = opaque_lane(lane_in);
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < Shape<W, RT, WV>::kTilesPerWave; ++i) {
    float s[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float v = 0.0f;
#pragma unroll
      for (int t = 0; t < RT; ++t) v += acc[i][t][e];
      s[e] = row16_sum(v);
    }
    if (r == 0) *reinterpret_cast<f32x4_t *>(dst + (ft0 + i) * 16 + 4 * g) = f32x4_t{s[0], s[1], s[2], s[3]};
  }
}

// acc (FP32) as bf16 activation rows of LDS (the layout the MFMA loop reads) and, transposed,
// as dst[f * ld + col0 + row]
template <int W, int RT, int WV>
__device__ __forceinline__ void store_rows(const f32x4_t (&acc)[Shape<W, RT, WV>::kTilesPerWave][RT], __bf16 *act,
                                           __bf16 *dst, int64_t ld, int64_t col0, int ft0, int lane_in) {
  const int lane = opaque_lane(lane_in);
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < Shape<W, RT, WV>::kTilesPerWave; ++i) {
    const int f0 = (ft0 + i) * 16 + 4 * g;
    const int off = act_off<W>(r, f0 >> 3) + (f0 & 7);
    bf16x4_t h[RT];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
#pragma unroll
      for (int e = 0; e < 4; ++e) h[t][e] = (__bf16)acc[i][t][e];
      *reinterpret_cast<bf16x4_t *>(act + off + t * 16 * W) = h[t];
    }
#pragma unroll
    for (int t = 0; t < RT; t += 2) store_transposed(h[t], h[t + 1], dst, ld, col0, f0, t, r);
  }
}

// WV waves of 64 lanes per work-group.  At width 256 eight waves share 64 rows, two per SIMD
// (each weight fragment feeds 4 MFMAs, against 2 with four waves on 32 rows and two
// work-groups per CU); at width 512 four waves on 64 rows, one per SIMD: the accumulators
// alone take half the registers of two.
template <int W, int RT, int O, int WV>
__global__ void __launch_bounds__(64 * WV, W == 512 ? 1 : 2) train_kernel(TrainArgs a) {
  using Sh = Shape<W, RT, WV>;
  constexpr int TPW = Sh::kTilesPerWave, kRows = Sh::kRows, NTH = 64 * WV;
  static_assert(RT % 2 == 0, "row tiles leave transposed in pairs");
  static_assert(NTH >= kNFreq * kMaxIn && NTH >= kRows * 3, "one thread per staged frequency and output");
  __shared__ __attribute__((aligned(16))) __bf16 act[kRows * W];
  __shared__ __attribute__((aligned(16))) float freq_s[kNFreq * kMaxIn];
  __shared__ float red[WV][kRows][4];
  constexpr int kDPerWave = TPW * RT * 64 * 4;  // SiLU'(z) values (bf16) per wave and layer
  __shared__ float dy_s[kRows][4];
  __shared__ float sq_s[kRows * 4];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ft0 = wave * TPW;
  const int tile = blockIdx.x, ntiles = gridDim.x;
  const int64_t grow0 = *a.base_row + (int64_t)tile * kRows;  // data row of the tile's row 0
  const int64_t col0 = (int64_t)tile * kRows;                 // its column in the step's buffers
  const int64_t M = a.M;
  if (a.step && tile == 0 && tid == 0) *a.step += 1;
  if (tid < kNFreq * kMaxIn) {
    const int q = tid >> 2, c = tid & 3;
    float fv = 0.0f, sc = 0.0f;
    if (c < a.d_in) {
      fv = a.Bf[q * a.d_in + c];
      sc = a.scale[c];
    }
    freq_s[tid] = fv * sc;
  }
  const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc((void *)a.Wp, 0, a.wp_bytes, 0x00020000);
  // weight fragments PD k-steps ahead: at width 256 a k-step is only TPW * RT = 8 MFMAs per
  // wave, so one k-step of cover leaves the L2 latency exposed; width 512 has no registers
  // left for more
  constexpr int PD = W == 256 ? 4 : 1;
  bf16x8_t ring[PD][TPW];
#pragma unroll
  for (int p = 0; p < PD; ++p)
#pragma unroll
    for (int i = 0; i < TPW; ++i) ring[p][i] = load_frag(wr, lane * 16, frag_bytes<W>(p, ft0 + i));
  __syncthreads();

  // Fourier features [sin | cos] of the tile's rows
  for (int task = tid; task < kRows * 8; task += NTH) {
    const int j = task >> 3, fb = task & 7;
    float xin[kMaxIn];
#pragma unroll
    for (int c = 0; c < kMaxIn; ++c) xin[c] = c < a.d_in ? a.x[(grow0 + j) * a.d_in + c] : 0.0f;
    bf16x8_t sn, cs;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const f32x4_t fr = *reinterpret_cast<const f32x4_t *>(freq_s + (fb * 8 + jj) * kMaxIn);
      float ph = 0.0f;
#pragma unroll
      for (int c = 0; c < kMaxIn; ++c) ph += fr[c] * xin[c];
      ph -= rintf(ph);
      sn[jj] = (__bf16)__builtin_amdgcn_sinf(ph);
      cs[jj] = (__bf16)__builtin_amdgcn_cosf(ph);
    }
    *reinterpret_cast<bf16x8_t *>(act + act_off<W>(j, fb)) = sn;
    *reinterpret_cast<bf16x8_t *>(act + act_off<W>(j, kNFreq / 8 + fb)) = cs;
  }
  __syncthreads();
  transpose_out<W, RT, NTH>(act, kFeat, a.FT, M, col0, tid);

  f32x4_t acc[TPW][RT];
  // Forward.  Layer 0 (K = 128) and the last layer are peeled off the loop so that no
  // accumulator array is carried around a loop holding two instantiations of the MFMA loop.
  layer_mma_ring<W, RT, kFeat, PD, WV>(wr, layer_bytes<W>(0), layer_bytes<W>(1), act, ring, acc, ft0, lane);
  for (int L = 0;; ++L) {
    __syncthreads();  // every wave has read this layer's inputs
    {
      const float *bl = a.params + b_off<W>(L);
      const int ln = opaque_lane(lane);
      const int r = ln & 15, g = ln >> 4;
      __bf16 *dact = a.DACT + (((int64_t)L * ntiles + tile) * WV + wave) * kDPerWave;
      __bf16 *ht = a.HT + (int64_t)L * W * M;
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
        const int f0 = (ft0 + i) * 16 + 4 * g;
        const f32x4_t bv = *reinterpret_cast<const f32x4_t *>(bl + f0);
        const int off = act_off<W>(r, f0 >> 3) + (f0 & 7);
        bf16x4_t h[RT];
#pragma unroll
        for (int t = 0; t < RT; ++t) {
          bf16x4_t d;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float z = acc[i][t][e] + bv[e];
            const float s = sigmoid_f(z);
            h[t][e] = (__bf16)(z * s);
            d[e] = (__bf16)(s * (1.0f + z * (1.0f - s)));
          }
          *reinterpret_cast<bf16x4_t *>(act + off + t * 16 * W) = h[t];
          *reinterpret_cast<bf16x4_t *>(dact + ((i * RT + t) * 64 + ln) * 4) = d;
        }
#pragma unroll
        for (int t = 0; t < RT; t += 2) store_transposed(h[t], h[t + 1], ht, M, col0, f0, t, r);
        __builtin_amdgcn_sched_barrier(0);  // one feature tile at a time: no loads hoisted
      }
    }
    __syncthreads();
    if (L + 2 == a.depth) break;
    layer_mma_ring<W, RT, W, PD, WV>(wr, layer_bytes<W>(L + 1), layer_bytes<W>(L + 2), act, ring, acc, ft0, lane);
  }
  {
    const int L = a.depth - 1;
    layer_mma_ring<W, RT, W, PD, WV>(wr, layer_bytes<W>(L), tlayer_bytes<W>(L, a.depth), act, ring, acc, ft0, lane);
    const float *bl = a.params + b_off<W>(L);
    __syncthreads();  // every wave has read the last layer's inputs
    // output layer (FP32): y = SiLU(z + b) Wout^T + bout, then the MSE gradient
    const float *wout = a.params + w_off<W>(a.depth);
    const float *bout = wout + O * W;
    {
      const int g = opaque_lane(lane) >> 4;
      float yp[RT][O];
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int o = 0; o < O; ++o) yp[t][o] = 0.0f;
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
        const int f0 = (ft0 + i) * 16 + 4 * g;
        const f32x4_t bv = *reinterpret_cast<const f32x4_t *>(bl + f0);
        f32x4_t wv[O];
#pragma unroll
        for (int o = 0; o < O; ++o) wv[o] = *reinterpret_cast<const f32x4_t *>(wout + o * W + f0);
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const float z = acc[i][t][e] + bv[e];
            const float hz = z * sigmoid_f(z);
#pragma unroll
            for (int o = 0; o < O; ++o) yp[t][o] = fmaf(hz, wv[o][e], yp[t][o]);
          }
        __builtin_amdgcn_sched_barrier(0);
      }
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
    }
    __syncthreads();
    if (tid < kRows * O) {
      const int n = tid / O, o = tid % O;
      float y = bout[o];
#pragma unroll
      for (int w = 0; w < WV; ++w) y += red[w][n][o];
      const float diff = y - a.t[(grow0 + n) * O + o];
      dy_s[n][o] = 2.0f * diff * a.loss_scale;
      sq_s[tid] = diff * diff * a.loss_scale;
    }
    __syncthreads();
    if (tid == 0) {
      float l = 0.0f;
      for (int k = 0; k < kRows * O; ++k) l += sq_s[k];
      a.loss_part[tile] = l;
    } else if (tid <= O) {
      float sb = 0.0f;
      for (int n = 0; n < kRows; ++n) sb += dy_s[n][tid - 1];
      a.bout_part[tile * 4 + tid - 1] = sb;
    }
    // dZ of the last hidden layer and the output weights' gradient partials
    float *wpart = a.wout_part + (int64_t)tile * O * W;
    const int ln = opaque_lane(lane);
    const int r = ln & 15, g = ln >> 4;
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      const int f0 = (ft0 + i) * 16 + 4 * g;
      const f32x4_t bv = *reinterpret_cast<const f32x4_t *>(bl + f0);
      f32x4_t wv[O];
#pragma unroll
      for (int o = 0; o < O; ++o) wv[o] = *reinterpret_cast<const f32x4_t *>(wout + o * W + f0);
      float ws[O][4];
#pragma unroll
      for (int o = 0; o < O; ++o)
#pragma unroll
        for (int e = 0; e < 4; ++e) ws[o][e] = 0.0f;
#pragma unroll
      for (int t = 0; t < RT; ++t) {
        float dyv[O];
#pragma unroll
        for (int o = 0; o < O; ++o) dyv[o] = dy_s[t * 16 + r][o];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float z = acc[i][t][e] + bv[e];
          const float s = sigmoid_f(z);
          const float hz = z * s;
          float dh = 0.0f;
#pragma unroll
          for (int o = 0; o < O; ++o) {
            dh = fmaf(dyv[o], wv[o][e], dh);
            ws[o][e] = fmaf(dyv[o], hz, ws[o][e]);
          }
          acc[i][t][e] = dh * (s * (1.0f + z * (1.0f - s)));
        }
      }
#pragma unroll
      for (int o = 0; o < O; ++o)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          ws[o][e] = row16_sum(ws[o][e]);
        }
      if (r == 0) {
#pragma unroll
        for (int o = 0; o < O; ++o)
          *reinterpret_cast<f32x4_t *>(wpart + o * W + f0) = f32x4_t{ws[o][0], ws[o][1], ws[o][2], ws[o][3]};
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  // backward through the hidden layers: acc holds dZ of layer l
  for (int l = a.depth - 1;; --l) {
    feature_sums<W, RT, WV>(acc, a.bias_part + ((int64_t)tile * a.depth + l) * W, ft0, lane);
    // every wave finished reading act before the last barrier
    store_rows<W, RT, WV>(acc, act, a.DZT + (int64_t)l * W * M, M, col0, ft0, lane);
    if (l == 0) break;
    __syncthreads();
    const int lb = tlayer_bytes<W>(l, a.depth);
    const int nlb = l > 1 ? tlayer_bytes<W>(l - 1, a.depth) : layer_bytes<W>(0);
    layer_mma_ring<W, RT, W, PD, WV>(wr, lb, nlb, act, ring, acc, ft0, lane);  // dH of layer l - 1
    __syncthreads();
    const __bf16 *dact = a.DACT + (((int64_t)(l - 1) * ntiles + tile) * WV + wave) * kDPerWave;
    const int ln = opaque_lane(lane);
#pragma unroll
    for (int i = 0; i < TPW; ++i)
#pragma unroll
      for (int t = 0; t < RT; ++t) {
        const bf16x4_t d = *reinterpret_cast<const bf16x4_t *>(dact + ((i * RT + t) * 64 + ln) * 4);
#pragma unroll
        for (int e = 0; e < 4; ++e) acc[i][t][e] *= (float)d[e];
      }
  }
}

// Adam on the FP32 master weights with the cosine learning-rate schedule evaluated from
// *step (the step's index; bias corrections for Adam step t0 + *step + 1).  Shared by
// train_adam_kernel and the reduction's fused update, so both give the same bits.
struct AdamParams {
  float *w, *m, *v;
  __bf16 *packed;  // forward and transposed bf16 packing of the hidden weights
  float b1, b2, eps, wd, lr_max, lr_min;
  int total, t0;
  const int *step;
};

struct AdamSchedule {
  float lr, bc1, bc2;
};

__device__ __forceinline__ AdamSchedule adam_schedule(const AdamParams &p) {
  const int k = *p.step;
  const double c = cos(3.141592653589793 * (double)k / (double)(p.total > 0 ? p.total : 1));
  AdamSchedule s;
  s.lr = (float)((double)p.lr_min + 0.5 * ((double)p.lr_max - (double)p.lr_min) * (1.0 + c));
  s.bc1 = 1.0f - powf(p.b1, (float)(p.t0 + k + 1));
  s.bc2 = 1.0f - powf(p.b2, (float)(p.t0 + k + 1));
  return s;
}

// Parameter i's weight and moments, loaded ahead of the gradient they are updated with.
struct AdamState {
  float w, m, v;
};
__device__ __forceinline__ AdamState adam_load(const AdamParams &p, int64_t i) { return {p.w[i], p.m[i], p.v[i]}; }

// updates parameter i (state st, as loaded) with gradient g; returns the new weight
__device__ __forceinline__ float adam_apply(const AdamParams &p, const AdamSchedule &s, int64_t i, float g,
                                            AdamState st) {
  const float gi = g + p.wd * st.w;
  const float mi = p.b1 * st.m + (1.0f - p.b1) * gi;
  const float vi = p.b2 * st.v + (1.0f - p.b2) * gi * gi;
  p.m[i] = mi;
  p.v[i] = vi;
  const float wi = st.w - s.lr * (mi / s.bc1) / (sqrtf(vi / s.bc2) + p.eps);
  p.w[i] = wi;
  return wi;
}

// index of element (m, k) of a rows x K matrix in the fragment-packed order of
// pyftle_mlp_pack_fused
__device__ __forceinline__ int64_t packed_index(int m, int k, int rows) {
  const int ft = m >> 4, ks = k >> 5, ln = (m & 15) + 16 * ((k & 31) >> 3);
  return (((int64_t)ks * (rows >> 4) + ft) * 64 + ln) * 8 + (k & 7);
}

// the bf16 packings of hidden weight (o, i) of layer l (fan inputs), whose forward packing
// starts at element pk: forward order, and for l > 0 the transposed layer the backward pass reads
__device__ __forceinline__ void pack_weight(__bf16 *packed, int W, int depth, int l, int fan, int64_t pk, int o,
                                            int i, float wi) {
  packed[pk + packed_index(o, i, W)] = (__bf16)wi;
  if (l > 0) {  // W_l^T: fan x W, element (i, o)
    const int64_t tb = (int64_t)W * kFeat + (int64_t)(depth - 1) * W * W + (int64_t)(l - 1) * W * W;
    packed[tb + packed_index(i, o, fan)] = (__bf16)wi;
  }
}

struct ReduceArgs {
  const float *part;       // weight-gradient partials, layer after layer: (splits, W, fan_l)
  int splits;
  const float *bias_part, *wout_part, *bout_part, *loss_part;
  int tiles, depth, W, O;
  int64_t n_weights;       // hidden-layer weights (the split-K part)
  int n_tiled;             // biases, output weights and output biases (the per-tile part)
  int weight_blocks;       // work-groups of the first part
  float *grads;
  float *loss;             // += the step's loss
  int64_t *base_row;       // += M
  int M;
  AdamParams adam;         // ADAM: the update applied to each summed gradient
};

// src[k * stride] summed over k = k0, k0 + step, ... below n, added in that order.  The loads
// go out 16 at a time: the partials were written by the previous launch and come from L2 or
// memory, so a term-by-term chain would wait one round trip per term.
__device__ __forceinline__ float strided_sum(const float *src, int64_t stride, int k0, int step, int n) {
  float s = 0.0f;
  for (int k = k0; k < n; k += 16 * step) {
    float v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int kk = k + j * step;
      v[j] = kk < n ? src[(int64_t)kk * stride] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) s += v[j];
  }
  return s;
}

// Gradients summed in a fixed order: the hidden weights over the GEMM splits (a thread per
// weight), the rest over the row tiles (16 parameters per work-group, 16 interleaved partial
// sums per parameter combined in order), the loss by one more work-group.  With ADAM each
// summed gradient is also applied to its parameter at once (the step's Adam update, as
// train_adam_kernel would apply it, its schedule evaluated once per work-group) and the hidden
// weights' bf16 packings are rewritten.
template <bool ADAM>
__global__ void __launch_bounds__(256) train_reduce_kernel(ReduceArgs a) {
  const int W = a.W;
  __shared__ AdamSchedule sched;
  if constexpr (ADAM) {
    if (threadIdx.x == 0) sched = adam_schedule(a.adam);
    __syncthreads();
  }
  if ((int)blockIdx.x == (int)gridDim.x - 1) {
    // the last work-group: the loss over the row tiles (thread j sums tiles j, j + 256, ...
    // in order, then thread 0 the 256 partial sums in order) and the batch offset
    __shared__ float lp[256];
    float l = 0.0f;
    for (int k = threadIdx.x; k < a.tiles; k += blockDim.x) l += a.loss_part[k];
    lp[threadIdx.x] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
      float t = 0.0f;
      for (int j = 0; j < (int)blockDim.x; ++j) t += lp[j];
      *a.loss += t;
      *a.base_row += a.M;
    }
    return;
  }
  if ((int)blockIdx.x < a.weight_blocks) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.n_weights) return;
    int64_t off = 0, poff = 0, woff = 0;  // flat offset, partial offset, weight index of layer l
    for (int l = 0; l < a.depth; ++l) {
      const int fan = l == 0 ? kFeat : W;
      const int64_t nw = (int64_t)W * fan;
      if (q < woff + nw) {
        const int64_t loc = q - woff;
        AdamState st{};  // loaded with the partials: one memory round trip, not two
        if constexpr (ADAM) st = adam_load(a.adam, off + loc);
        const float gsum = strided_sum(a.part + poff + loc, nw, 0, 1, a.splits);
        a.grads[off + loc] = gsum;
        if constexpr (ADAM) {
          const float wi = adam_apply(a.adam, sched, off + loc, gsum, st);
          pack_weight(a.adam.packed, W, a.depth, l, fan, woff, (int)(loc / fan), (int)(loc % fan), wi);
        }
        return;
      }
      off += nw + W;
      poff += (int64_t)a.splits * nw;
      woff += nw;
    }
    return;
  }
  __shared__ float ps[16][16];
  const int qi = threadIdx.x & 15, ph = threadIdx.x >> 4;
  const int q = ((int)blockIdx.x - a.weight_blocks) * 16 + qi;
  const float *src = nullptr;
  int64_t stride = 0, dst = 0;
  if (q < a.n_tiled) {
    const int nb = a.depth * W;
    int64_t out_off = 0;
    for (int l = 0; l < a.depth; ++l) out_off += (int64_t)W * (l == 0 ? kFeat : W) + W;
    if (q < nb) {
      const int l = q / W, o = q % W;
      src = a.bias_part + (int64_t)l * W + o;
      stride = (int64_t)a.depth * W;
      dst = (l == 0 ? (int64_t)W * kFeat : (int64_t)W * kFeat + W + (int64_t)(l - 1) * ((int64_t)W * W + W) +
                                               (int64_t)W * W) + o;
    } else if (q < nb + a.O * W) {
      const int loc = q - nb;
      src = a.wout_part + loc;
      stride = (int64_t)a.O * W;
      dst = out_off + loc;
    } else {
      const int o = q - nb - a.O * W;
      src = a.bout_part + o;
      stride = 4;
      dst = out_off + (int64_t)a.O * W + o;
    }
  }
  AdamState st{};
  if constexpr (ADAM) {
    if (ph == 0 && src) st = adam_load(a.adam, dst);
  }
  ps[ph][qi] = src ? strided_sum(src, stride, ph, 16, a.tiles) : 0.0f;
  __syncthreads();
  if (ph == 0 && src) {
    float t = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) t += ps[j][qi];
    a.grads[dst] = t;
    if constexpr (ADAM) adam_apply(a.adam, sched, dst, t, st);
  }
}

__global__ void __launch_bounds__(256) pack_train_kernel(const float *params, int W, int depth, __bf16 *out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t off = 0, pk = 0;  // parameter offset and packed forward offset (elements) of layer l
  for (int l = 0; l < depth; ++l) {
    const int fan = l == 0 ? kFeat : W;
    const int64_t nw = (int64_t)W * fan;
    if (q < pk + nw) {
      const int64_t loc = q - pk;
      pack_weight(out, W, depth, l, fan, pk, (int)(loc / fan), (int)(loc % fan), params[off + loc]);
      return;
    }
    off += nw + W;
    pk += nw;
  }
}

// Adam over every parameter (gradients from train_reduce_kernel<false>), writing the hidden
// weights' forward and transposed bf16 packings.
__global__ void __launch_bounds__(256) train_adam_kernel(AdamParams ad, const float *g, int64_t n, int W,
                                                         int depth) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float wi = adam_apply(ad, adam_schedule(ad), p, g[p], adam_load(ad, p));
  int64_t off = 0, pk = 0;
  for (int l = 0; l < depth; ++l) {
    const int fan = l == 0 ? kFeat : W;
    const int64_t nw = (int64_t)W * fan;
    if (p < off + nw) {
      const int64_t loc = p - off;
      pack_weight(ad.packed, W, depth, l, fan, pk, (int)(loc / fan), (int)(loc % fan), wi);
      return;
    }
    if (p < off + nw + W) return;  // a bias
    off += nw + W;
    pk += nw;
  }
}

template <int W, int RT, int O, int WV>
int launch_train(const TrainArgs &a, hipStream_t st) {
  train_kernel<W, RT, O, WV><<<a.M / (16 * RT), 64 * WV, 0, st>>>(a);
  return check_launch("mlp_train_grads");
}

int64_t align256(int64_t n) { return (n + 255) / 256 * 256; }

struct WorkLayout {
  int64_t ft, ht, dzt, dact, part, bias, wout, bout, loss, total;
};

WorkLayout work_layout(int64_t M, int W, int depth, int O, int splits, int rows_per_tile) {
  WorkLayout w{};
  const int64_t tiles = M / rows_per_tile;
  int64_t nw = 0;
  for (int l = 0; l < depth; ++l) nw += (int64_t)W * (l == 0 ? kFeat : W);
  int64_t o = 0;
  w.ft = o; o += align256(2 * kFeat * M);
  w.ht = o; o += align256(2 * (int64_t)(depth - 1) * W * M);
  w.dzt = o; o += align256(2 * (int64_t)depth * W * M);
  w.dact = o; o += align256(2 * (int64_t)(depth - 1) * W * M);
  w.part = o; o += align256(4 * (int64_t)splits * nw);
  w.bias = o; o += align256(4 * tiles * depth * W);
  w.wout = o; o += align256(4 * tiles * O * W);
  w.bout = o; o += align256(4 * tiles * 4);
  w.loss = o; o += align256(4 * tiles);
  w.total = o;
  return w;
}

// rows per work-group: 64 (eight waves at width 256, four at width 512)
int tile_rows(int) { return 64; }

// forward layers then transposed layers 1..depth-1, bf16
int packed_train_bytes(int W, int depth) { return 2 * (W * kFeat + 2 * (depth - 1) * W * W); }

}  // namespace
}  // namespace pyftle

using namespace pyftle;

extern "C" {

int64_t pyftle_mlp_train_work_bytes(int32_t M, int32_t width, int32_t depth, int32_t O, int32_t splits) {
  if (M <= 0 || (width != 256 && width != 512) || depth < 2 || depth > mlpf::kMaxDepth || O < 1 || O > 3 ||
      splits < 1)
    return -1;
  return work_layout(M, width, depth, O, splits, tile_rows(width)).total;
}

namespace pyftle {
namespace {

// the three launches of a step: train_kernel, the grouped weight-gradient GEMM and the
// reduction (with the Adam update when `adam` is given)
int train_step_launches(const pyftle_mlp_train_args *t, const AdamParams *adam, hipStream_t st);

}  // namespace
}  // namespace pyftle

int pyftle_mlp_train_grads(const pyftle_mlp_train_args *t, void *stream) {
  return train_step_launches(t, nullptr, as_stream(stream));
}

int pyftle_mlp_train_step(const pyftle_mlp_train_args *t, float *m, float *v, float beta1, float beta2, float eps,
                          float weight_decay, float lr_max, float lr_min, int32_t total_steps, int32_t adam_step0,
                          void *stream) {
  if (!t) return fail("mlp_train_step: NULL arguments");
  if (!m || !v || !t->step) return fail("mlp_train_step: NULL Adam moments or step index");
  AdamParams ad{};
  ad.w = const_cast<float *>(t->params);
  ad.m = m;
  ad.v = v;
  ad.packed = static_cast<__bf16 *>(const_cast<void *>(t->W_packed));
  ad.b1 = beta1;
  ad.b2 = beta2;
  ad.eps = eps;
  ad.wd = weight_decay;
  ad.lr_max = lr_max;
  ad.lr_min = lr_min;
  ad.total = total_steps;
  ad.t0 = adam_step0;
  ad.step = t->step;
  return train_step_launches(t, &ad, as_stream(stream));
}

}  // extern "C"

namespace pyftle {
namespace {

int train_step_launches(const pyftle_mlp_train_args *t, const AdamParams *adam, hipStream_t st) {
  if (!t) return fail("mlp_train_grads: NULL arguments");
  if (!t->x || !t->targets || !t->base_row || !t->scale || !t->Bf || !t->W_packed || !t->params || !t->grads ||
      !t->loss || !t->work)
    return fail("mlp_train_grads: NULL pointer");
  if (t->n_freq != mlpf::kNFreq || (t->width != 256 && t->width != 512))
    return fail("mlp_train_grads: needs 64 Fourier frequencies and hidden width 256 or 512");
  if (t->depth < 2 || t->depth > mlpf::kMaxDepth) return fail("mlp_train_grads: depth must be in [2, 8]");
  if (t->d_in < 1 || t->d_in > mlpf::kMaxIn || t->d_out < 1 || t->d_out > 3)
    return fail("mlp_train_grads: 1 <= d_in <= 4 and 1 <= d_out <= 3");
  const int R = tile_rows(t->width);
  const int S = t->splits;
  if (t->M <= 0 || t->M % R || S < 1 || t->M % (128 * S))
    return fail("mlp_train_grads: M must be a multiple of " + std::to_string(R) + " and of 128 * splits");
  if ((uintptr_t)t->params % 16 || (uintptr_t)t->W_packed % 16 || (uintptr_t)t->work % 256)
    return fail("mlp_train_grads: params and W_packed 16-byte, work 256-byte aligned");
  const int W = t->width, depth = t->depth, O = t->d_out;
  const WorkLayout wl = work_layout(t->M, W, depth, O, S, R);
  char *wk = static_cast<char *>(t->work);
  TrainArgs a{};
  a.x = t->x;
  a.t = t->targets;
  a.base_row = t->base_row;
  a.M = t->M;
  a.d_in = t->d_in;
  a.depth = depth;
  a.scale = t->scale;
  a.Bf = t->Bf;
  a.Wp = static_cast<const __bf16 *>(t->W_packed);
  a.wp_bytes = packed_train_bytes(W, depth);
  a.params = t->params;
  a.loss_scale = 1.0f / ((float)t->M * (float)O);
  a.FT = reinterpret_cast<__bf16 *>(wk + wl.ft);
  a.HT = reinterpret_cast<__bf16 *>(wk + wl.ht);
  a.DZT = reinterpret_cast<__bf16 *>(wk + wl.dzt);
  a.DACT = reinterpret_cast<__bf16 *>(wk + wl.dact);
  a.bias_part = reinterpret_cast<float *>(wk + wl.bias);
  a.wout_part = reinterpret_cast<float *>(wk + wl.wout);
  a.bout_part = reinterpret_cast<float *>(wk + wl.bout);
  a.loss_part = reinterpret_cast<float *>(wk + wl.loss);
  a.step = t->step;
  int rc;
  if (W == 256) {
    rc = O == 1 ? launch_train<256, 4, 1, 8>(a, st) : O == 2 ? launch_train<256, 4, 2, 8>(a, st)
                                                               : launch_train<256, 4, 3, 8>(a, st);
  } else {
    rc = O == 1 ? launch_train<512, 4, 1, 4>(a, st) : O == 2 ? launch_train<512, 4, 2, 4>(a, st)
                                                               : launch_train<512, 4, 3, 4>(a, st);
  }
  if (rc) return rc;
  // weight gradients: dW_l (W x fan) = DZT_l (W x M) HT_{l-1}^T, both K-contiguous, split over K,
  // every layer in one launch
  float *part = reinterpret_cast<float *>(wk + wl.part);
  GemmArgs gm[kGemmGroupMax];
  int64_t poff = 0, n_weights = 0, n_params = 0;
  for (int l = 0; l < depth; ++l) {
    const int fan = l == 0 ? mlpf::kFeat : W;
    GemmArgs &g = gm[l];
    g = GemmArgs{};
    g.A = a.DZT + (int64_t)l * W * t->M;
    g.B = l == 0 ? a.FT : a.HT + (int64_t)(l - 1) * W * t->M;
    g.lda = g.ldb = t->M;
    g.M = W;
    g.N = fan;
    g.K = t->M;
    g.k_per_split = t->M / S;
    g.C32 = part + poff;
    g.ldc = fan;
    g.z_rows = W;
    poff += (int64_t)S * W * fan;
    n_weights += (int64_t)W * fan;
    n_params += (int64_t)W * fan + W;
  }
  rc = launch_gemm_group_f32(gm, depth, st);
  if (rc) return rc;
  ReduceArgs r{};
  r.part = part;
  r.splits = S;
  r.bias_part = a.bias_part;
  r.wout_part = a.wout_part;
  r.bout_part = a.bout_part;
  r.loss_part = a.loss_part;
  r.tiles = t->M / R;
  r.depth = depth;
  r.W = W;
  r.O = O;
  r.n_weights = n_weights;
  r.n_tiled = depth * W + O * W + O;
  r.weight_blocks = (int)((n_weights + 255) / 256);
  r.grads = t->grads;
  r.loss = t->loss;
  r.base_row = t->base_row;
  r.M = t->M;
  const int blocks = r.weight_blocks + (r.n_tiled + 15) / 16 + 1;  // + the loss's work-group
  if (adam) {
    r.adam = *adam;
    train_reduce_kernel<true><<<blocks, 256, 0, st>>>(r);
  } else {
    train_reduce_kernel<false><<<blocks, 256, 0, st>>>(r);
  }
  (void)n_params;
  return check_launch("mlp_train_grads reduce");
}

}  // namespace
}  // namespace pyftle

extern "C" {

int pyftle_mlp_train_adam(float *params, float *m, float *v, const float *grads, int32_t width, int32_t depth,
                          int32_t d_out, float beta1, float beta2, float eps, float weight_decay, float lr_max,
                          float lr_min, int32_t total_steps, int32_t adam_step0, const int32_t *step,
                          void *packed, void *stream) {
  if (!params || !m || !v || !grads || !step || !packed) return fail("mlp_train_adam: NULL pointer");
  if ((width != 256 && width != 512) || depth < 2 || depth > mlpf::kMaxDepth || d_out < 1 || d_out > 3)
    return fail("mlp_train_adam: width 256 or 512, depth in [2, 8], d_out in [1, 3]");
  int64_t n = 0;
  for (int l = 0; l < depth; ++l) n += (int64_t)width * (l == 0 ? mlpf::kFeat : width) + width;
  n += (int64_t)d_out * width + d_out;
  AdamParams ad{};
  ad.w = params;
  ad.m = m;
  ad.v = v;
  ad.packed = static_cast<__bf16 *>(packed);
  ad.b1 = beta1;
  ad.b2 = beta2;
  ad.eps = eps;
  ad.wd = weight_decay;
  ad.lr_max = lr_max;
  ad.lr_min = lr_min;
  ad.total = total_steps;
  ad.t0 = adam_step0;
  ad.step = step;
  train_adam_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(ad, grads, n, width, depth);
  return check_launch("mlp_train_adam");
}

int pyftle_mlp_pack_train(const float *params, int32_t width, int32_t depth, void *packed, void *stream) {
  if (!params || !packed) return fail("mlp_pack_train: NULL pointer");
  if ((width != 256 && width != 512) || depth < 2 || depth > mlpf::kMaxDepth)
    return fail("mlp_pack_train: width 256 or 512, depth in [2, 8]");
  const int64_t n = (int64_t)width * mlpf::kFeat + (int64_t)(depth - 1) * width * width;
  pack_train_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(params, width, depth,
                                                                               static_cast<__bf16 *>(packed));
  return check_launch("mlp_pack_train");
}

}  // extern "C"
