// Fused surrogate inference: Fourier features -> depth x (256, SiLU) -> O outputs,
// with forward-mode tangents, in one kernel.
//
// No reference counterpart: the learned flow-map approximation of the RK4 path
// (integrate.py:265-342).  The layer-by-layer kernels of mlp.hip write every
// activation to HBM and read it back (about 87 flops per byte at width 256,
// below the MFMA ridge); here a wave keeps its seeds' activations in registers
// for the whole network and only the weights move.
//
// Formulation: each layer computes out^T = W in^T, so the weights are the MFMA
// A operand (output features x input features) and the activations the B
// operand (input features x seeds).  With v_mfma_f32_16x16x32_bf16 a lane
// holds, for its seed (lane % 16), the D values of features 4*(lane/16) + e of
// each 16-row tile; two adjacent tiles give the 8 values a lane needs as the
// next layer's B operand, provided the next layer's input-feature order is
// permuted inside each block of 32 (pyftle_mlp_pack_fused writes the weights in
// that order).  Tangent rows share the weight fragment: each A fragment read
// from LDS feeds 1 + n_tangents MFMAs.
//
// Work-group: 4 waves x 16 seeds per tile (x 3 seed groups without tangents),
// persistent over tiles.  Weights stream through LDS in chunks of 128 output features,
// double-buffered: the next chunk is copied by direct-to-LDS loads, spread over the
// k-steps, while the current one is multiplied.  Each chunk is multiplied as two
// 64-feature halves with alternating accumulator sets, so the SiLU epilogue of one half
// is issued between the MFMAs of the next.  The output layer (O <= 3) runs in
// FP32 on the last hidden layer's registers, reduced across the 4 lanes of a
// seed with cross-lane shuffles.

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cstdint>

#include "common.hpp"

namespace pyftle {
namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kWidth = 256;     // hidden width
constexpr int kFeat = 128;      // Fourier features: 64 frequencies, sin and cos
constexpr int kNFreq = kFeat / 2;
constexpr int kChunk = 128;     // output features per weight chunk
constexpr int kSeedsPerWave = 16;
constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;
constexpr int kMaxDepth = 8;
constexpr int kMaxIn = 4;
constexpr int kBufElems = kChunk * kWidth;  // one weight chunk (rows of K bf16, swizzled)

struct FusedArgs {
  const float *x;       // (M, d_in) inputs
  const float *scale;   // (d_in) input scaling before the Fourier projection
  const float *Bf;      // (64, d_in) frequencies, in revolutions per unit
  const __bf16 *W;      // packed hidden weights (pyftle_mlp_pack_fused order)
  int w_bytes;          // their size
  const float *bias;    // (depth, 256)
  const float *Wout;    // (O, 256)
  const float *bout;    // (O)
  float *Y;             // (1 + n_tangents, M, O)
  float out_scale;
  int d_in;
  int M;
  int depth;
};

__device__ __forceinline__ int64_t layer_offset(int L) {
  return L == 0 ? 0 : (int64_t)kWidth * kFeat + (int64_t)(L - 1) * kWidth * kWidth;
}

// Chunk c = (layer c / 2, output half c % 2) of the packed weights, copied global -> LDS
// by direct-to-LDS buffer loads (no registers): each wave instruction writes 64 lanes x
// 16 bytes contiguously.  LDS rows hold K bf16 unpadded; the 16-byte unit u of row R is
// stored at unit u ^ (R % 16), so the 16 rows read by one MFMA fragment land in distinct
// banks.  Row R % 16 of instruction i depends only on i and the lane, so the per-lane
// offsets repeat with period 16 / rows-per-instruction and the rest is a scalar offset.
template <int K>
__device__ __forceinline__ void chunk_dma(__amdgpu_buffer_rsrc_t w, int c, __bf16 *buf, int wave, int lane,
                                          int i0 = 0, int n = 64) {
  constexpr int kUnits = K / 8;              // 16-byte units per row
  constexpr int kRowsPerInst = 64 / kUnits;  // rows per wave instruction (1 KB)
  constexpr int kInstPerWave = kChunk / kRowsPerInst / kWaves;
  constexpr int kPatterns = 16 / kRowsPerInst;
  const int chunk_bytes = (int)((layer_offset(c >> 1) + (int64_t)(c & 1) * kChunk * K) * 2);
  const int rr = lane / kUnits, p = lane % kUnits;
#pragma unroll
  for (int i = 0; i < kInstPerWave; ++i) {
    if (i < i0 || i >= i0 + n) continue;
    const int blk = wave * kInstPerWave + i;  // 1 KB block of the chunk
    const int voff = rr * K * 2 + (p ^ (kRowsPerInst * (i % kPatterns) + rr)) * 16;
    __builtin_amdgcn_raw_ptr_buffer_load_lds(w, (__attribute__((address_space(3))) void *)(buf + blk * 512), 16,
                                             voff, chunk_bytes + blk * 1024, 0, 0);
  }
}

// Part `part` of `parts` of chunk c's copy (the instructions of one wave split evenly).
__device__ __forceinline__ void chunk_dma_part(__amdgpu_buffer_rsrc_t w, int c, __bf16 *buf, int wave, int lane,
                                               int part, int parts) {
  if (c < 2) {
    const int per = (kChunk / (64 / (kFeat / 8)) / kWaves) / parts;
    chunk_dma<kFeat>(w, c, buf, wave, lane, part * per, per);
  } else {
    const int per = (kChunk / (64 / (kWidth / 8)) / kWaves) / parts;
    chunk_dma<kWidth>(w, c, buf, wave, lane, part * per, per);
  }
}

// Wait for this wave's outstanding global loads (including the direct-to-LDS copies),
// then for the whole work-group.
__device__ __forceinline__ void wait_loads_and_sync() {
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  __syncthreads();
}

// acc[b][t] = W_chunk rows row0 + 16 t .. (t < 4: one sub-chunk of 64 output features) x
// in[b] over all K input features (16 seeds per row block).  The 4 weight fragments of
// k-step ks+1 are read from LDS before the MFMAs of k-step ks.  Piece ks of the previous
// sub-chunk's epilogue, epi(ks), is issued between the MFMAs of k-step ks: it reads only
// the other accumulator set, so its VALU work runs while the matrix core is busy.
template <int K, int NB, typename Epi>
__device__ __forceinline__ void sub_mma(const __bf16 *buf, int row0, const bf16x8_t (&in)[NB][8],
                                        f32x4_t (&acc)[NB][4], int lane, Epi &&epi) {
  const int r = lane & 15, g = lane >> 4;
  const __bf16 *base = buf + (row0 + r) * K;
  const int sw = g ^ r;  // unit ks*4 + g of row row0 + 16 t + r sits at unit (ks*4) ^ sw
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[b][t] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  bf16x8_t frag[2][4];
#pragma unroll
  for (int t = 0; t < 4; ++t) frag[0][t] = *reinterpret_cast<const bf16x8_t *>(base + t * 16 * K + sw * 8);
#pragma unroll
  for (int ks = 0; ks < K / 32; ++ks) {
    if (ks + 1 < K / 32) {
#pragma unroll
      for (int t = 0; t < 4; ++t)
        frag[(ks + 1) & 1][t] =
            *reinterpret_cast<const bf16x8_t *>(base + t * 16 * K + (((ks + 1) * 4) ^ sw) * 8);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int b = 0; b < NB; ++b)
        acc[b][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(frag[ks & 1][t], in[b][ks], acc[b][t], 0, 0, 0);
    epi(ks);
#pragma unroll
    for (int i = 0; i < 4 * NB; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
      __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);  // then up to 4 VALU
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// 1 / (1 + e^-z) with the hardware exp2 and reciprocal (about 1 ulp each; the results are
// rounded to bf16 or summed in FP32 afterwards).  e^-z overflows to +inf for z < -88 and the
// reciprocal then gives 0, the correct limit.
__device__ __forceinline__ float sigmoid_f(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.4426950408889634f));
}

// Epilogues, one piece at a time: piece p of a sub-chunk covers elements
// [p*EP, p*EP + EP) of its 16 per-lane values (tile t = p*EP / 4), EP = 16 / pieces.
// Row blocks of a wave: G seed groups of 16, each followed by its NT tangent blocks
// (block q * (1 + NT) + 1 + a holds d/d x_a of group q).
//
// Hidden layer: h = SiLU(z + bias), tangents * SiLU'(z), packed as the next layer's B
// operand: output feature mc*128 + tt*16 + 4g + e (tt = 4 h + t) goes to k-step
// mc*4 + tt/2, element (tt%2)*4 + e.
template <int G, int NT, int PIECES>
__device__ __forceinline__ void hidden_piece(const f32x4_t (&acc)[G * (1 + NT)][4], const float *bias_l,
                                             int mc, int h, int p, bf16x8_t (&out)[G * (1 + NT)][8], int lane) {
  constexpr int EP = 16 / PIECES;
  const int g = lane >> 4;
  const int t = (p * EP) >> 2, e0 = (p * EP) & 3, tt = h * 4 + t;
  const f32x4_t bv = *reinterpret_cast<const f32x4_t *>(bias_l + mc * kChunk + tt * 16 + 4 * g);
  const int ks = mc * 4 + (tt >> 1);
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int b0 = q * (1 + NT);
#pragma unroll
    for (int e = e0; e < e0 + EP; ++e) {
      const float z = acc[b0][t][e] + bv[e];
      const float s = sigmoid_f(z);
      const int j = (tt & 1) * 4 + e;
      out[b0][ks][j] = (__bf16)(z * s);
      if (NT > 0) {
        const float d = s * (1.0f + z * (1.0f - s));
#pragma unroll
        for (int a = 1; a <= NT; ++a) out[b0 + a][ks][j] = (__bf16)(acc[b0 + a][t][e] * d);
      }
    }
  }
}

// Last hidden layer: the output layer's partial sums over this lane's features.
template <int G, int NT, int O, int PIECES>
__device__ __forceinline__ void output_piece(const f32x4_t (&acc)[G * (1 + NT)][4], const float *bias_l,
                                             const float *wout, int mc, int h, int p,
                                             float (&yp)[G * (1 + NT)][O], int lane) {
  constexpr int EP = 16 / PIECES;
  const int g = lane >> 4;
  const int t = (p * EP) >> 2, e0 = (p * EP) & 3, tt = h * 4 + t;
  const int f0 = mc * kChunk + tt * 16 + 4 * g;
  const f32x4_t bv = *reinterpret_cast<const f32x4_t *>(bias_l + f0);
  f32x4_t wv[O];
#pragma unroll
  for (int o = 0; o < O; ++o) wv[o] = *reinterpret_cast<const f32x4_t *>(wout + o * kWidth + f0);
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int b0 = q * (1 + NT);
#pragma unroll
    for (int e = e0; e < e0 + EP; ++e) {
      const float z = acc[b0][t][e] + bv[e];
      const float s = sigmoid_f(z);
      const float hz = z * s;
      const float d = s * (1.0f + z * (1.0f - s));
#pragma unroll
      for (int o = 0; o < O; ++o) {
        yp[b0][o] = fmaf(hz, wv[o][e], yp[b0][o]);
#pragma unroll
        for (int a = 1; a <= NT; ++a) yp[b0 + a][o] = fmaf(acc[b0 + a][t][e] * d, wv[o][e], yp[b0 + a][o]);
      }
    }
  }
}

// One layer (K input features) of one tile: chunk 2L (output features 0..127) from LDS
// buffer 0, chunk 2L+1 (128..255) from buffer 1, each as two 64-feature sub-chunks with
// the accumulators alternating between accA and accB.  The epilogue of each sub-chunk
// runs under the MFMAs of the next one, except the last (the next layer needs all of this
// layer's outputs).  The copy of chunk 2L+1 is issued before chunk 2L is multiplied and
// the copy of the chunk after 2L+1 (the next layer's first, or chunk 0 of the next tile)
// before chunk 2L+1.
template <int K, int G, int NT, int O, bool LAST>
__device__ __forceinline__ void run_layer(int L, int nchunks, __amdgpu_buffer_rsrc_t wr, __bf16 *wbuf0,
                                          __bf16 *wbuf1, const float *bias_s, const float *wout_s,
                                          const bf16x8_t (&in)[G * (1 + NT)][8],
                                          bf16x8_t (&out)[G * (1 + NT)][8], float (&yp)[G * (1 + NT)][O],
                                          f32x4_t (&accA)[G * (1 + NT)][4], f32x4_t (&accB)[G * (1 + NT)][4],
                                          int wave, int lane) {
  constexpr int NB = G * (1 + NT), P = K / 32;
  const float *bl = bias_s + L * kWidth;
  auto piece = [&](const f32x4_t(&acc)[NB][4], int mc, int h, int p) {
    if constexpr (LAST) output_piece<G, NT, O, P>(acc, bl, wout_s, mc, h, p, yp, lane);
    else hidden_piece<G, NT, P>(acc, bl, mc, h, p, out, lane);
  };
  // the copies are spread over the k-steps of the first sub-chunk of each chunk
  const int c1 = 2 * L + 1, c2 = (2 * L + 2) % nchunks;
  sub_mma<K, NB>(wbuf0, 0, in, accA, lane, [&](int p) { chunk_dma_part(wr, c1, wbuf1, wave, lane, p, P); });
  sub_mma<K, NB>(wbuf0, 64, in, accB, lane, [&](int p) { piece(accA, 0, 0, p); });
  wait_loads_and_sync();
  sub_mma<K, NB>(wbuf1, 0, in, accA, lane, [&](int p) {
    chunk_dma_part(wr, c2, wbuf0, wave, lane, p, P);
    piece(accB, 0, 1, p);
  });
  sub_mma<K, NB>(wbuf1, 64, in, accB, lane, [&](int p) { piece(accA, 1, 0, p); });
#pragma unroll
  for (int p = 0; p < P; ++p) piece(accB, 1, 1, p);
  wait_loads_and_sync();
}

// Persistent: work-group b handles seed tiles b, b + gridDim.x, ...  The weight chunks
// cycle 0, 1, ..., 2*depth - 1, 0, 1, ... across tiles; chunk c always lives in LDS
// buffer c % 2 (2*depth is even), and the copy of the chunk after the current one is
// in flight while the current one is multiplied.
template <int G, int NT, int O>
__global__ void __launch_bounds__(kThreads, 1) fused_mlp_kernel(FusedArgs a) {
  constexpr int NB = G * (1 + NT);
  constexpr int kTileSeeds = kWaves * G * kSeedsPerWave;
  // two separate arrays (not halves of one) so the compiler can tell the copy into one
  // buffer from the reads of the other and does not wait for the copy before the reads
  __shared__ __attribute__((aligned(16))) __bf16 wbuf0[kBufElems];
  __shared__ __attribute__((aligned(16))) __bf16 wbuf1[kBufElems];
  __shared__ __attribute__((aligned(16))) float bias_s[kMaxDepth * kWidth];
  __shared__ __attribute__((aligned(16))) float wout_s[O * kWidth];
  __shared__ __attribute__((aligned(16))) float freq_s[kNFreq * kMaxIn];  // B * scale, zero-padded
  __shared__ __attribute__((aligned(16))) float dfreq_s[kNFreq * kMaxIn]; // 2 pi B * scale

  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar offsets, M0
  const int r = lane & 15, g = lane >> 4;
  const int nchunks = 2 * a.depth;
  const int64_t ntiles = ((int64_t)a.M + kTileSeeds - 1) / kTileSeeds;

  const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc((void *)a.W, 0, a.w_bytes, 0x00020000);
  chunk_dma<kFeat>(wr, 0, wbuf0, wave, lane);
  {
    // per-work-group constants: all global loads issued before the LDS stores
    float bv[kMaxDepth], wv[O], fv = 0.0f;
#pragma unroll
    for (int i = 0; i < kMaxDepth; ++i) bv[i] = i < a.depth ? a.bias[i * kWidth + tid] : 0.0f;
#pragma unroll
    for (int o = 0; o < O; ++o) wv[o] = a.Wout[o * kWidth + tid];
    const int q = tid >> 2, c = tid & 3;  // kThreads == kNFreq * kMaxIn
    float sc = 0.0f;
    if (c < a.d_in) {
      fv = a.Bf[q * a.d_in + c];
      sc = a.scale[c];
    }
#pragma unroll
    for (int i = 0; i < kMaxDepth; ++i) bias_s[i * kWidth + tid] = bv[i];
#pragma unroll
    for (int o = 0; o < O; ++o) wout_s[o * kWidth + tid] = wv[o];
    freq_s[tid] = fv * sc;
    dfreq_s[tid] = 6.283185307179586f * fv * sc;
  }

  // seed of group q of this lane in tile T: T * kTileSeeds + (wave * G + q) * 16 + r
  int64_t tile = blockIdx.x;
  float xin[G][kMaxIn];
#pragma unroll
  for (int q = 0; q < G; ++q) {
    const int64_t seed = tile * kTileSeeds + (wave * G + q) * kSeedsPerWave + r;
#pragma unroll
    for (int c = 0; c < kMaxIn; ++c) xin[q][c] = (seed < a.M && c < a.d_in) ? a.x[seed * a.d_in + c] : 0.0f;
  }
  wait_loads_and_sync();

  for (; tile < ntiles; tile += gridDim.x) {
    const int64_t seed0 = tile * kTileSeeds + wave * G * kSeedsPerWave + r;  // group q: + 16 q

    // Fourier features of this lane's seeds, already in the B-operand layout:
    // k-step ks, element j <-> feature 32*ks + 8*g + j ([sin | cos] of 64 frequencies).
    bf16x8_t in[NB][8];
#pragma unroll
    for (int ks = 0; ks < kNFreq / 32; ++ks)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int fq = ks * 32 + g * 8 + j;
        const f32x4_t fr = *reinterpret_cast<const f32x4_t *>(freq_s + fq * kMaxIn);
        f32x4_t dq = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
        if (NT > 0) dq = *reinterpret_cast<const f32x4_t *>(dfreq_s + fq * kMaxIn);
#pragma unroll
        for (int q = 0; q < G; ++q) {
          float ph = 0.0f;  // revolutions
#pragma unroll
          for (int c = 0; c < kMaxIn; ++c) ph += fr[c] * xin[q][c];
          ph -= rintf(ph);
          const float sn = __builtin_amdgcn_sinf(ph), cs = __builtin_amdgcn_cosf(ph);
          const int b0 = q * (1 + NT);
          in[b0][ks][j] = (__bf16)sn;
          in[b0][ks + kNFreq / 32][j] = (__bf16)cs;
#pragma unroll
          for (int a2 = 1; a2 <= NT; ++a2) {
            in[b0 + a2][ks][j] = (__bf16)(cs * dq[a2 - 1]);
            in[b0 + a2][ks + kNFreq / 32][j] = (__bf16)(-sn * dq[a2 - 1]);
          }
        }
      }
    // the next tile's inputs load while this tile is multiplied
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int64_t nseed = seed0 + q * kSeedsPerWave + (int64_t)gridDim.x * kTileSeeds;
#pragma unroll
      for (int c = 0; c < kMaxIn; ++c)
        xin[q][c] = (nseed < a.M && c < a.d_in) ? a.x[nseed * a.d_in + c] : 0.0f;
    }

    f32x4_t accA[NB][4], accB[NB][4];
    bf16x8_t out[NB][8];
    float yp[NB][O];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int o = 0; o < O; ++o) yp[b][o] = 0.0f;

    run_layer<kFeat, G, NT, O, false>(0, nchunks, wr, wbuf0, wbuf1, bias_s, wout_s, in, out, yp, accA, accB,
                                      wave, lane);
    for (int L = 1; L < a.depth; ++L) {
#pragma unroll
      for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int k = 0; k < 8; ++k) in[b][k] = out[b][k];
      if (L == a.depth - 1)
        run_layer<kWidth, G, NT, O, true>(L, nchunks, wr, wbuf0, wbuf1, bias_s, wout_s, in, out, yp, accA, accB,
                                          wave, lane);
      else
        run_layer<kWidth, G, NT, O, false>(L, nchunks, wr, wbuf0, wbuf1, bias_s, wout_s, in, out, yp, accA,
                                           accB, wave, lane);
    }

    // the 4 lanes of a seed (lane % 16 equal) hold disjoint feature subsets
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int o = 0; o < O; ++o) {
        float v = yp[b][o];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        yp[b][o] = v;
      }
    if (g == 0) {
#pragma unroll
      for (int q = 0; q < G; ++q) {
        const int64_t seed = seed0 + q * kSeedsPerWave;
        if (seed < a.M) {
#pragma unroll
          for (int a2 = 0; a2 <= NT; ++a2)
#pragma unroll
            for (int o = 0; o < O; ++o) {
              const float v = yp[q * (1 + NT) + a2][o];
              a.Y[((int64_t)a2 * a.M + seed) * O + o] = (a2 == 0 ? v + a.bout[o] : v) * a.out_scale;
            }
        }
      }
    }
  }
}

// Input-feature order inside each block of 32 for layers fed by a hidden layer:
// position p = 8*g + j holds feature 4*g + j (j < 4) or 16 + 4*g + (j - 4).
__device__ __forceinline__ int fused_perm(int p) {
  const int g = p >> 3, j = p & 7;
  return j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4);
}

__global__ void pack_fused_kernel(const float *W, int rows, int K, int permute, __bf16 *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * K) return;
  const int o = (int)(i / K), k = (int)(i % K);
  const int src = permute ? (k & ~31) + fused_perm(k & 31) : k;
  out[i] = (__bf16)W[(int64_t)o * K + src];
}

static_assert(kThreads == kNFreq * kMaxIn && kThreads == kWidth, "constant staging maps one thread per entry");

template <int G, int NT, int O>
void launch_fused(const FusedArgs &a, hipStream_t st) {
  // one work-group per CU (the LDS weight buffers admit one); each loops over seed tiles
  const int64_t tile_seeds = (int64_t)kWaves * G * kSeedsPerWave;
  const int64_t tiles = ((int64_t)a.M + tile_seeds - 1) / tile_seeds;
  const unsigned grid = (unsigned)std::min<int64_t>(tiles, device_cu_count());
  fused_mlp_kernel<G, NT, O><<<grid, kThreads, 0, st>>>(a);
}

}  // namespace
}  // namespace pyftle

using namespace pyftle;

extern "C" {

int pyftle_mlp_pack_fused(const float *W, int32_t rows, int32_t K, int32_t permute, void *out, void *stream) {
  if (!W || !out) return fail("mlp_pack_fused: NULL pointer");
  if (rows < 1 || K < 32 || K % 32) return fail("mlp_pack_fused: K must be a positive multiple of 32");
  const int64_t n = (int64_t)rows * K;
  pack_fused_kernel<<<(unsigned)((n + 255) / 256), 256, 0, as_stream(stream)>>>(W, rows, K, permute,
                                                                                (__bf16 *)out);
  return check_launch("mlp_pack_fused");
}

int pyftle_mlp_fused_forward(const float *x, int32_t d_in, int32_t M, const float *scale, const float *Bf,
                             int32_t n_freq, const void *W_packed, const float *bias, int32_t depth,
                             int32_t width, const float *Wout, const float *bout, int32_t O,
                             float out_scale, int32_t n_tangents, float *Y, void *stream) {
  if (!x || !scale || !Bf || !W_packed || !bias || !Wout || !bout || !Y)
    return fail("mlp_fused_forward: NULL pointer");
  if (n_freq != kNFreq || width != kWidth)
    return fail("mlp_fused_forward: needs 64 Fourier frequencies and hidden width 256");
  if (depth < 2 || depth > kMaxDepth) return fail("mlp_fused_forward: depth must be in [2, 8]");
  if (d_in < 1 || d_in > kMaxIn || M < 0 || n_tangents < 0 || n_tangents > d_in)
    return fail("mlp_fused_forward: bad input sizes");
  if ((uintptr_t)W_packed % 16) return fail("mlp_fused_forward: packed weights must be 16-byte aligned");
  if (M == 0) return PYFTLE_OK;
  const int w_bytes = 2 * (kWidth * kFeat + (depth - 1) * kWidth * kWidth);
  FusedArgs a{x, scale, Bf, (const __bf16 *)W_packed, w_bytes, bias, Wout, bout, Y, out_scale, d_in, M, depth};
  hipStream_t st = as_stream(stream);
  // row blocks per wave: 3 groups of 16 seeds without tangents, 1 group + 2 tangents with
  if (O == 2 && n_tangents == 0) launch_fused<3, 0, 2>(a, st);
  else if (O == 2 && n_tangents == 2) launch_fused<1, 2, 2>(a, st);
  else if (O == 3 && n_tangents == 0) launch_fused<3, 0, 3>(a, st);
  else return fail("mlp_fused_forward: supported (outputs, tangents): (2, 0), (2, 2), (3, 0)");
  return check_launch("mlp_fused_forward");
}

}  // extern "C"
