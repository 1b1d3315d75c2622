// Flow-map surrogate kernels: Fourier-feature MLP with BF16 MFMA GEMMs and
// FP32 accumulation (gfx950 v_mfma_f32_32x32x16_bf16), fused epilogues,
// Fourier encoding with analytic input tangents, the narrow output layer,
// MSE loss and Adam.  No reference counterpart (the surrogate is this
// engine's addition); the oracle is a NumPy FP64 restatement.
//
// GEMM: C[M,N] = sum_k A[m,k] B[k,n] over a 128x128 output tile per
// work-group (4 waves, 2x2 of 64x64, each 2x2 MFMA 32x32 blocks), K staged
// through LDS 32 at a time.  Operand layouts: A "k-contiguous" (row-major
// MxK) or "m-contiguous" (A^T stored row-major), B "k-contiguous" (B^T stored
// row-major, i.e. a weight stored [out][in]) or "n-contiguous" (row-major KxN).
// MFMA 32x32x16 operand layout: lane l holds row (l % 32) and k-slice
// 8*(l / 32) .. +8; the 32x32 FP32 result gives lane l column (l % 32) and
// rows 8*(e / 4) + 4*(l / 32) + e % 4 for element e of its 16.
#include <hip/hip_bf16.h>

#include <cmath>
#include <string>

#include "common.hpp"
#include "mlp_gemm.hpp"

namespace pyftle {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

enum { EPI_F32 = 0, EPI_BIAS_SILU = 1, EPI_DSILU = 2, EPI_BIAS = 3 };

// GemmArgs: mlp_gemm.hpp

// Logistic sigmoid with the hardware exp2 and reciprocal (about 1 ulp each; results are
// rounded to bf16 afterwards).  e^-z overflows to +inf for z < -88 and the reciprocal
// then gives 0, the correct limit.
__device__ __forceinline__ float sigmoid_hw(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.4426950408889634f));
}
__device__ __forceinline__ float silu_f(float z) { return z * sigmoid_hw(z); }
__device__ __forceinline__ float dsilu_f(float z) {
  const float s = sigmoid_hw(z);
  return s * (1.0f + z * (1.0f - s));
}

constexpr int BM = 128, BN = 128, BK = 64, LDSK = BK + 8;  // padded LDS rows (bank spread)
// The training step's weight-gradient group stages K 128 deep: twice the MFMAs per barrier
// and per global-load round trip (its work-groups have a CU each and hide nothing else).
constexpr int BK_GROUP = 128;

// A 128 x 64 (rows x k) operand tile held in registers between its global
// load and its LDS store, so the next tile's loads overlap the MFMAs of the
// current one.  KC: the source is k-contiguous (element (r,k) at
// src[r*ld + k]); otherwise r-contiguous (element (r,k) at src[k*ld + r]).
template <bool KC, int BKT = BK>
struct TileLoader {
  static constexpr int kReps = 128 * BKT / 8 / 256;  // 16-byte chunks per thread
  static constexpr int kChunksPerRow = BKT / 8;
  bf16x8 v[kReps];
  __device__ __forceinline__ void load(const __bf16 *src, int64_t ld, int r0, int k0, int tid) {
#pragma unroll
    for (int rep = 0; rep < kReps; ++rep) {
      const int c = tid + 256 * rep;  // 128 * BKT / 8 chunks of 8 elements
      if (KC) {
        const int row = c / kChunksPerRow, kq = (c % kChunksPerRow) * 8;
        v[rep] = *reinterpret_cast<const bf16x8 *>(src + (int64_t)(r0 + row) * ld + k0 + kq);
      } else {
        const int k = c >> 4, rq = (c & 15) * 8;
        v[rep] = *reinterpret_cast<const bf16x8 *>(src + (int64_t)(k0 + k) * ld + r0 + rq);
      }
    }
  }
  __device__ __forceinline__ void store(__bf16 (*sm)[BKT + 8], int tid) const {
#pragma unroll
    for (int rep = 0; rep < kReps; ++rep) {
      const int c = tid + 256 * rep;
      if (KC) {
        const int row = c / kChunksPerRow, kq = (c % kChunksPerRow) * 8;
        *reinterpret_cast<bf16x8 *>(&sm[row][kq]) = v[rep];
      } else {
        const int k = c >> 4, rq = (c & 15) * 8;
#pragma unroll
        for (int e = 0; e < 8; ++e) sm[rq + e][k] = v[rep][e];
      }
    }
  }
};

template <bool A_KC, bool B_KC, int EPI, int BKT = BK>
__device__ __forceinline__ void gemm_tile(const GemmArgs &g, const int bx, const int by, const int bz) {
  // One LDS buffer: operand tiles in the main loop, the output tile after it.
  constexpr int LK = BKT + 8;  // padded LDS rows (bank spread)
  __shared__ __attribute__((aligned(16))) __bf16 smem[(BM + BN) * LK];
  __bf16(*As)[LK] = reinterpret_cast<__bf16(*)[LK]>(smem);
  __bf16(*Bs)[LK] = reinterpret_cast<__bf16(*)[LK]>(smem + BM * LK);
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = by * BM, n0 = bx * BN;
  const int kb = bz * g.k_per_split;
  const int ke = min(g.K, kb + g.k_per_split);
  f32x16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;

  const int lr = lane & 31, lk = (lane >> 5) * 8;
  TileLoader<A_KC, BKT> la;
  TileLoader<B_KC, BKT> lb;
  la.load(g.A, g.lda, m0, kb, tid);
  lb.load(g.B, g.ldb, n0, kb, tid);
  la.store(As, tid);
  lb.store(Bs, tid);
  __syncthreads();
  for (int k0 = kb; k0 < ke; k0 += BKT) {
    const bool more = k0 + BKT < ke;
    if (more) {
      la.load(g.A, g.lda, m0, k0 + BKT, tid);
      lb.load(g.B, g.ldb, n0, k0 + BKT, tid);
    }
#pragma unroll
    for (int ks = 0; ks < BKT; ks += 16) {
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8 *>(&As[wm * 64 + i * 32 + lr][ks + lk]);
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[wn * 64 + j * 32 + lr][ks + lk]);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
    if (more) {
      la.store(As, tid);
      lb.store(Bs, tid);
      __syncthreads();
    }
  }

  if constexpr (EPI == EPI_F32) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int row = m0 + wm * 64 + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
          const int col = n0 + wn * 64 + j * 32 + (lane & 31);
          g.C32[(int64_t)bz * g.M * g.ldc + (int64_t)row * g.ldc + col] = acc[i][j][e];
        }
  } else {
    // Stage the bf16 tile in LDS (the operand buffers are free after the
    // main loop), then write 16-byte chunks per lane, coalesced along rows.
    static_assert(BM * (BN + 8) <= (BM + BN) * LK, "the output tile fits in the operand buffer");
    __bf16(*Ct)[BN + 8] = reinterpret_cast<__bf16(*)[BN + 8]>(smem);  // EPI_BIAS_SILU: Z, then SiLU(Z)
#pragma unroll
    for (int pass = 0; pass < (EPI == EPI_BIAS_SILU ? 2 : 1); ++pass) {
      if (EPI == EPI_BIAS_SILU && pass == 0 && !g.Z) continue;
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int e = 0; e < 16; ++e) {
            const int r = wm * 64 + i * 32 + (e >> 2) * 8 + (lane >> 5) * 4 + (e & 3);
            const int cc = wn * 64 + j * 32 + (lane & 31);
            float v = acc[i][j][e];
            if (EPI == EPI_BIAS_SILU || EPI == EPI_BIAS) v += (g.bias ? g.bias[n0 + cc] : 0.0f);
            if (EPI == EPI_BIAS_SILU && pass == 1) v = silu_f(v);
            Ct[r][cc] = (__bf16)v;
          }
      __syncthreads();
      __bf16 *dst = (EPI == EPI_BIAS_SILU && pass == 0) ? g.Z : g.Cb;
      const int zoff = (EPI == EPI_DSILU) ? (m0 / g.z_rows) * g.z_rows : 0;
#pragma unroll
      for (int rep = 0; rep < 8; ++rep) {
        const int c = tid + 256 * rep;  // 2048 chunks of 8
        const int r = c >> 4, cq = (c & 15) * 8;
        bf16x8 v = *reinterpret_cast<const bf16x8 *>(&Ct[r][cq]);
        const int64_t o = (int64_t)(m0 + r) * g.ldc + n0 + cq;
        if (EPI == EPI_DSILU) {
          const bf16x8 z = *reinterpret_cast<const bf16x8 *>(g.Z + (int64_t)(m0 + r - zoff) * g.ldc + n0 + cq);
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] = (__bf16)((float)v[q] * dsilu_f((float)z[q]));
        }
        *reinterpret_cast<bf16x8 *>(dst + o) = v;
      }
      __syncthreads();
    }
  }
}

template <bool A_KC, bool B_KC, int EPI>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
  gemm_tile<A_KC, B_KC, EPI>(g, blockIdx.x, blockIdx.y, blockIdx.z);
}

struct GemmGroup {
  GemmArgs g[kGemmGroupMax];
  int start[kGemmGroupMax + 1];  // first work-group of each member
  int n;
  int xcd;                       // deal the tiles of a split to one XCD (see gemm_group_kernel)
};

// Work-groups are dealt to the 8 XCDs round robin (work-group b to XCD b % 8), and each XCD
// has its own L2.  With xcd set (every member's work-group count and start a multiple of 8,
// its split count too), the tiles of one split -- which share their operand rows -- all go
// to one XCD: work-group b of a member is XCD b % 8, slot b / 8 there, and takes split
// (b % 8) + 8 (slot / tiles), tile slot % tiles.
__global__ void __launch_bounds__(256, 2) gemm_group_kernel(GemmGroup G) {
  const int b = blockIdx.x;
  int j = 0;
  while (j + 1 < G.n && b >= G.start[j + 1]) ++j;
  const GemmArgs &g = G.g[j];
  const int nx = g.N / BN, ny = g.M / BM, nt = nx * ny;
  const int loc = b - G.start[j];
  int tile, split;
  if (G.xcd) {
    const int slot = loc >> 3;
    split = (loc & 7) + 8 * (slot / nt);
    tile = slot % nt;
  } else {
    tile = loc % nt;
    split = loc / nt;
  }
  gemm_tile<true, true, EPI_F32, BK_GROUP>(g, tile % nx, tile / nx, split);
}

int launch_gemm_group_f32(const GemmArgs *members, int n, hipStream_t st) {
  if (n < 1 || n > kGemmGroupMax) return fail("gemm group: 1..8 members");
  GemmGroup G{};
  G.n = n;
  G.xcd = 1;
  int total = 0;
  for (int j = 0; j < n; ++j) {
    if (members[j].k_per_split % BK_GROUP || members[j].K % members[j].k_per_split)
      return fail("gemm group: each split must cover a multiple of 128 of K");
    G.g[j] = members[j];
    G.start[j] = total;
    const int splits = members[j].K / members[j].k_per_split;
    if (splits % 8) G.xcd = 0;  // then also the starts stay multiples of 8
    total += (members[j].N / BN) * (members[j].M / BM) * splits;
  }
  G.start[n] = total;
  gemm_group_kernel<<<total, 256, 0, st>>>(G);
  return check_launch("gemm group");
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------

// Fourier features of v = (in * scale): z = [sin(2 pi B v), cos(2 pi B v)],
// B is (F, d) FP32.  With n_tan > 0 also writes d(z)/d(in_a) for the first
// n_tan input coordinates into rows M*(1+a) .. of `out` (tangent blocks).
__global__ void __launch_bounds__(256) fourier_kernel(const float *in, int d, const float *scale,
                                                      const float *Bm, int F, int M, int ld_out,
                                                      int n_tan, __bf16 *out) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)M * F) return;
  const int m = (int)(idx / F), f = (int)(idx % F);
  float ph = 0.0f;  // phase in revolutions
  for (int a = 0; a < d; ++a) ph += Bm[f * d + a] * in[(int64_t)m * d + a] * scale[a];
  ph -= rintf(ph);  // the hardware sine takes revolutions in a bounded range
  const float s = __builtin_amdgcn_sinf(ph);  // sin(2 pi ph)
  const float c = __builtin_amdgcn_cosf(ph);
  out[(int64_t)m * ld_out + f] = (__bf16)s;
  out[(int64_t)m * ld_out + F + f] = (__bf16)c;
  const float tp = 6.283185307179586f;
  for (int a = 0; a < n_tan; ++a) {
    const float dph = tp * Bm[f * d + a] * scale[a];
    const int64_t r = (int64_t)(a + 1) * M + m;
    out[r * ld_out + f] = (__bf16)(c * dph);
    out[r * ld_out + F + f] = (__bf16)(-s * dph);
  }
}

// Narrow output layer y[m, o] = sum_k h[m, k] W[o, k] (+ b[o]), FP32 out.
// One thread per row, 16-byte loads of the row, W staged in LDS (O*K <= 4096);
// `rows_with_bias` rows get the bias (tangent rows do not).
__global__ void __launch_bounds__(256) linear_out_kernel(const __bf16 *H, int64_t ldh, int K,
                                                         const float *W, const float *b, int O, int M,
                                                         int rows_with_bias, float *Y) {
  __shared__ float Ws[4096];
  for (int i = threadIdx.x; i < O * K; i += blockDim.x) Ws[i] = W[i];
  __syncthreads();
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  const __bf16 *row = H + m * ldh;
  for (int k = 0; k < K; k += 8) {
    const bf16x8 v = *reinterpret_cast<const bf16x8 *>(row + k);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float x = (float)v[q];
#pragma unroll
      for (int o = 0; o < 4; ++o)
        if (o < O) acc[o] = fmaf(x, Ws[o * K + k + q], acc[o]);
    }
  }
#pragma unroll
  for (int o = 0; o < 4; ++o)
    if (o < O) Y[m * O + o] = acc[o] + ((m < rows_with_bias && b) ? b[o] : 0.0f);
}

// Backward of the output layer for one row: dH[m, k] = sum_o dY[m, o] W[o, k]
// times SiLU'(Z[m, k]) (the last hidden layer's pre-activation), as bf16.
__global__ void __launch_bounds__(256) linear_out_bwd_kernel(const float *dY, const float *W, int O,
                                                             int K, int M, const __bf16 *Z,
                                                             int64_t ld, __bf16 *dZ) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)M * K) return;
  const int m = (int)(idx / K), k = (int)(idx % K);
  float s = 0.0f;
  for (int o = 0; o < O; ++o) s += dY[(int64_t)m * O + o] * W[o * K + k];
  dZ[(int64_t)m * ld + k] = (__bf16)(s * dsilu_f((float)Z[(int64_t)m * ld + k]));
}

// Partial output-layer weight gradient over a chunk of rows:
// dWp[c, o, k] = sum_{m in chunk c} dY[m, o] H[m, k]; dbp[c, o] = sum dY[m, o].
__global__ void __launch_bounds__(256) linear_out_wgrad_kernel(const float *dY, const __bf16 *H,
                                                               int64_t ldh, int K, int O, int M,
                                                               int rows, float *dWp, float *dbp) {
  const int o = blockIdx.y;
  const int c = blockIdx.z;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int mb = c * rows, me = min(M, mb + rows);
  if (k < K) {
    float s = 0.0f;
    for (int m = mb; m < me; ++m) s += dY[(int64_t)m * O + o] * (float)H[(int64_t)m * ldh + k];
    dWp[((int64_t)c * O + o) * K + k] = s;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float s = 0.0f;
    for (int m = mb; m < me; ++m) s += dY[(int64_t)m * O + o];
    dbp[(int64_t)c * O + o] = s;
  }
}

// Sum the split partials: out[i] = sum_s part[s*n + i].
__global__ void __launch_bounds__(256) reduce_splits_kernel(const float *part, int splits, int64_t n,
                                                            float *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int k = 0; k < splits; ++k) s += part[(int64_t)k * n + i];
  out[i] = s;
}

// Same sum with many partials: a block reduces 4 outputs, 64 threads per output each
// summing every 64th partial in order, then a fixed-order tree in LDS (deterministic,
// like the one-thread-per-output kernel).
__global__ void __launch_bounds__(256) reduce_splits_wide_kernel(const float *part, int splits, int64_t n,
                                                                 float *out) {
  __shared__ float red[4][64];
  const int l = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int64_t i = (int64_t)blockIdx.x * 4 + q;
  float s = 0.0f;
  if (i < n)
    for (int k = l; k < splits; k += 64) s += part[(int64_t)k * n + i];
  red[q][l] = s;
  __syncthreads();
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) {
    if (l < w) red[q][l] += red[q][l + w];
    __syncthreads();
  }
  if (l == 0 && i < n) out[i] = red[q][0];
}

// Partial column sums over a chunk of rows: part[c, n] = sum_{m in chunk c} X[m, n].
__global__ void __launch_bounds__(256) colsum_kernel(const __bf16 *X, int64_t ld, int M, int N, int rows,
                                                     float *part) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (n >= N) return;
  const int mb = c * rows, me = min(M, mb + rows);
  float s = 0.0f;
  for (int m = mb; m < me; ++m) s += (float)X[(int64_t)m * ld + n];
  part[(int64_t)c * N + n] = s;
}

// MSE loss gradient: dY = 2 (Y - T) / (M * O); loss partial sums per block.
__global__ void __launch_bounds__(256) mse_kernel(const float *Y, const float *T, int64_t n, float scale,
                                                  float *dY, float *loss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.0f;
  if (i < n) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d * scale;
    e = d * d * scale;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, e);
}

// Adam (Kingma & Ba 2015) on FP32 master weights; writes the bf16 copy.
__global__ void __launch_bounds__(256) adam_kernel(float *w, float *m, float *v, const float *g, int64_t n,
                                                   float lr, float b1, float b2, float eps, float bc1,
                                                   float bc2, float wd, __bf16 *wb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * w[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (__bf16)wi;
}

// Adam with the cosine learning-rate schedule evaluated on the device from the number
// of completed steps k (*step): lr = lr_min + (lr_max - lr_min) (1 + cos(pi k / total)) / 2,
// bias corrections for Adam step t0 + k + 1 (t0: steps taken before this schedule).  Nothing changes on the host between steps, so a
// training step can be captured once in a graph and replayed.
__global__ void __launch_bounds__(256) adam_scheduled_kernel(float *w, float *m, float *v, const float *g,
                                                             int64_t n, float b1, float b2, float eps,
                                                             float wd, float lr_max, float lr_min,
                                                             int total, int t0, const int *step, __bf16 *wb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = *step;
  const double c = cos(3.141592653589793 * (double)k / (double)(total > 0 ? total : 1));
  const float lr = (float)((double)lr_min + 0.5 * ((double)lr_max - (double)lr_min) * (1.0 + c));
  const float bc1 = 1.0f - powf(b1, (float)(t0 + k + 1));
  const float bc2 = 1.0f - powf(b2, (float)(t0 + k + 1));
  const float gi = g[i] + wd * w[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (__bf16)wi;
}

__global__ void step_increment_kernel(int *step) { *step += 1; }

__global__ void __launch_bounds__(256) cast_bf16_kernel(const float *x, int64_t n, __bf16 *y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (__bf16)x[i];
}

static inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

template <bool A_KC, bool B_KC>
static void launch_gemm(const GemmArgs &g, int epi, int splits, hipStream_t st) {
  dim3 grid(g.N / BN, g.M / BM, splits);
  switch (epi) {
    case EPI_F32: gemm_bf16_kernel<A_KC, B_KC, EPI_F32><<<grid, 256, 0, st>>>(g); break;
    case EPI_BIAS_SILU: gemm_bf16_kernel<A_KC, B_KC, EPI_BIAS_SILU><<<grid, 256, 0, st>>>(g); break;
    case EPI_DSILU: gemm_bf16_kernel<A_KC, B_KC, EPI_DSILU><<<grid, 256, 0, st>>>(g); break;
    default: gemm_bf16_kernel<A_KC, B_KC, EPI_BIAS><<<grid, 256, 0, st>>>(g); break;
  }
}

}  // namespace pyftle

using namespace pyftle;

extern "C" {

// C = A B with bf16 operands and FP32 accumulation; see pyftle_hip.h.
int pyftle_mlp_gemm(const void *A, int64_t lda, int32_t a_k_contig, const void *B, int64_t ldb,
                    int32_t b_k_contig, int32_t M, int32_t N, int32_t K, int32_t splits,
                    int32_t epilogue, const float *bias, void *Z, int32_t z_rows, void *C, int64_t ldc,
                    void *stream) {
  if (!A || !B || !C) return fail("mlp_gemm: NULL operand");
  if (M <= 0 || N <= 0 || K <= 0) return fail("mlp_gemm: sizes must be positive");
  if (M % BM || N % BN || K % BK) return fail("mlp_gemm: M, N must be multiples of 128 and K of 64");
  if (splits < 1 || K % (splits * BK)) return fail("mlp_gemm: K must split into multiples of 64");
  if (splits > 1 && epilogue != EPI_F32) return fail("mlp_gemm: split-K needs the FP32 epilogue");
  if ((epilogue == EPI_BIAS_SILU && !bias) || (epilogue == EPI_DSILU && !Z))
    return fail("mlp_gemm: epilogue operand missing");
  if (epilogue == EPI_DSILU && (z_rows < BM || z_rows % BM || M % z_rows))
    return fail("mlp_gemm: z_rows must be a multiple of 128 dividing M");
  if (ldc < N) return fail("mlp_gemm: ldc < N");
  if ((a_k_contig ? lda < K : lda < M) || (b_k_contig ? ldb < K : ldb < N) || lda % 8 || ldb % 8)
    return fail("mlp_gemm: leading dimensions too small or not multiples of 8");
  GemmArgs g{};
  g.A = (const __bf16 *)A;
  g.B = (const __bf16 *)B;
  g.lda = lda;
  g.ldb = ldb;
  g.M = M;
  g.N = N;
  g.K = K;
  g.k_per_split = K / splits;
  g.C32 = epilogue == EPI_F32 ? (float *)C : nullptr;
  g.Cb = epilogue == EPI_F32 ? nullptr : (__bf16 *)C;
  g.Z = (__bf16 *)Z;
  g.z_rows = z_rows > 0 ? z_rows : M;
  g.bias = bias;
  g.ldc = ldc;
  hipStream_t st = as_stream(stream);
  if (a_k_contig && b_k_contig) launch_gemm<true, true>(g, epilogue, splits, st);
  else if (a_k_contig && !b_k_contig) launch_gemm<true, false>(g, epilogue, splits, st);
  else if (!a_k_contig && b_k_contig) launch_gemm<false, true>(g, epilogue, splits, st);
  else launch_gemm<false, false>(g, epilogue, splits, st);
  return check_launch("mlp_gemm");
}

int pyftle_mlp_fourier(const float *in, int32_t d, const float *scale, const float *Bm, int32_t F,
                       int32_t M, int32_t ld_out, int32_t n_tangents, void *out, void *stream) {
  if (!in || !scale || !Bm || !out) return fail("mlp_fourier: NULL pointer");
  if (d < 1 || F < 1 || M < 0 || n_tangents < 0 || n_tangents > d || ld_out < 2 * F)
    return fail("mlp_fourier: bad sizes");
  if (M == 0) return PYFTLE_OK;
  fourier_kernel<<<nb256((int64_t)M * F), 256, 0, as_stream(stream)>>>(in, d, scale, Bm, F, M, ld_out,
                                                                       n_tangents, (__bf16 *)out);
  return check_launch("mlp_fourier");
}

int pyftle_mlp_linear_out(const void *H, int64_t ldh, int32_t K, const float *W, const float *b,
                          int32_t O, int32_t M, int32_t rows_with_bias, float *Y, void *stream) {
  if (!H || !W || !Y) return fail("mlp_linear_out: NULL pointer");
  if (K < 1 || O < 1 || O > 4 || O * K > 4096 || K % 8 || ldh % 8 || M < 0 || ldh < K)
    return fail("mlp_linear_out: needs O <= 4, O*K <= 4096, K and ldh multiples of 8");
  if (M == 0) return PYFTLE_OK;
  linear_out_kernel<<<(unsigned)((M + 255) / 256), 256, 0, as_stream(stream)>>>(
      (const __bf16 *)H, ldh, K, W, b, O, M, rows_with_bias, Y);
  return check_launch("mlp_linear_out");
}

int pyftle_mlp_linear_out_backward(const float *dY, const float *W, int32_t O, int32_t K, int32_t M,
                                   const void *H, const void *Z, int64_t ld, int32_t chunk_rows,
                                   float *dW_part, float *db_part, void *dZ, void *stream) {
  if (!dY || !W || !H || !Z || !dW_part || !db_part || !dZ)
    return fail("mlp_linear_out_backward: NULL pointer");
  if (O < 1 || K < 1 || M < 1 || ld < K || chunk_rows < 1) return fail("mlp_linear_out_backward: bad sizes");
  hipStream_t st = as_stream(stream);
  const int chunks = (M + chunk_rows - 1) / chunk_rows;
  linear_out_bwd_kernel<<<nb256((int64_t)M * K), 256, 0, st>>>(dY, W, O, K, M, (const __bf16 *)Z, ld,
                                                               (__bf16 *)dZ);
  linear_out_wgrad_kernel<<<dim3((K + 255) / 256, O, chunks), 256, 0, st>>>(
      dY, (const __bf16 *)H, ld, K, O, M, chunk_rows, dW_part, db_part);
  return check_launch("mlp_linear_out_backward");
}

int pyftle_mlp_reduce_splits(const float *part, int32_t splits, int64_t n, float *out, void *stream) {
  if (!part || !out || splits < 1 || n < 0) return fail("mlp_reduce_splits: bad arguments");
  if (n == 0) return PYFTLE_OK;
  if (splits >= 32 && n * splits <= (int64_t)1 << 26)
    reduce_splits_wide_kernel<<<(unsigned)((n + 3) / 4), 256, 0, as_stream(stream)>>>(part, splits, n, out);
  else
    reduce_splits_kernel<<<nb256(n), 256, 0, as_stream(stream)>>>(part, splits, n, out);
  return check_launch("mlp_reduce_splits");
}

int pyftle_mlp_colsum(const void *X, int64_t ld, int32_t M, int32_t N, int32_t chunk_rows, float *part,
                      void *stream) {
  if (!X || !part || M < 1 || N < 1 || ld < N || chunk_rows < 1) return fail("mlp_colsum: bad arguments");
  const int chunks = (M + chunk_rows - 1) / chunk_rows;
  colsum_kernel<<<dim3((N + 255) / 256, chunks), 256, 0, as_stream(stream)>>>((const __bf16 *)X, ld, M, N,
                                                                               chunk_rows, part);
  return check_launch("mlp_colsum");
}

int pyftle_mlp_mse(const float *Y, const float *T, int64_t n, float scale, float *dY, float *loss,
                   void *stream) {
  if (!Y || !T || !dY || n < 0) return fail("mlp_mse: bad arguments");
  if (n == 0) return PYFTLE_OK;
  mse_kernel<<<nb256(n), 256, 0, as_stream(stream)>>>(Y, T, n, scale, dY, loss);
  return check_launch("mlp_mse");
}

int pyftle_mlp_adam(float *w, float *m, float *v, const float *g, int64_t n, float lr, float beta1,
                    float beta2, float eps, int32_t step, float weight_decay, void *w_bf16,
                    void *stream) {
  if (!w || !m || !v || !g || n < 0 || step < 1) return fail("mlp_adam: bad arguments");
  if (n == 0) return PYFTLE_OK;
  const float bc1 = 1.0f - powf(beta1, (float)step);
  const float bc2 = 1.0f - powf(beta2, (float)step);
  adam_kernel<<<nb256(n), 256, 0, as_stream(stream)>>>(w, m, v, g, n, lr, beta1, beta2, eps, bc1, bc2,
                                                       weight_decay, (__bf16 *)w_bf16);
  return check_launch("mlp_adam");
}

int pyftle_mlp_adam_scheduled(float *w, float *m, float *v, const float *g, int64_t n, float beta1, float beta2,
                              float eps, float weight_decay, float lr_max, float lr_min, int32_t total_steps,
                              int32_t adam_step0, int32_t *step, void *w_bf16, void *stream) {
  if (!w || !m || !v || !g || !step || n < 0 || total_steps < 1 || adam_step0 < 0)
    return fail("mlp_adam_scheduled: bad arguments");
  hipStream_t st = as_stream(stream);
  if (n > 0)
    adam_scheduled_kernel<<<nb256(n), 256, 0, st>>>(w, m, v, g, n, beta1, beta2, eps, weight_decay, lr_max,
                                                    lr_min, total_steps, adam_step0, step, (__bf16 *)w_bf16);
  step_increment_kernel<<<1, 1, 0, st>>>(step);
  return check_launch("mlp_adam_scheduled");
}

int pyftle_mlp_cast_bf16(const float *x, int64_t n, void *y, void *stream) {
  if (!x || !y || n < 0) return fail("mlp_cast_bf16: bad arguments");
  if (n == 0) return PYFTLE_OK;
  cast_bf16_kernel<<<nb256(n), 256, 0, as_stream(stream)>>>(x, n, (__bf16 *)y);
  return check_launch("mlp_cast_bf16");
}

}  // extern "C"
