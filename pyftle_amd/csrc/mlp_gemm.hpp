// The BF16 MFMA GEMM of mlp.hip as seen by the other MLP sources (grouped launches of
// the fused training step).
#pragma once

#include <hip/hip_bf16.h>

#include "common.hpp"

namespace pyftle {

struct GemmArgs {
  const __bf16 *A;
  const __bf16 *B;
  int64_t lda, ldb;   // leading dimension of the stored matrices
  int M, N, K;
  int k_per_split;    // K range of one split (grid z)
  float *C32;         // EPI_F32: (split, M, N) partials / result
  __bf16 *Cb;         // bf16 output
  __bf16 *Z;          // EPI_BIAS_SILU: pre-activation out; EPI_DSILU: pre-activation in
  int z_rows;         // EPI_DSILU: Z row = row % z_rows (tangent blocks stacked under primal rows)
  const float *bias;  // (N,)
  int64_t ldc;
};

// One launch over several FP32 (split-K partial) GEMMs with k-contiguous A and B (A
// stored M x K, B stored N x K): every work-group takes one 128 x 128 tile of one split of
// one member.  Members are checked by the caller as pyftle_mlp_gemm checks them.
constexpr int kGemmGroupMax = 8;
int launch_gemm_group_f32(const GemmArgs *members, int n, hipStream_t st);

}  // namespace pyftle
