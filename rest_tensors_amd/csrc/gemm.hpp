// Host-side launch helpers around the FP64 MFMA tile engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mfma_f64.hpp"

namespace rta {

// Column-major operand description for one GEMM (possibly batched).
struct GemmDesc {
  bool ta = false, tb = false;
  int64_t M = 0, N = 0, K = 0;
  const double* A = nullptr;
  int64_t lda = 0, sA = 0;
  const double* B = nullptr;
  int64_t ldb = 0, sB = 0;
  double* C = nullptr;
  int64_t c_rs = 1, c_cs = 0, sC = 0;  // C(m,n) of batch b at b*sC+m*c_rs+n*c_cs
  double alpha = 1.0, beta = 0.0;
  int64_t batch = 1;
};

// Tile edge for the engine: 128 (big problems) or 64; 0 picks by size.
int gemm_pick_tile(int64_t M, int64_t N, int64_t batch);

// Strided-batched GEMM: C_b = alpha*op(A_b)*op(B_b) + beta*C_b for all b.
int gemm_batched(const GemmDesc& d, hipStream_t stream, const char* timer_name,
                 int tile = 0);

// Number of compute units of the current device (cached).
int device_cu_count();

// Whether a descriptor can run on the k-major (TN) direct-to-LDS engine:
// ta && !tb, K % 16 == 0, even leading dims and batch strides, 16-byte
// aligned A and B.
bool tn_eligible(const GemmDesc& d);

// Batches per workgroup for gemm_reduce_sym on an M x M output with inner
// dimension K per batch (balances the grid over the CUs).
int64_t reduce_bchunk(int64_t M, int64_t K, int64_t nbatch, int tile);

// Batch-reducing GEMM with symmetric output:
//   out(m,n) = alpha * sum_b op(A_b)(m,:) . op(B_b)(:,n), M == N, only the
// upper tiles are computed, then summed over chunks of `bchunk` batches in a
// fixed order and mirrored.  out is M x M column-major with leading dim ldo.
// `partial` must hold nchunks*M*M doubles (nchunks = ceil(batch/bchunk)).
int gemm_reduce_sym(const GemmDesc& d, int64_t bchunk, double* partial,
                    double* out, int64_t ldo, hipStream_t stream,
                    const char* timer_name, int tile = 128);

}  // namespace rta
