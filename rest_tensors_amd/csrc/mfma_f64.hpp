// FP64 MFMA GEMM tile engine for gfx950 (CDNA4).
//
// One workgroup = 4 wavefronts (256 lanes) computes a BM x BN tile of
//   C_slot(m,n) = sum_{b in batch range} sum_k opA(A_b)(m,k) * opB(B_b)(k,n)
// with v_mfma_f64_16x16x4f64.  The batch range is one batch for an ordinary
// strided-batched GEMM, or a chunk of batches that is summed inside the
// workgroup ("reduce" mode: K = sum_P B_P^T B_P treats the aux index P as an
// outer reduction index; partial tiles land in a slot per chunk and a second
// kernel sums the slots in a fixed order, so results are deterministic).
//
// Data movement: each k-chunk (BK deep) of both operands is loaded from
// global memory into registers one iteration ahead, then written to LDS in a
// k-major layout S[k][r] so that the 16 lanes that feed one MFMA row read 16
// consecutive doubles.  The wave grid is 2 x 2, each wave owns a
// (BM/2) x (BN/2) sub-tile = (BM/32) x (BN/32) MFMA blocks.
//
// MFMA operand order: the B fragment is passed as src0 and the A fragment as
// src1, so each accumulator holds C^T blocks.  For the FP64 16x16x4 MFMA,
// accumulator register r of lane l holds output row 4*r + l/16 and column
// l%16 (measured on gfx950; it differs from the FP32 variants' 4*(l/16) + r),
// so here lane l owns row m = l%16 of columns n = l/16 + 4*r.  Consecutive
// lanes own consecutive rows of the column-major output: stores coalesce.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rta {

typedef double f64x4 __attribute__((ext_vector_type(4)));

struct GemmParams {
  const double* A;
  const double* B;
  double* C;
  int64_t M, N, K;
  int64_t lda, ldb;
  int64_t sA, sB;          // batch strides of A and B (elements)
  int64_t c_rs, c_cs, sC;  // C(m,n) of output slot s at s*sC + m*c_rs + n*c_cs
  double alpha, beta;
  int64_t nbatch;   // number of batches
  int64_t bchunk;   // batches summed by one workgroup (1 = plain batched)
  int64_t nchunks;  // ceil(nbatch / bchunk) = number of output slots
  int tiles_m, tiles_n;
  int64_t ntiles;   // tiles per slot (tiles_m*tiles_n, or the upper count)
  int64_t nwork;    // ntiles * nchunks
  int reduce;       // 1: raw partial sums into slot `chunk` (no alpha/beta)
  int upper;        // 1: only tiles with tm <= tn (symmetric output)
};

// Bijective block-id remap so that consecutive work items land on the same
// XCD (workgroups are dispatched round-robin over the 8 XCDs); work items of
// one batch chunk then share that XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nblk) {
  const int64_t nx = 8;
  int64_t q = nblk / nx, r = nblk % nx;
  int64_t x = bid % nx, i = bid / nx;
  int64_t base = (x < r) ? x * (q + 1) : r * (q + 1) + (x - r) * q;
  return base + i;
}

// t -> (tm, tn) with tm <= tn, enumerating columns of the upper triangle.
__device__ __forceinline__ void upper_tile(int64_t t, int& tm, int& tn) {
  int64_t j = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while (j * (j + 1) / 2 > t) --j;
  while ((j + 1) * (j + 2) / 2 <= t) ++j;
  tn = (int)j;
  tm = (int)(t - j * (j + 1) / 2);
}

// Loads an R x BK operand tile (rows r0.., k0..) into registers.
// kcontig=false: elem(r,k) = P[r + k*ld];  kcontig=true: elem(r,k) = P[k + r*ld]
template <int R, int BK, bool KCONTIG>
__device__ __forceinline__ void load_tile(const double* __restrict__ P,
                                          int64_t ld, int64_t r0, int64_t rmax,
                                          int64_t k0, int64_t kmax,
                                          double (&reg)[R * BK / 256]) {
  static_assert(KCONTIG || 256 % R == 0, "row-contiguous tiles need R | 256");
  const int t = threadIdx.x;
#pragma unroll
  for (int p = 0; p < R * BK / 256; ++p) {
    int r, k;
    if (KCONTIG) {
      k = t % BK;
      r = t / BK + p * (256 / BK);
    } else {
      r = t % R;
      k = t / R + p * (256 / R);
    }
    int64_t gr = r0 + r, gk = k0 + k;
    double v = 0.0;
    if (gr < rmax && gk < kmax)
      v = KCONTIG ? P[gk + gr * ld] : P[gr + gk * ld];
    reg[p] = v;
  }
}

template <int R, int BK, bool KCONTIG, int PAD>
__device__ __forceinline__ void store_tile(double* __restrict__ S,
                                           const double (&reg)[R * BK / 256]) {
  const int t = threadIdx.x;
#pragma unroll
  for (int p = 0; p < R * BK / 256; ++p) {
    int r, k;
    if (KCONTIG) {
      k = t % BK;
      r = t / BK + p * (256 / BK);
    } else {
      r = t % R;
      k = t / R + p * (256 / R);
    }
    S[k * (R + PAD) + r] = reg[p];
  }
}

constexpr int kGemmPad = 1;

// One work item w (tile x batch chunk) of a GEMM; every thread of the 256-wide
// workgroup must call it (it synchronizes).  As/Bs: LDS of
// BK*(BM+1) and BK*(BN+1) doubles.
template <int BM, int BN, int BK, bool TA, bool TB, int WAVES_M = 2>
__device__ __forceinline__ void dgemm_tile(const GemmParams& g, int64_t w,
                                           double* __restrict__ As,
                                           double* __restrict__ Bs) {
  constexpr int PAD = kGemmPad;
  // 4 waves as WAVES_M x (4 / WAVES_M): narrow tiles (one side 32 or 48 wide)
  // stack the waves along the wide side
  constexpr int WAVES_N = 4 / WAVES_M;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int FM = WM / 16, FN = WN / 16;
  static_assert(WAVES_M * WAVES_N == 4 && WM % 16 == 0 && WN % 16 == 0, "wave layout");
  static_assert(BK % 4 == 0 && (BM * BK) % 256 == 0 && (BN * BK) % 256 == 0,
                "tile shape");
  const int64_t chunk = w / g.ntiles;
  const int64_t tid = w % g.ntiles;
  int tm, tn;
  if (g.upper) {
    upper_tile(tid, tm, tn);
  } else {
    tm = (int)(tid % g.tiles_m);
    tn = (int)(tid / g.tiles_m);
  }
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int64_t b0 = chunk * g.bchunk;
  const int64_t b1 = (b0 + g.bchunk < g.nbatch) ? b0 + g.bchunk : g.nbatch;
  const int64_t nk = (g.K + BK - 1) / BK;
  const int64_t niter = (b1 - b0) * nk;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wave_m = wave % WAVES_M, wave_n = wave / WAVES_M;

  f64x4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  double ra[BM * BK / 256], rb[BN * BK / 256];
  auto gload = [&](int64_t it) {
    int64_t b = b0 + it / nk;
    int64_t k0 = (it % nk) * BK;
    load_tile<BM, BK, TA>(g.A + b * g.sA, g.lda, m0, g.M, k0, g.K, ra);
    load_tile<BN, BK, !TB>(g.B + b * g.sB, g.ldb, n0, g.N, k0, g.K, rb);
  };

  if (niter > 0) gload(0);
  for (int64_t it = 0; it < niter; ++it) {
    __syncthreads();
    store_tile<BM, BK, TA, PAD>(As, ra);
    store_tile<BN, BK, !TB, PAD>(Bs, rb);
    __syncthreads();
    if (it + 1 < niter) gload(it + 1);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      double a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i)
        a[i] = As[kr * (BM + PAD) + wave_m * WM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < FN; ++j)
        b[j] = Bs[kr * (BN + PAD) + wave_n * WN + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] =
              __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
  }

  // epilogue: acc[i][j][r] is C(m, n) with
  //   m = m0 + wave_m*WM + i*16 + lane%16,  n = n0 + wave_n*WN + j*16 + lane/16 + 4*r
  double* out = g.C + chunk * g.sC;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int64_t m = m0 + wave_m * WM + i * 16 + (lane & 15);
    if (m >= g.M) continue;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t n = n0 + wave_n * WN + j * 16 + (lane >> 4) + 4 * r;
        if (n >= g.N) continue;
        double* p = out + m * g.c_rs + n * g.c_cs;
        double v = acc[i][j][r];
        if (g.reduce) {
          *p = v;
        } else if (g.beta == 0.0) {
          *p = g.alpha * v;
        } else {
          *p = g.alpha * v + g.beta * (*p);
        }
      }
    }
  }
}

template <int BM, int BN, int BK, bool TA, bool TB>
__global__ __launch_bounds__(256) void dgemm_mfma_kernel(GemmParams g) {
  __shared__ double As[BK * (BM + kGemmPad)];
  __shared__ double Bs[BK * (BN + kGemmPad)];
  const int64_t w = xcd_remap(blockIdx.x, gridDim.x);
  if (w >= g.nwork) return;
  dgemm_tile<BM, BN, BK, TA, TB>(g, w, As, Bs);
}

// Persistent grouped GEMM: descs[d] is one (batched or batch-reducing) GEMM
// with descs[d].nwork work items; offs holds the exclusive prefix sums of
// nwork (offs[ndesc] = total).  Workgroups stride over the global work list.
template <int BM, int BN, int BK, bool TA, bool TB, int WAVES_M = 2>
__global__ __launch_bounds__(256) void dgemm_grouped_kernel(
    const GemmParams* __restrict__ descs, const int64_t* __restrict__ offs,
    int ndesc) {
  __shared__ double As[BK * (BM + kGemmPad)];
  __shared__ double Bs[BK * (BN + kGemmPad)];
  const int64_t total = offs[ndesc];
  for (int64_t gw = blockIdx.x; gw < total; gw += gridDim.x) {
    int lo = 0, hi = ndesc;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (offs[mid] <= gw) lo = mid;
      else hi = mid;
    }
    const GemmParams g = descs[lo];
    dgemm_tile<BM, BN, BK, TA, TB, WAVES_M>(g, gw - offs[lo], As, Bs);
    __syncthreads();  // LDS is reused by the next work item
  }
}

}  // namespace rta
