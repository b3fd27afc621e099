// FP64 MFMA "TN" GEMM engine for gfx950: both operands stored k-major.
//
//   C_slot(m,n) = sum_{b in batch range} sum_k A_b[k + m*lda] * B_b[k + n*ldb]
//
// Every GEMM of the RI hot path has this shape once the intermediates are
// stored k-major:  K = sum_P B_P^T B_P,  ao2mo  X^T_P = A_P^T C_occ  and
// B_P = (X^T_P)^T C,  reference ao2mo  T_P = A_P^T C  and  ri3mo_P = T_P^T C.
//
// Data path: operand tiles go global -> LDS with global_load_lds_dwordx4
// (no register staging), NSTAGE k-chunks of BK (16 or 32) in flight.  A tile
// row holds BK doubles = BK/2 units of 16 bytes; unit c of row r is stored at
// position c ^ swz(r) (swz = (r>>1)&7 for 8-unit rows, r&15 for 16-unit
// rows), which makes the ds_read_b128 fragment reads of 16 consecutive rows
// cover all 64 LDS banks exactly once per 256 bytes: conflict-free.  A lane reads two consecutive k of its row
// with one ds_read_b128 and feeds them to two MFMAs (the k-slot order inside
// an MFMA is immaterial as long as both operands use the same one).
//
// Requirements (checked by the host launcher): K % BK == 0, lda, ldb, sA, sB
// even, A and B 16-byte aligned.  Rows beyond M/N are clamped to the last
// valid row on load and masked on store.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mfma_f64.hpp"

namespace rta {

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glb_void_t;
typedef double f64x2 __attribute__((ext_vector_type(2)));

struct TnParams {
  const double* A;
  const double* B;
  double* C;
  int64_t M, N, K;
  int64_t lda, ldb;
  int64_t sA, sB;
  int64_t c_rs, c_cs, sC;
  double alpha, beta;
  int64_t nbatch, bchunk, nchunks;
  int tiles_m, tiles_n;
  int64_t ntiles, nwork;
  int reduce, upper;
};

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  if constexpr (N == 0) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  else if constexpr (N == 2) asm volatile("s_waitcnt vmcnt(2)" ::: "memory");
  else if constexpr (N == 4) asm volatile("s_waitcnt vmcnt(4)" ::: "memory");
  else if constexpr (N == 8) asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
  else if constexpr (N == 12) asm volatile("s_waitcnt vmcnt(12)" ::: "memory");
  else if constexpr (N == 16) asm volatile("s_waitcnt vmcnt(16)" ::: "memory");
  else if constexpr (N == 24) asm volatile("s_waitcnt vmcnt(24)" ::: "memory");
  else if constexpr (N == 32) asm volatile("s_waitcnt vmcnt(32)" ::: "memory");
  else static_assert(N == 0, "unsupported vmcnt");
}

template <int BK>
__device__ __forceinline__ int tn_swz(int r) {
  return BK == 16 ? ((r >> 1) & 7) : (r & 15);
}

// Issues the direct-to-LDS loads of one BK-deep k-chunk of an R-row operand
// tile: R*BK/2 units of 16 B in NT = R*BK/128 instructions of 64 lanes.  When
// NT is a multiple of 4 every wave issues NT/4 of them; otherwise (R = 48:
// NT = 6 for BK = 16) instruction q*4 + wave goes to wave `wave`, and the
// waves carry different counts, which only the vmcnt(0) wait of a two-stage
// pipeline allows.
template <int R, int BK>
__device__ __forceinline__ void tn_issue(const double* __restrict__ base,
                                         int64_t ld, int64_t r0, int64_t rmax,
                                         int64_t koff, double* lds_tile) {
  constexpr int UPR = BK / 2;                 // units per row
  constexpr int NT = R * UPR / 64;            // instructions per workgroup
  static_assert(NT * 64 == R * UPR, "operand rows");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  auto one = [&](int inst) {
    const int u = inst * 64 + lane;
    const int row = u / UPR, pos = u % UPR;
    const int c = pos ^ tn_swz<BK>(row);
    int64_t gr = r0 + row;
    gr = gr < rmax ? gr : rmax - 1;
    const double* src = base + gr * ld + koff + 2 * c;
    __builtin_amdgcn_global_load_lds((const glb_void_t*)src,
                                     (lds_void_t*)(lds_tile + inst * 128), 16,
                                     0, 0);
  };
  if constexpr (NT % 4 == 0) {
#pragma unroll
    for (int q = 0; q < NT / 4; ++q) one(wave * (NT / 4) + q);
  } else {
#pragma unroll
    for (int q = 0; q < (NT + 3) / 4; ++q)
      if (q * 4 + wave < NT) one(q * 4 + wave);
  }
}

// One work item w (tile x batch chunk) of a TN GEMM; all 256 threads call it.
// smem: NSTAGE (BM + BN) BK doubles.  WAVES_M of the 4 waves along M.
template <int BM, int BN, int BK, int NSTAGE, int WAVES_M = 2, bool COLSKIP = false>
__device__ __forceinline__ void tn_tile(const TnParams& g, int64_t w, double* __restrict__ smem) {
  static_assert(BK == 16 || BK == 32, "BK");
  constexpr int WAVES_N = 4 / WAVES_M;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int FM = WM / 16, FN = WN / 16;
  static_assert(WAVES_M * WAVES_N == 4 && WM % 16 == 0 && WN % 16 == 0, "wave layout");
  static_assert(((BM * BK / 2) % 256 == 0 && (BN * BK / 2) % 256 == 0) || NSTAGE == 2,
                "uneven load counts per wave need the two-stage (vmcnt 0) pipeline");
  constexpr int STAGE = (BM + BN) * BK;  // doubles per stage
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
  const int64_t nk = g.K / BK;
  const int64_t niter = (b1 - b0) * nk;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int wave_m = wave % WAVES_M, wave_n = wave / WAVES_M;

  // Upper mode: a wave tile strictly below the diagonal is never read.  Its
  // wave keeps loading and synchronising but issues no LDS reads or MFMAs,
  // which leaves its SIMD's MFMA pipe to the co-resident workgroup.
  // COLSKIP: the 16-column MFMA blocks of the wave that lie wholly at or past
  // N (a narrow N in a wide tile) get no fragment reads and no MFMAs; a wave
  // with none left idles like the one below the diagonal.
  int fn_act = FN;
  if constexpr (COLSKIP) {
    const int64_t left = g.N - (n0 + wave_n * WN);
    fn_act = left <= 0 ? 0 : (left >= WN ? FN : (int)((left + 15) / 16));
  }
  const bool idle =
      (g.upper && (m0 + wave_m * WM >= n0 + wave_n * WN + WN)) || (COLSKIP && fn_act == 0);

  f64x4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

  // next chunk to load: batch ib, k offset ik, LDS stage islot (counters
  // advanced per issue; no division in the loop)
  int64_t ib = b0, ik = 0;
  int islot = 0;
  auto issue = [&]() {
    double* st = smem + islot * STAGE;
    tn_issue<BM, BK>(g.A + ib * g.sA, g.lda, m0, g.M, ik, st);
    tn_issue<BN, BK>(g.B + ib * g.sB, g.ldb, n0, g.N, ik, st + BM * BK);
    ik += BK;
    if (ik == nk * BK) {
      ik = 0;
      ++ib;
    }
    islot = (islot + 1 == NSTAGE) ? 0 : islot + 1;
  };

  // prologue: NSTAGE-1 chunks in flight
#pragma unroll
  for (int s = 0; s < NSTAGE - 1; ++s)
    if (s < niter) issue();

  constexpr int PER_STAGE = (BM + BN) * BK / 512;  // load instructions per wave
  for (int64_t it = 0; it < niter; ++it) {
    // this wave's loads of chunk `it` have landed once at most the loads of
    // the younger chunks are outstanding
    if constexpr (NSTAGE > 2) {
      const int64_t younger = (niter - 1 - it) < (NSTAGE - 2) ? (niter - 1 - it)
                                                              : (NSTAGE - 2);
      if (younger >= 2) wait_vmcnt<2 * PER_STAGE>();
      else if (younger == 1) wait_vmcnt<PER_STAGE>();
      else wait_vmcnt<0>();
    } else {
      wait_vmcnt<0>();
    }
    // every wave's chunk `it` is in LDS and chunk it-1 is consumed (its LDS
    // reads completed before its MFMAs).  A bare s_barrier: __syncthreads()
    // would also drain vmcnt and serialize the load pipeline.
    __builtin_amdgcn_s_barrier();
    if (it + NSTAGE - 1 < niter) issue();  // chunk it + NSTAGE - 1
    if (idle) continue;

    const double* As = smem + (int)(it % NSTAGE) * STAGE;
    const double* Bs = As + BM * BK;
    // fragments of k-pair group kp (16 B per lane per 16-row block); two sets
    // in flight: the reads of group kp+1 are issued before the MFMAs of kp,
    // and a scheduling barrier keeps the compiler from hoisting every group's
    // reads (which would need 128 VGPRs and push the accumulators out)
    f64x2 a[2][FM], b[2][FN];
    auto load = [&](int kp, int set) {
      const int c = kp * 4 + (lane >> 4);
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int r = wave_m * WM + i * 16 + (lane & 15);
        a[set][i] = *(const f64x2*)(As + r * BK + 2 * (c ^ tn_swz<BK>(r)));
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        if (COLSKIP && j >= fn_act) continue;
        const int r = wave_n * WN + j * 16 + (lane & 15);
        b[set][j] = *(const f64x2*)(Bs + r * BK + 2 * (c ^ tn_swz<BK>(r)));
      }
    };
    load(0, 0);
#pragma unroll
    for (int kp = 0; kp < BK / 8; ++kp) {
      if (kp + 1 < BK / 8) load(kp + 1, (kp + 1) & 1);
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            if (COLSKIP && j >= fn_act) continue;
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[kp & 1][j][s], a[kp & 1][i][s],
                                                             acc[i][j], 0, 0, 0);
          }
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  // epilogue: lane l owns row m = .. + l%16 of columns n = .. + l/16 + 4*r
  if (idle) return;
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
        const double v = acc[i][j][r];
        if (g.reduce) *p = v;
        else if (g.beta == 0.0) *p = g.alpha * v;
        else *p = g.alpha * v + g.beta * (*p);
      }
    }
  }
}

template <int BM, int BN, int BK, int NSTAGE>
__global__ __launch_bounds__(256) void dgemm_tn_kernel(TnParams g) {
  extern __shared__ double smem[];
  const int64_t w = xcd_remap(blockIdx.x, gridDim.x);
  if (w >= g.nwork) return;
  tn_tile<BM, BN, BK, NSTAGE>(g, w, smem);
}

// Persistent grouped TN GEMM over a work list: descs[d] has descs[d].nwork
// items, offs the exclusive prefix sums (offs[ndesc] = total).
template <int BM, int BN, int BK, int NSTAGE, int WAVES_M, bool COLSKIP = false>
__global__ __launch_bounds__(256) void dgemm_grouped_tn_kernel(const TnParams* __restrict__ descs,
                                                               const int64_t* __restrict__ offs,
                                                               int ndesc) {
  extern __shared__ double smem[];
  const int64_t total = offs[ndesc];
  for (int64_t gw = blockIdx.x; gw < total; gw += gridDim.x) {
    int lo = 0, hi = ndesc;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (offs[mid] <= gw) lo = mid;
      else hi = mid;
    }
    const TnParams g = descs[lo];
    tn_tile<BM, BN, BK, NSTAGE, WAVES_M, COLSKIP>(g, gw - offs[lo], smem);
    __syncthreads();  // every wave is done with LDS before the next item's loads
  }
}

}  // namespace rta
