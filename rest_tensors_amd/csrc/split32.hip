// RI-K in split FP32 (RT_PREC_SPLIT_FP32): every B value is split into two
// FP32 numbers, hi = fl32(b) and lo = fl32(b - hi) (b = hi + lo to ~2^-48),
// and K = sum_P B_P^T B_P is formed from the three products that matter,
//   K ~ sum hi^T hi + (hi^T lo + lo^T hi)          (lo^T lo ~ 2^-48, dropped)
// on v_mfma_f32_16x16x4f32 with FP32 accumulation.  The hi*hi sums and the
// cross sums use separate FP32 accumulators (the cross terms are 2^-24 of
// the others and would vanish in a shared one), and both are folded into
// FP64 after every 32-deep k-chunk so FP32 rounding acts on 32-term sums only.
//
// On MI355X an FP32 MFMA does only twice the multiply-adds of an FP64 one per
// cycle, so three products make this form slower than the FP64 build; it is
// kept as the reference point the precision comparison names.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.hpp"
#include "gemm.hpp"
#include "mfma_f64.hpp"
#include "mfma_tn.hpp"
#include "sliced.hpp"
#include "split32.hpp"
#include "timer.hpp"

namespace rta {

typedef float s32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void s32_lds_t;
typedef __attribute__((address_space(1))) void s32_glb_t;

// One thread per (m, i, P) value: planes S[((P*nmo + i)*2 + {0,1})*noccp + m]
// (hi, lo), zeros for the padded m in [nocc, noccp).
__global__ __launch_bounds__(256) void split32_kernel(const double* __restrict__ B,
                                                      int64_t nocc, int64_t noccp,
                                                      int64_t ncol, float* __restrict__ S) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= ncol * noccp) return;
  const int64_t col = e / noccp, m = e % noccp;
  const double v = (m < nocc) ? B[col * nocc + m] : 0.0;
  const float hi = (float)v;
  const float lo = (float)(v - (double)hi);
  S[(col * 2 + 0) * noccp + m] = hi;
  S[(col * 2 + 1) * noccp + m] = lo;
}

struct Split32Params {
  const float* S;
  double* part;  // partial slots, slot stride nmo*nmo
  int64_t nmo, noccp, naux, bchunk;
  int64_t ntiles, nwork;
};

// 512 threads = 8 waves (2 x 4), 128 x 128 upper tile, wave tile 64 x 32
// (4 x 2 blocks of 16 x 16).  One stage = one 32-deep k-chunk of one aux slice,
// both planes of both operands: a row holds 16 units of 16 bytes (unit
// plane*8 + q = floats 4q..4q+3 of the plane), stored at unit ^ (row & 15).
// Lane group g = lane/16 reads quads g and 4+g; element e of quad j feeds the
// MFMA k-step (j, e), so one b128 read serves 4 MFMAs and the lanes of the 4
// groups together cover the 32 k of the chunk (the same assignment for A and
// B, so the k order does not matter).  B fragment = src0, A fragment = src1:
// register r of lane l holds K(row = .. + l%16, col = .. + 4*(l/16) + r).
__global__ __launch_bounds__(512) void rik_split32_kernel(Split32Params g) {
  constexpr int ROWS = 128, UPR = 16, OPU = ROWS * UPR, STAGE = 2 * OPU;
  constexpr int FM = 4, FN = 2;
  constexpr int PER_STAGE = STAGE / 512;  // direct-to-LDS loads per thread per stage
  extern __shared__ uint4 s32_lds[];

  const int64_t w = xcd_remap(blockIdx.x, gridDim.x);
  if (w >= g.nwork) return;
  const int64_t chunk = w / g.ntiles;
  int tm, tn;
  upper_tile(w % g.ntiles, tm, tn);
  const int64_t m0 = (int64_t)tm * ROWS, n0 = (int64_t)tn * ROWS;
  const int64_t b0 = chunk * g.bchunk;
  const int64_t b1 = (b0 + g.bchunk < g.naux) ? b0 + g.bchunk : g.naux;
  const int64_t nkc = g.noccp / 32;
  const int64_t niter = (b1 - b0) * nkc;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wave_m = wave & 1, wave_n = wave >> 1;
  const int l15 = lane & 15, grp = lane >> 4;

  s32x4_t hh[FM][FN], hx[FM][FN];
  double acc[FM][FN][4];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      hh[i][j] = s32x4_t{0.f, 0.f, 0.f, 0.f};
      hx[i][j] = s32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
    }

  auto issue = [&](int64_t it) {
    const int64_t P = b0 + it / nkc;
    const int64_t kc = it % nkc;
    uint4* st = s32_lds + (it & 1) * STAGE;
#pragma unroll
    for (int op = 0; op < 2; ++op) {
      const int64_t base = op == 0 ? m0 : n0;
#pragma unroll
      for (int q = 0; q < OPU / 512; ++q) {
        const int inst = wave * (OPU / 512) + q;
        const int u = inst * 64 + lane;
        const int row = u >> 4, pos = u & 15;
        const int gu = pos ^ (row & 15);
        const int plane = gu >> 3, quad = gu & 7;
        int64_t i = base + row;
        i = i < g.nmo ? i : g.nmo - 1;
        const float* src = g.S + ((P * g.nmo + i) * 2 + plane) * g.noccp + kc * 32 + quad * 4;
        __builtin_amdgcn_global_load_lds((const s32_glb_t*)src,
                                         (s32_lds_t*)(st + op * OPU + inst * 64), 16, 0, 0);
      }
    }
  };

  if (niter > 0) issue(0);
  for (int64_t it = 0; it < niter; ++it) {
    if (it + 1 < niter) {
      issue(it + 1);  // buffer (it+1)&1 was released by the barrier ending step it-1
      wait_vmcnt<PER_STAGE>();
    } else {
      wait_vmcnt<0>();
    }
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
    const uint4* As = s32_lds + (it & 1) * STAGE;
    const uint4* Bs = As + OPU;
#pragma unroll
    for (int jq = 0; jq < 2; ++jq) {
      const int quad = jq * 4 + grp;
      s32x4_t ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int r = wave_m * 64 + i * 16 + l15;
        ah[i] = __builtin_bit_cast(s32x4_t, As[r * UPR + ((0 * 8 + quad) ^ l15)]);
        al[i] = __builtin_bit_cast(s32x4_t, As[r * UPR + ((1 * 8 + quad) ^ l15)]);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int r = wave_n * 32 + j * 16 + l15;
        bh[j] = __builtin_bit_cast(s32x4_t, Bs[r * UPR + ((0 * 8 + quad) ^ l15)]);
        bl[j] = __builtin_bit_cast(s32x4_t, Bs[r * UPR + ((1 * 8 + quad) ^ l15)]);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            hh[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bh[j][e], ah[i][e], hh[i][j], 0, 0, 0);
            hx[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bl[j][e], ah[i][e], hx[i][j], 0, 0, 0);
            hx[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bh[j][e], al[i][e], hx[i][j], 0, 0, 0);
          }
    }
    // fold the chunk's FP32 sums into FP64
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[i][j][r] += (double)hh[i][j][r] + (double)hx[i][j][r];
        hh[i][j] = s32x4_t{0.f, 0.f, 0.f, 0.f};
        hx[i][j] = s32x4_t{0.f, 0.f, 0.f, 0.f};
      }
    __builtin_amdgcn_s_barrier();  // this stage's buffer may be refilled
  }

  double* out = g.part + chunk * g.nmo * g.nmo;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int64_t row = m0 + wave_m * 64 + i * 16 + l15;
    if (row >= g.nmo) continue;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t col = n0 + wave_n * 32 + j * 16 + 4 * grp + r;
        if (col < g.nmo) out[row + col * g.nmo] = acc[i][j][r];
      }
  }
}

int rik_split32(const double* B, int64_t nocc, int64_t nmo, int64_t naux, double* K,
                hipStream_t s) {
  const int64_t noccp = std::max<int64_t>(32, (nocc + 31) / 32 * 32);
  float* S = nullptr;
  RTA_TRY(scratch_get(SCRATCH_LOWP_A,
                      sizeof(float) * 2 * (size_t)noccp * nmo * std::max<int64_t>(naux, 1),
                      (void**)&S));
  const int64_t ncol = nmo * naux;
  if (ncol > 0) {
    KernelTimer kt("rik_slice", s);
    hipLaunchKernelGGL(split32_kernel, dim3((unsigned)((ncol * noccp + 255) / 256)),
                       dim3(256), 0, s, B, nocc, noccp, ncol, S);
    RTA_HIP(hipGetLastError());
  }
  const int64_t bchunk = std::max<int64_t>(1, reduce_bchunk(nmo, 3 * noccp / 2, naux, 128));
  const int64_t nchunks = naux > 0 ? (naux + bchunk - 1) / bchunk : 1;
  double* part = nullptr;
  RTA_TRY(scratch_get(SCRATCH_GEMM_PARTIAL, sizeof(double) * nchunks * nmo * nmo,
                      (void**)&part));
  Split32Params p;
  p.S = S;
  p.part = part;
  p.nmo = nmo;
  p.noccp = (nocc > 0 && naux > 0) ? noccp : 0;  // no k-chunks: zero partials
  p.naux = naux;
  p.bchunk = bchunk;
  const int64_t tiles = (nmo + 127) / 128;
  p.ntiles = tiles * (tiles + 1) / 2;
  p.nwork = p.ntiles * nchunks;
  const size_t lds = 2 * 2 * 128 * 16 * 16;
  static bool attr = false;
  if (!attr) {
    RTA_HIP(hipFuncSetAttribute((const void*)rik_split32_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  {
    KernelTimer kt("rik_split32", s);
    hipLaunchKernelGGL(rik_split32_kernel, dim3((unsigned)p.nwork), dim3(512), lds, s, p);
    RTA_HIP(hipGetLastError());
  }
  return reduce_partials_sym(part, nchunks, nmo, K, false, s);
}

}  // namespace rta
