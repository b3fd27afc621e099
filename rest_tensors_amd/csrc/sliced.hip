// Sliced-precision K build (RT_PREC_OZAKI_BF16, RT_PREC_BF16X3); the
// arithmetic is described in sliced.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "gemm.hpp"
#include "mfma_f64.hpp"
#include "sliced.hpp"
#include "timer.hpp"

namespace rta {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_v_t;
typedef __attribute__((address_space(1))) void glb_v_t;

__device__ __forceinline__ uint16_t f32_to_bf16_exact(float f) {
  // callers pass values with <= 8 significant bits: truncation is exact
  return (uint16_t)(__float_as_uint(f) >> 16);
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

__device__ __forceinline__ double slc_wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// One wave per column (i, P): slices of B[:, i, P] (nocc values, padded with
// zeros to noccp).
template <int NS, bool ALIGNED>
__global__ __launch_bounds__(256) void slice_kernel(const double* __restrict__ B,
                                                    int64_t nocc, int64_t noccp,
                                                    int64_t ncol,
                                                    uint16_t* __restrict__ S,
                                                    int32_t* __restrict__ E) {
  const int lane = threadIdx.x & 63;
  const int64_t col = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (col >= ncol) return;
  const double* v = B + col * nocc;
  uint16_t* out = S + col * NS * noccp;
  int e = 0;
  if (ALIGNED) {
    double mx = 0.0;
    for (int64_t m = lane; m < nocc; m += 64) mx = fmax(mx, fabs(v[m]));
    mx = slc_wave_max(mx);
    if (mx > 0.0) frexp(mx, &e);  // mx < 2^e
    if (lane == 0) E[col] = e;
  } else if (lane == 0) {
    E[col] = 0;
  }
  for (int64_t m = lane; m < noccp; m += 64) {
    const double x = (m < nocc) ? v[m] : 0.0;
    if (ALIGNED) {
      double u = ldexp(x, 7 - e);  // |u| <= 128
      double sg = rint(u);
      out[m] = f32_to_bf16_exact((float)sg);
      double rem = u - sg;  // |rem| <= 1/2, exact
#pragma unroll
      for (int s = 1; s < NS; ++s) {
        u = rem * 256.0;
        sg = rint(u);
        out[s * noccp + m] = f32_to_bf16_exact((float)sg);
        rem = u - sg;
      }
    } else {
      const uint16_t h = f32_to_bf16_rne((float)x);
      const double rest = x - (double)bf16_to_f32(h);
      out[m] = h;
      if (NS > 1) out[noccp + m] = f32_to_bf16_rne((float)rest);
    }
  }
}

template <int NS>
__device__ __forceinline__ int slc_swz(int r) {
  return NS == 4 ? (r & 15) : ((r >> 1) & 7);
}

struct SlicedParams {
  const uint16_t* S;
  const int32_t* E;
  double* part;  // partial slots, slot stride nmo*nmo
  int64_t nmo, noccp, naux;
  int64_t bchunk;
  int tiles;     // tiles per side
  int64_t ntiles, nwork;
};

// 128 x 128 upper tile of K per workgroup (8 waves as 2 x 4, wave tile 64 x 32),
// stage = one 32-deep k-chunk of one aux slice, all NS slices of both operands.
// The B fragment is MFMA src0 and the A fragment src1, so accumulator register
// r of lane l holds K(row = .. + l%16, col = .. + 4*(l/16) + r) (the FP32
// accumulator layout of the 16x16 MFMAs): consecutive lanes own consecutive
// rows and the column-major stores coalesce.
template <int NS, bool ALIGNED>
__global__ __launch_bounds__(512) void rik_sliced_kernel(SlicedParams g) {
  constexpr int UPR = 4 * NS;          // 16-byte units per tile row per stage
  constexpr int ROWS = 128;
  constexpr int STAGE_UNITS = 2 * ROWS * UPR;
  constexpr int FM = 4, FN = 2;
  constexpr int NL = ALIGNED ? NS : 2;  // product levels kept
  extern __shared__ uint4 lds[];       // 2 stages

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

  f32x4_t acc[NL][FM][FN];
  double acc64[FM][FN][4];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int q = 0; q < NL; ++q) acc[q][i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[i][j][r] = 0.0;
    }

  auto issue = [&](int64_t it) {
    const int64_t P = b0 + it / nkc;
    const int64_t kc = it % nkc;
    uint4* st = lds + (it & 1) * STAGE_UNITS;
#pragma unroll
    for (int op = 0; op < 2; ++op) {
      const int64_t base = op == 0 ? m0 : n0;
#pragma unroll
      for (int q = 0; q < ROWS * UPR / 512; ++q) {
        const int inst = wave * (ROWS * UPR / 512) + q;
        const int u = inst * 64 + lane;
        const int row = u / UPR, pos = u % UPR;
        const int gu = pos ^ slc_swz<NS>(row);
        const int s = gu >> 2, c = gu & 3;
        int64_t i = base + row;
        i = i < g.nmo ? i : g.nmo - 1;
        const uint16_t* src =
            g.S + ((P * g.nmo + i) * NS + s) * g.noccp + kc * 32 + c * 8;
        __builtin_amdgcn_global_load_lds((const glb_v_t*)src,
                                         (lds_v_t*)(st + op * ROWS * UPR + inst * 64),
                                         16, 0, 0);
      }
    }
  };

  constexpr int PER_STAGE = 2 * ROWS * UPR / 512;  // load instructions per wave
  if (niter > 0) issue(0);
  for (int64_t it = 0; it < niter; ++it) {
    if (it + 1 < niter) {
      // the previous stage's buffer is free once every wave passed the barrier
      // that ended its compute (below)
      issue(it + 1);
      if constexpr (PER_STAGE == 8) asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
      else asm volatile("s_waitcnt vmcnt(4)" ::: "memory");
    } else {
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    __builtin_amdgcn_s_barrier();
    const uint4* As = lds + (it & 1) * STAGE_UNITS;
    const uint4* Bs = As + ROWS * UPR;
    const int kg = lane >> 4;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      bf16x8_t a[FM];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int r = wave_m * 64 + i * 16 + (lane & 15);
        uint4 v = As[r * UPR + ((s * 4 + kg) ^ slc_swz<NS>(r))];
        a[i] = __builtin_bit_cast(bf16x8_t, v);
      }
#pragma unroll
      for (int t = 0; t < NS; ++t) {
        const int q = ALIGNED ? s + t : (s + t == 0 ? 0 : 1);
        if (ALIGNED ? (s + t >= NS) : (s + t >= 2)) continue;
        bf16x8_t b[FN];
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const int r = wave_n * 32 + j * 16 + (lane & 15);
          uint4 v = Bs[r * UPR + ((t * 4 + kg) ^ slc_swz<NS>(r))];
          b[j] = __builtin_bit_cast(bf16x8_t, v);
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[q][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j], a[i],
                                                                   acc[q][i][j], 0, 0, 0);
      }
    }
    // fold into FP64 after the aux slice's last k-chunk, and every 8 chunks
    // (256 k) so that the FP32 level sums stay exact integers below 2^24
    const int64_t kc_now = it % nkc;
    if (kc_now == nkc - 1 || (kc_now & 7) == 7) {
      const int64_t P = b0 + it / nkc;
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int64_t row = m0 + wave_m * 64 + i * 16 + (lane & 15);
        const int ei = ALIGNED ? g.E[P * g.nmo + (row < g.nmo ? row : g.nmo - 1)] : 0;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t col = n0 + wave_n * 32 + j * 16 + 4 * (lane >> 4) + r;
            double v;
            if (ALIGNED) {
              v = (double)acc[0][i][j][r];
#pragma unroll
              for (int q = 1; q < NL; ++q) v += ldexp((double)acc[q][i][j][r], -8 * q);
              const int ej = g.E[P * g.nmo + (col < g.nmo ? col : g.nmo - 1)];
              v = ldexp(v, ei + ej - 14);
            } else {
              v = (double)acc[0][i][j][r] + (double)acc[1][i][j][r];
            }
            acc64[i][j][r] += v;
          }
#pragma unroll
          for (int q = 0; q < NL; ++q) acc[q][i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        }
      }
    }
    __builtin_amdgcn_s_barrier();  // this stage's buffer may be refilled
  }

  double* out = g.part + chunk * g.nmo * g.nmo;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int64_t row = m0 + wave_m * 64 + i * 16 + (lane & 15);
    if (row >= g.nmo) continue;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t col = n0 + wave_n * 32 + j * 16 + 4 * (lane >> 4) + r;
        if (col < g.nmo) out[row + col * g.nmo] = acc64[i][j][r];
      }
  }
}

int slice_b(const double* B, int64_t nocc, int64_t nmo, int64_t naux, int ns,
            bool aligned, uint16_t* S, int32_t* E, hipStream_t s) {
  const int64_t noccp = (nocc + 31) / 32 * 32;
  const int64_t ncol = nmo * naux;
  const unsigned grid = (unsigned)((ncol + 3) / 4);
  KernelTimer kt("rik_slice", s);
  if (aligned && ns == 4)
    hipLaunchKernelGGL((slice_kernel<4, true>), dim3(grid), dim3(256), 0, s, B, nocc,
                       noccp, ncol, S, E);
  else if (!aligned && ns == 2)
    hipLaunchKernelGGL((slice_kernel<2, false>), dim3(grid), dim3(256), 0, s, B, nocc,
                       noccp, ncol, S, E);
  else
    return fail(RT_ERR_ARG, "unsupported slicing (%d slices, aligned %d)", ns,
                (int)aligned);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

template <int NS, bool AL>
static int launch_sliced(const SlicedParams& p, hipStream_t s) {
  const size_t lds = 2 * (size_t)2 * 128 * 4 * NS * 16;
  static bool attr = false;
  if (!attr) {
    RTA_HIP(hipFuncSetAttribute((const void*)rik_sliced_kernel<NS, AL>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  hipLaunchKernelGGL((rik_sliced_kernel<NS, AL>), dim3((unsigned)p.nwork), dim3(512),
                     lds, s, p);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// out (upper, mirrored) = [out +] sum of the slots' upper triangles
__global__ __launch_bounds__(256) void reduce_sym_kernel2(const double* __restrict__ part,
                                                          int64_t nslot, int64_t M,
                                                          double* __restrict__ out,
                                                          int accumulate) {
  __shared__ double T[32][33];
  int tm, tn;
  upper_tile(blockIdx.x, tm, tn);
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t slot = M * M;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int64_t m = (int64_t)tm * 32 + tx, n = (int64_t)tn * 32 + ty + 8 * q;
    double s = 0.0;
    if (m < M && n < M && m <= n) {
      const double* p = part + m + n * M;
      for (int64_t k = 0; k < nslot; ++k) s += p[k * slot];
      if (accumulate) s += out[m + n * M];
      out[m + n * M] = s;
    }
    T[ty + 8 * q][tx] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int64_t n = (int64_t)tn * 32 + tx, m = (int64_t)tm * 32 + ty + 8 * q;
    if (m < M && n < M && m < n) out[n + m * M] = T[tx][ty + 8 * q];
  }
}

int reduce_partials_sym(const double* part, int64_t nslot, int64_t M, double* K,
                        bool accumulate, hipStream_t s) {
  const int64_t t32 = (M + 31) / 32;
  KernelTimer kt("reduce_sym", s);
  hipLaunchKernelGGL(reduce_sym_kernel2, dim3((unsigned)(t32 * (t32 + 1) / 2)), dim3(256),
                     0, s, part, nslot, M, K, accumulate ? 1 : 0);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

int rik_sliced(const uint16_t* S, const int32_t* E, int64_t nocc, int64_t nmo,
               int64_t naux, int ns, bool aligned, double* K, hipStream_t s) {
  const int64_t noccp = (nocc + 31) / 32 * 32;
  const int64_t bchunk = reduce_bchunk(nmo, noccp, naux, 128);
  const int64_t nchunks = (naux + bchunk - 1) / bchunk;
  double* part = nullptr;
  RTA_TRY(scratch_get(SCRATCH_GEMM_PARTIAL, sizeof(double) * nchunks * nmo * nmo,
                      (void**)&part));
  SlicedParams p;
  p.S = S;
  p.E = E;
  p.part = part;
  p.nmo = nmo;
  p.noccp = noccp;
  p.naux = naux;
  p.bchunk = bchunk;
  p.tiles = (int)((nmo + 127) / 128);
  p.ntiles = (int64_t)p.tiles * (p.tiles + 1) / 2;
  p.nwork = p.ntiles * nchunks;
  {
    KernelTimer kt(aligned ? "rik_ozaki" : "rik_bf16x3", s);
    if (aligned && ns == 4) RTA_TRY((launch_sliced<4, true>(p, s)));
    else if (!aligned && ns == 2) RTA_TRY((launch_sliced<2, false>(p, s)));
    else return fail(RT_ERR_ARG, "unsupported slicing");
  }
  const int64_t t32 = (nmo + 31) / 32;
  KernelTimer kt("reduce_sym", s);
  hipLaunchKernelGGL(reduce_sym_kernel2, dim3((unsigned)(t32 * (t32 + 1) / 2)), dim3(256),
                     0, s, part, nchunks, nmo, K, 0);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

}  // namespace rta
