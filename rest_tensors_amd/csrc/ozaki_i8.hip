// RI-K on INT8 MFMA: Ozaki-scheme error-free slicing of B into int8 planes,
// exact int32 products on v_mfma_i32_16x16x64_i8, FP64 folding.
//
// Slicing of row i of one exponent group g (the aux slices
// P in [g*gp, (g+1)*gp), all m < nocc):
//   e = exponent with max|B[m,i,P]| < 2^e over the group, u = B 2^-e in (-1,1);
//   x0 = 127 u,        sigma_0 = rint(x0), r_0 = x0 - sigma_0  (|r_0| <= 1/2)
//   x_s = 254 r_{s-1}, sigma_s = rint(x_s), r_s = x_s - sigma_s (s = 1..3)
//   so  B = 2^e / 127 * sum_s sigma_s 254^-s  up to 2^e/127 * 254^-3 / 2,
//   every |sigma_s| <= 127 (exact int8).
// K[i,j] over the group = 2^(e_i+e_j) / 127^2 * sum_q 254^-q L_q with the level
// sums L_q = sum_{s+t=q} sum_{m,P} sigma_s[m,i,P] sigma_t[m,j,P], q <= 3
// (10 slice products; the dropped levels q >= 4 weigh 254^-4).  Each L_q is an
// exact integer: at most 4 products of |sigma| <= 127 over gp*noccp <= 8192
// terms, below 2^29, so the int32 accumulators never overflow; the group's
// partial is folded into FP64 once per group.
//
// Work decomposition as in the FP64 reduce-mode GEMM: one workgroup per
// (128x128 upper tile, exponent group), one partial slot per group, slots
// summed in a fixed order by reduce_partials_sym (deterministic, symmetric).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "gemm.hpp"
#include "mfma_f64.hpp"
#include "mfma_tn.hpp"
#include "ozaki_i8.hpp"
#include "sliced.hpp"
#include "timer.hpp"

namespace rta {

typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void oz_lds_t;
typedef __attribute__((address_space(1))) void oz_glb_t;

__device__ __forceinline__ double oz_wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// One workgroup per (row i, group g): every thread holds up to 16 pairs of
// values (m, m+1) of the group in registers (gp * noccp <= 8192 values), the
// group max comes from a workgroup reduction, then each pair is cut into its
// four slices and stored as 2-byte words (zeros for the padded m >= nocc).
// Byte layout of S: ((P*nmo + i)*nkc + kc)*256 + s*64 + (m % 64), kc = m / 64.
template <bool VEC>
__global__ __launch_bounds__(256) void oz_slice_kernel(
    const double* __restrict__ B, int64_t nocc, int64_t nkc, int64_t nmo,
    int64_t naux, int64_t gp, int8_t* __restrict__ S, int32_t* __restrict__ E) {
  __shared__ double wmax[4];
  const int64_t i = blockIdx.x % nmo, g = blockIdx.x / nmo;
  const int64_t p0 = g * gp, p1 = (p0 + gp < naux) ? p0 + gp : naux;
  const int hp = (int)(nkc * 32);              // m pairs per aux slice
  const int nitems = (int)((p1 - p0) * hp);    // <= kOzMaxGroupK / 2
  const int tid = threadIdx.x;
  double2 v[16];
  double mx = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int item = k * 256 + tid;
    v[k] = double2{0.0, 0.0};
    if (item < nitems) {
      const int64_t P = p0 + item / hp;
      const int64_t m = 2 * (item % hp);
      const double* src = B + (P * nmo + i) * nocc + m;
      if (VEC && m + 1 < nocc) {
        v[k] = *(const double2*)src;
      } else {
        if (m < nocc) v[k].x = src[0];
        if (m + 1 < nocc) v[k].y = src[1];
      }
    }
    mx = fmax(mx, fmax(fabs(v[k].x), fabs(v[k].y)));
  }
  mx = oz_wave_max(mx);
  if ((tid & 63) == 0) wmax[tid >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
  int e = 0;
  if (mx > 0.0) frexp(mx, &e);  // mx < 2^e
  if (tid == 0) E[g * nmo + i] = e;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int item = k * 256 + tid;
    if (item >= nitems) continue;
    const int64_t P = p0 + item / hp;
    const int64_t m = 2 * (item % hp);
    uint16_t* o = (uint16_t*)(S + ((P * nmo + i) * nkc + (m >> 6)) * 256 + (m & 63));
    double x0 = __dmul_rn(ldexp(v[k].x, -e), 127.0);  // ldexp exact
    double x1 = __dmul_rn(ldexp(v[k].y, -e), 127.0);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const double s0 = rint(x0), s1 = rint(x1);
      o[s * 32] = (uint16_t)((uint8_t)(int8_t)(int)s0 | ((uint16_t)(uint8_t)(int8_t)(int)s1 << 8));
      x0 = __dmul_rn(__dsub_rn(x0, s0), 254.0);
      x1 = __dmul_rn(__dsub_rn(x1, s1), 254.0);
    }
  }
}

struct OzParams {
  const int8_t* S;
  const int32_t* E;
  double* part;  // partial slots (one per group of this launch), stride nmo*nmo
  int64_t nmo, nkc, naux, gp;
  int64_t g0;    // first group of this launch
  int64_t ntiles, nwork;
};

// 256 threads = 4 waves (2 x 2), 128 x 128 upper tile, wave tile 64 x 64
// (4 x 4 MFMA blocks of 16 x 16); one workgroup per (tile, exponent group).
// One stage = one 64-deep k-chunk of one aux slice with all 4 slices of both
// operands (2 x 128 rows x 256 bytes); two stages in LDS, filled straight
// from global memory by global_load_lds.  LDS unit (16 bytes) u of row r
// holds global unit u ^ (r & 15): the 16 lanes of a ds_read_b128 phase read
// 16 rows at one logical unit, which the XOR spreads over all 64 banks.
// The B fragment is MFMA src0 and the A fragment src1, so accumulator
// register r of lane l holds K(row = .. + l%16, col = .. + 4*(l/16) + r).
// Both fragments use the same lane -> k assignment, so the k order inside an
// MFMA does not matter.  The int32 level sums cover the whole group and are
// folded into FP64 once, after the k loop.
__global__ __launch_bounds__(256) void rik_oz_i8_kernel(OzParams g) {
  constexpr int ROWS = 128, UPR = 16, OPU = ROWS * UPR, STAGE = 2 * OPU;
  extern __shared__ uint4 oz_lds[];

  const int64_t w = xcd_remap(blockIdx.x, gridDim.x);
  if (w >= g.nwork) return;
  const int64_t gl = w / g.ntiles;
  const int64_t grp = g.g0 + gl;
  int tm, tn;
  upper_tile(w % g.ntiles, tm, tn);
  const int64_t m0 = (int64_t)tm * ROWS, n0 = (int64_t)tn * ROWS;
  const int64_t b0 = grp * g.gp;
  const int64_t b1 = (b0 + g.gp < g.naux) ? b0 + g.gp : g.naux;
  const int64_t niter = (b1 - b0) * g.nkc;

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wave_m = wave & 1, wave_n = wave >> 1;

  i32x4_t acc[4][4][4];  // [level][i][j]
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[q][i][j] = i32x4_t{0, 0, 0, 0};

  // per-lane byte offsets of the 16 direct-to-LDS loads (8 per operand)
  // relative to the (P, kc) slab S + (P*nmo*nkc + kc)*256
  uint32_t goff[2][OPU / 256];
#pragma unroll
  for (int op = 0; op < 2; ++op) {
    const int64_t base = op == 0 ? m0 : n0;
#pragma unroll
    for (int q = 0; q < OPU / 256; ++q) {
      const int u = (wave * (OPU / 256) + q) * 64 + lane;
      const int row = u >> 4, pos = u & 15;
      int64_t i = base + row;
      i = i < g.nmo ? i : g.nmo - 1;
      goff[op][q] = (uint32_t)(i * g.nkc * 256 + (pos ^ (row & 15)) * 16);
    }
  }
  const int64_t pstride = g.nmo * g.nkc * 256;
  int64_t iP = b0, ikc = 0;  // next (P, kc) to load
  auto issue = [&](int slot) {
    const char* slab = (const char*)g.S + iP * pstride + ikc * 256;
    uint4* st = oz_lds + slot * STAGE;
#pragma unroll
    for (int op = 0; op < 2; ++op)
#pragma unroll
      for (int q = 0; q < OPU / 256; ++q)
        __builtin_amdgcn_global_load_lds(
            (const oz_glb_t*)(slab + goff[op][q]),
            (oz_lds_t*)(st + op * OPU + (wave * (OPU / 256) + q) * 64), 16, 0, 0);
    if (++ikc == g.nkc) {
      ikc = 0;
      ++iP;
    }
  };

  // LDS fragment offsets (units): row r = wave_x*64 + i*16 + l15 holds
  // logical unit c at c ^ l15, and (s*4 + kg) ^ l15 = ((s ^ (l15>>2))*4) + (kg ^ (l15&3))
  const int l15 = lane & 15, kg = lane >> 4;
  const int arow = (wave_m * 64 + l15) * UPR, brow = (wave_n * 64 + l15) * UPR;
  const int kx = kg ^ (l15 & 3), hx = l15 >> 2;

  // L2 prefetch of the slab PF k-chunks ahead: each lane touches one 128-byte
  // line (2 per row of 256 bytes, 512 lines for both operands, 2 loads per
  // wave) through global_load_lds into a 1 KB per-wave scratch area of LDS
  // that is never read, so no VGPR is tied up by a load in flight.
  constexpr int PF = 3;
  int64_t pP = b0, pkc = 0;  // next (P, kc) to prefetch
  auto prefetch = [&]() {
    const char* slab = (const char*)g.S + pP * pstride + pkc * 256;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int line = (wave * 2 + q) * 64 + lane;  // 0..511
      const int op = line >> 8, row = (line & 255) >> 1, half = line & 1;
      int64_t i = (op == 0 ? m0 : n0) + row;
      i = i < g.nmo ? i : g.nmo - 1;
      __builtin_amdgcn_global_load_lds((const oz_glb_t*)(slab + i * g.nkc * 256 + half * 128),
                                       (oz_lds_t*)(oz_lds + 2 * STAGE + wave * 64), 16, 0, 0);
    }
    if (++pkc == g.nkc) {
      pkc = 0;
      ++pP;
    }
  };
  // k-chunk j is prefetched in iteration j - PF (the prologue covers 1..PF-1)
  if (niter > 0) {
    for (int j = 0; j < PF && j < niter; ++j) {
      if (j > 0) prefetch();
      else if (++pkc == g.nkc) {  // chunk 0 is loaded for real right away
        pkc = 0;
        ++pP;
      }
    }
    issue(0);
  }
  for (int64_t it = 0; it < niter; ++it) {
    // stage it landed (this wave's loads; the prefetch issued after it may
    // still be in flight), then every wave's loads landed and every wave
    // finished reading stage it-1, whose buffer is refilled next
    if (it >= 1 && it - 1 + PF < niter) wait_vmcnt<2>();
    else wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");  // no LDS access moves across the barrier
    const uint4* As = oz_lds + (it & 1) * STAGE;
    const uint4* Bs = As + OPU;
    i32x4_t a[4][4], b[4][4];
    auto load_a = [&](int s) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        a[s][i] = __builtin_bit_cast(i32x4_t, As[arow + i * 16 * UPR + ((s ^ hx) * 4 + kx)]);
    };
    auto load_b = [&](int t) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        b[t][j] = __builtin_bit_cast(i32x4_t, Bs[brow + j * 16 * UPR + ((t ^ hx) * 4 + kx)]);
    };
    auto prod = [&](int s, int t) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[s + t][i][j] =
              __builtin_amdgcn_mfma_i32_16x16x64_i8(b[t][j], a[s][i], acc[s + t][i][j], 0, 0, 0);
    };
    // slices enter in pairs (a_s, b_s); each pair unlocks the products that
    // need it, so the first MFMAs wait for 8 LDS reads, not 20
    load_a(0);
    load_b(0);
    if (it + 1 < niter) issue((int)((it + 1) & 1));
    if (it + PF < niter) prefetch();
    load_a(1);
    load_b(1);
    prod(0, 0);
    load_a(2);
    load_b(2);
    prod(0, 1);
    prod(1, 0);
    prod(1, 1);
    load_a(3);
    load_b(3);
    prod(0, 2);
    prod(2, 0);
    prod(1, 2);
    prod(2, 1);
    prod(0, 3);
    prod(3, 0);
  }

  // fold: K_part = 2^(e_i+e_j) / 127^2 * sum_q 254^-q L_q (all zero if niter == 0)
  double* out = g.part + gl * g.nmo * g.nmo;
  const int32_t* Eg = g.E + grp * g.nmo;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int64_t row = m0 + wave_m * 64 + i * 16 + (lane & 15);
    if (row >= g.nmo) continue;
    const int ei = niter > 0 ? Eg[row] : 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t col = n0 + wave_n * 64 + j * 16 + 4 * (lane >> 4) + r;
        if (col >= g.nmo) continue;
        const int ej = niter > 0 ? Eg[col] : 0;
        double v = (double)acc[3][i][j][r];
        v = (double)acc[2][i][j][r] + v * (1.0 / 254.0);
        v = (double)acc[1][i][j][r] + v * (1.0 / 254.0);
        v = (double)acc[0][i][j][r] + v * (1.0 / 254.0);
        out[row + col * g.nmo] = ldexp(v * (1.0 / 16129.0), ei + ej);
      }
  }
}

int64_t oz_group_size(int64_t nocc, int64_t nmo, int64_t naux, int64_t ncu) {
  const int64_t nkc = std::max<int64_t>(1, (nocc + 63) / 64);
  const int64_t tiles = (nmo + 127) / 128;
  const int64_t ntiles = tiles * (tiles + 1) / 2;
  // the int32 bound caps the group; below that, keep >= 2 workgroups per CU
  const int64_t gmax = std::max<int64_t>(1, kOzMaxGroupK / (64 * nkc));
  const int64_t gfill = std::max<int64_t>(1, naux * ntiles / (2 * std::max<int64_t>(ncu, 1)));
  return std::max<int64_t>(1, std::min<int64_t>({gmax, gfill, std::max<int64_t>(naux, 1)}));
}

int oz_slice(const double* B, int64_t nocc, int64_t nmo, int64_t naux, int64_t gp,
             int8_t* S, int32_t* E, hipStream_t s) {
  const int64_t nkc = std::max<int64_t>(1, (nocc + 63) / 64);
  const int64_t ngroups = (naux + gp - 1) / gp;
  if (gp * nkc * 64 > kOzMaxGroupK)
    return fail(RT_ERR_ARG, "exponent group of %lld x %lld values exceeds %lld",
                (long long)gp, (long long)(nkc * 64), (long long)kOzMaxGroupK);
  const bool vec = (nocc % 2 == 0) && (((uintptr_t)B & 15) == 0);
  KernelTimer kt("rik_slice", s);
  if (vec)
    hipLaunchKernelGGL(oz_slice_kernel<true>, dim3((unsigned)(nmo * ngroups)), dim3(256), 0,
                       s, B, nocc, nkc, nmo, naux, gp, S, E);
  else
    hipLaunchKernelGGL(oz_slice_kernel<false>, dim3((unsigned)(nmo * ngroups)), dim3(256), 0,
                       s, B, nocc, nkc, nmo, naux, gp, S, E);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

int rik_ozaki_i8(const double* B, int64_t nocc, int64_t nmo, int64_t naux, double* K,
                 hipStream_t s) {
  const int64_t nkc = std::max<int64_t>(1, (nocc + 63) / 64);
  const int64_t gp = oz_group_size(nocc, nmo, naux, device_cu_count());
  const int64_t ngroups = naux > 0 ? (naux + gp - 1) / gp : 1;
  int8_t* S = nullptr;
  int32_t* E = nullptr;
  RTA_TRY(scratch_get(SCRATCH_LOWP_A, (size_t)std::max<int64_t>(naux, 1) * nmo * nkc * 256,
                      (void**)&S));
  RTA_TRY(scratch_get(SCRATCH_LOWP_B, sizeof(int32_t) * (size_t)ngroups * nmo,
                      (void**)&E));
  if (naux > 0 && nocc > 0) RTA_TRY(oz_slice(B, nocc, nmo, naux, gp, S, E, s));

  // one partial slot per group; launches cover at most kOzMaxSlotBytes of
  // slots each, and every launch after the first accumulates into K
  const int64_t tiles = (nmo + 127) / 128;
  const int64_t ntiles = tiles * (tiles + 1) / 2;
  const int64_t slot_bytes = (int64_t)sizeof(double) * nmo * nmo;
  const int64_t per_launch =
      std::max<int64_t>(1, std::min<int64_t>(ngroups, kOzMaxSlotBytes / slot_bytes));
  double* part = nullptr;
  RTA_TRY(scratch_get(SCRATCH_GEMM_PARTIAL, (size_t)slot_bytes * per_launch, (void**)&part));
  const size_t lds = 2 * 2 * 128 * 16 * 16 + 4 * 1024;  // 2 stages + prefetch scratch
  static bool attr = false;
  if (!attr) {
    RTA_HIP(hipFuncSetAttribute((const void*)rik_oz_i8_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  for (int64_t g0 = 0; g0 < ngroups; g0 += per_launch) {
    const int64_t ng = std::min<int64_t>(per_launch, ngroups - g0);
    OzParams p;
    p.S = S;
    p.E = E;
    p.part = part;
    p.nmo = nmo;
    p.nkc = (naux > 0 && nocc > 0) ? nkc : 0;
    p.naux = naux;
    p.gp = gp;
    p.g0 = g0;
    p.ntiles = ntiles;
    p.nwork = ntiles * ng;
    {
      KernelTimer kt("rik_oz_i8", s);
      hipLaunchKernelGGL(rik_oz_i8_kernel, dim3((unsigned)p.nwork), dim3(256), lds, s, p);
      RTA_HIP(hipGetLastError());
    }
    RTA_TRY(reduce_partials_sym(part, ng, nmo, K, g0 > 0, s));
  }
  return RT_OK;
}

}  // namespace rta

extern "C" int64_t rt_ozaki_i8_group_size(int64_t nocc, int64_t nmo, int64_t naux) {
  return rta::oz_group_size(nocc, nmo, naux, rta::device_cu_count());
}
