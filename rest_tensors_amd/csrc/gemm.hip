// FP64 GEMM entry points (A6: _dgemm / _dgemm_full / general_dgemm_f) and the
// launch helpers used by ao2mo, density and RI-K.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "common.hpp"
#include "gemm.hpp"
#include "mfma_tn.hpp"
#include "timer.hpp"

namespace rta {

template <int BM, int BN>
static void launch_variant(const GemmParams& g, bool ta, bool tb,
                           unsigned grid, hipStream_t s) {
  constexpr int BK = 16;
  if (!ta && !tb)
    hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, BK, false, false>),
                       dim3(grid), dim3(256), 0, s, g);
  else if (ta && !tb)
    hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, BK, true, false>),
                       dim3(grid), dim3(256), 0, s, g);
  else if (!ta && tb)
    hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, BK, false, true>),
                       dim3(grid), dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, BK, true, true>),
                       dim3(grid), dim3(256), 0, s, g);
}

static int launch(const GemmParams& g, bool ta, bool tb, int tile,
                  hipStream_t s, const char* timer_name) {
  if (g.nwork <= 0) return RT_OK;
  RTA_CHECK(g.nwork < (int64_t(1) << 31), "GEMM grid too large (%lld tiles)",
            (long long)g.nwork);
  KernelTimer kt(timer_name, s);
  if (tile == 128)
    launch_variant<128, 128>(g, ta, tb, (unsigned)g.nwork, s);
  else
    launch_variant<64, 64>(g, ta, tb, (unsigned)g.nwork, s);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// ---- TN fast path (both operands k-major, direct-to-LDS pipeline) --------
static int g_num_cu = 0;
int device_cu_count() {
  if (g_num_cu == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) ==
            hipSuccess && n > 0)
      g_num_cu = n;
    else
      g_num_cu = 256;
  }
  return g_num_cu;
}

// Engine configuration knobs (environment, read once): RTA_TN_STAGES in
// {2,3,4} (k-chunks in flight), RTA_TN_BK in {16,32} (k-chunk depth; unset:
// 16 for 128-row tiles, 32 for 64-row tiles).  The 128-row default, BK 16 with
// 2 stages, takes 64 KB of LDS so two workgroups share a CU and one's barrier
// waits hide behind the other's MFMAs (228 VGPRs: two waves per SIMD).
static int env_int(const char* name, int def) {
  const char* e = getenv(name);
  return e ? atoi(e) : def;
}
static int tn_stages() {
  static int st = [] {
    int v = env_int("RTA_TN_STAGES", 2);
    return (v == 2 || v == 3 || v == 4) ? v : 2;
  }();
  return st;
}
static int tn_bk_pref() {
  static int bk = [] {
    int v = env_int("RTA_TN_BK", 0);
    return (v == 16 || v == 32) ? v : 0;
  }();
  return bk;
}

bool tn_eligible(const GemmDesc& d) {
  return d.ta && !d.tb && d.K > 0 && d.K % 16 == 0 && d.lda % 2 == 0 &&
         d.ldb % 2 == 0 && d.sA % 2 == 0 && d.sB % 2 == 0 &&
         ((uintptr_t)d.A & 15) == 0 && ((uintptr_t)d.B & 15) == 0;
}

template <int BM, int BN, int BK, int NS>
static int tn_launch_cfg(const TnParams& p, hipStream_t s) {
  const size_t lds = (size_t)NS * (BM + BN) * BK * sizeof(double);
  static bool attr_set = false;
  if (!attr_set) {
    RTA_HIP(hipFuncSetAttribute((const void*)dgemm_tn_kernel<BM, BN, BK, NS>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    attr_set = true;
  }
  hipLaunchKernelGGL((dgemm_tn_kernel<BM, BN, BK, NS>), dim3((unsigned)p.nwork),
                     dim3(256), lds, s, p);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

static int tn_launch(const GemmParams& g, int tile, hipStream_t s,
                     const char* timer_name) {
  if (g.nwork <= 0) return RT_OK;
  RTA_CHECK(g.nwork < (int64_t(1) << 31), "GEMM grid too large");
  TnParams p;
  p.A = g.A;
  p.B = g.B;
  p.C = g.C;
  p.M = g.M;
  p.N = g.N;
  p.K = g.K;
  p.lda = g.lda;
  p.ldb = g.ldb;
  p.sA = g.sA;
  p.sB = g.sB;
  p.c_rs = g.c_rs;
  p.c_cs = g.c_cs;
  p.sC = g.sC;
  p.alpha = g.alpha;
  p.beta = g.beta;
  p.nbatch = g.nbatch;
  p.bchunk = g.bchunk;
  p.nchunks = g.nchunks;
  p.tiles_m = g.tiles_m;
  p.tiles_n = g.tiles_n;
  p.ntiles = g.ntiles;
  p.nwork = g.nwork;
  p.reduce = g.reduce;
  p.upper = g.upper;
  KernelTimer kt(timer_name, s);
  const int bk = tn_bk_pref() ? tn_bk_pref() : (tile == 128 ? 16 : 32);
  const bool bk32 = bk == 32 && p.K % 32 == 0;
  if (tile == 128) {
    if (bk32) return tn_launch_cfg<128, 128, 32, 2>(p, s);  // 2 x 64 KB LDS
    switch (tn_stages()) {
      case 3: return tn_launch_cfg<128, 128, 16, 3>(p, s);
      case 4: return tn_launch_cfg<128, 128, 16, 4>(p, s);
      default: return tn_launch_cfg<128, 128, 16, 2>(p, s);
    }
  }
  if (bk32) return tn_launch_cfg<64, 64, 32, 3>(p, s);
  return tn_launch_cfg<64, 64, 16, 4>(p, s);
}

int gemm_pick_tile(int64_t M, int64_t N, int64_t batch) {
  int64_t t128 = ((M + 127) / 128) * ((N + 127) / 128) * batch;
  if (M >= 128 && N >= 128 && t128 >= 512) return 128;
  return 64;
}

static void fill_common(GemmParams& g, const GemmDesc& d) {
  g.A = d.A;
  g.B = d.B;
  g.C = d.C;
  g.M = d.M;
  g.N = d.N;
  g.K = d.K;
  g.lda = d.lda;
  g.ldb = d.ldb;
  g.sA = d.sA;
  g.sB = d.sB;
  g.c_rs = d.c_rs;
  g.c_cs = d.c_cs;
  g.sC = d.sC;
  g.alpha = d.alpha;
  g.beta = d.beta;
  g.nbatch = d.batch;
}

// C = beta*C for the k == 0 case (BLAS: no product term).
__global__ void scale_kernel(double* C, int64_t M, int64_t N, int64_t c_rs,
                             int64_t c_cs, int64_t sC, int64_t batch,
                             double beta) {
  int64_t total = M * N * batch;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t m = e % M, n = (e / M) % N, b = e / (M * N);
    double* p = C + b * sC + m * c_rs + n * c_cs;
    *p = (beta == 0.0) ? 0.0 : beta * (*p);
  }
}

int gemm_batched(const GemmDesc& d, hipStream_t stream, const char* timer_name,
                 int tile) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return RT_OK;
  if (d.K <= 0) {
    int64_t total = d.M * d.N * d.batch;
    unsigned grid = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(scale_kernel, dim3(grid), dim3(256), 0, stream, d.C,
                       d.M, d.N, d.c_rs, d.c_cs, d.sC, d.batch, d.beta);
    RTA_HIP(hipGetLastError());
    return RT_OK;
  }
  if (tile == 0) tile = gemm_pick_tile(d.M, d.N, d.batch);
  GemmParams g{};
  fill_common(g, d);
  g.bchunk = 1;
  g.nchunks = d.batch;
  g.tiles_m = (int)((d.M + tile - 1) / tile);
  g.tiles_n = (int)((d.N + tile - 1) / tile);
  g.ntiles = (int64_t)g.tiles_m * g.tiles_n;
  g.nwork = g.ntiles * g.nchunks;
  g.reduce = 0;
  g.upper = 0;
  if (tn_eligible(d)) return tn_launch(g, tile, stream, timer_name);
  return launch(g, d.ta, d.tb, tile, stream, timer_name);
}

// Sums `nslot` partial M x M matrices (slot stride M*M, column-major) over
// the upper triangle in slot order and writes alpha*sum to out (ld ldo) and
// its mirror; 32 x 32 tiles, transposed through LDS for the lower half.
__global__ __launch_bounds__(256) void reduce_sym_kernel(
    const double* __restrict__ part, int64_t nslot, int64_t M, double alpha,
    double* __restrict__ out, int64_t ldo) {
  __shared__ double T[32][33];
  int tm, tn;
  upper_tile(blockIdx.x, tm, tn);
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const int64_t slot = M * M;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int64_t m = (int64_t)tm * 32 + tx, n = (int64_t)tn * 32 + ty + 8 * q;
    double s = 0.0;
    if (m < M && n < M && m <= n) {
      const double* p = part + m + n * M;
      for (int64_t k = 0; k < nslot; ++k) s += p[k * slot];
      s *= alpha;
      out[m + n * ldo] = s;
    }
    T[ty + 8 * q][tx] = s;
  }
  __syncthreads();
  // mirror: out(n, m) = value(m, n) for m < n
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int64_t n = (int64_t)tn * 32 + tx, m = (int64_t)tm * 32 + ty + 8 * q;
    if (m < M && n < M && m < n) out[n + m * ldo] = T[tx][ty + 8 * q];
  }
}

int gemm_reduce_sym(const GemmDesc& d, int64_t bchunk, double* partial,
                    double* out, int64_t ldo, hipStream_t stream,
                    const char* timer_name, int tile) {
  RTA_CHECK(d.M == d.N, "symmetric reduce needs M == N");
  if (d.M <= 0) return RT_OK;
  if (bchunk <= 0) bchunk = d.batch;
  GemmParams g{};
  fill_common(g, d);
  g.C = partial;
  g.c_rs = 1;
  g.c_cs = d.M;
  g.sC = d.M * d.M;
  g.bchunk = bchunk;
  g.nchunks = (d.batch + bchunk - 1) / bchunk;
  g.tiles_m = g.tiles_n = (int)((d.M + tile - 1) / tile);
  g.ntiles = (int64_t)g.tiles_m * (g.tiles_m + 1) / 2;
  g.nwork = g.ntiles * g.nchunks;
  g.reduce = 1;
  g.upper = 1;
  if (d.K <= 0 || d.batch <= 0) {
    // no product terms: one all-zero slot (a fill kernel, not a memset, so
    // the sequence stays capturable in a HIP graph)
    unsigned grid = (unsigned)std::min<int64_t>((d.M * d.M + 255) / 256, 65536);
    hipLaunchKernelGGL(scale_kernel, dim3(grid), dim3(256), 0, stream, partial,
                       d.M, d.M, (int64_t)1, d.M, (int64_t)0, (int64_t)1, 0.0);
    RTA_HIP(hipGetLastError());
    g.nchunks = 1;
  } else if (tn_eligible(d)) {
    RTA_TRY(tn_launch(g, tile, stream, timer_name));
  } else {
    RTA_TRY(launch(g, d.ta, d.tb, tile, stream, timer_name));
  }
  int64_t t32 = (d.M + 31) / 32;
  {
    KernelTimer kt("reduce_sym", stream);
    hipLaunchKernelGGL(reduce_sym_kernel, dim3((unsigned)(t32 * (t32 + 1) / 2)),
                       dim3(256), 0, stream, partial, g.nchunks, d.M, d.alpha,
                       out, ldo);
  }
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// Batches per work item for a batch-reducing GEMM: balances ntiles*nchunks
// work items of bchunk batches each over the CUs (one 128x128 workgroup per
// CU), charging each extra chunk the time to stream its M x M partial.
int64_t reduce_bchunk(int64_t M, int64_t K, int64_t nbatch, int tile) {
  const int64_t t = (M + tile - 1) / tile;
  const int64_t ntiles = t * (t + 1) / 2;
  const int64_t slots = device_cu_count();
  // microseconds: one batch of one tile on one CU (128 flop/clk at 2.4 GHz),
  // and one M x M partial streamed at ~5 TB/s
  const double t_unit = 2.0 * tile * tile * (double)K / (128.0 * 2.4e3);
  const double t_part = (double)M * M * 8.0 / 5e6;
  int64_t best = nbatch;
  double best_cost = 1e300;
  for (int64_t bc = 1; bc <= nbatch; ++bc) {
    int64_t nch = (nbatch + bc - 1) / bc;
    if (bc > 1 && (nbatch + bc - 2) / (bc - 1) == nch) continue;  // same split
    int64_t rounds = (ntiles * nch + slots - 1) / slots;
    double cost = (double)rounds * bc * t_unit + (double)nch * t_part;
    if (cost < best_cost) {
      best_cost = cost;
      best = bc;
    }
  }
  return best;
}

static int parse_op(char c, bool* t) {
  if (c == 'N' || c == 'n') {
    *t = false;
    return RT_OK;
  }
  if (c == 'T' || c == 't' || c == 'C' || c == 'c') {
    *t = true;
    return RT_OK;
  }
  return fail(RT_ERR_ARG, "op must be 'N' or 'T', got '%c' (0x%02x)", c,
              (unsigned char)c);
}

}  // namespace rta

using namespace rta;

extern "C" int rt_dgemm_strided_batched_dev(
    char opa, char opb, int64_t m, int64_t n, int64_t k, double alpha,
    const double* A, int64_t lda, int64_t stride_a, const double* B,
    int64_t ldb, int64_t stride_b, double beta, double* C, int64_t ldc,
    int64_t stride_c, int64_t batch, void* stream) {
  GemmDesc d;
  RTA_TRY(parse_op(opa, &d.ta));
  RTA_TRY(parse_op(opb, &d.tb));
  RTA_CHECK(m >= 0 && n >= 0 && k >= 0 && batch >= 0, "negative dimension");
  // BLAS leading-dimension rules
  RTA_CHECK(lda >= std::max<int64_t>(1, d.ta ? k : m), "lda too small");
  RTA_CHECK(ldb >= std::max<int64_t>(1, d.tb ? n : k), "ldb too small");
  RTA_CHECK(ldc >= std::max<int64_t>(1, m), "ldc too small");
  d.M = m;
  d.N = n;
  d.K = k;
  d.A = A;
  d.lda = lda;
  d.sA = stride_a;
  d.B = B;
  d.ldb = ldb;
  d.sB = stride_b;
  d.C = C;
  d.c_rs = 1;
  d.c_cs = ldc;
  d.sC = stride_c;
  d.alpha = alpha;
  d.beta = beta;
  d.batch = batch;
  return gemm_batched(d, (hipStream_t)stream, "dgemm");
}

extern "C" int rt_dgemm_dev(char opa, char opb, int64_t m, int64_t n,
                            int64_t k, double alpha, const double* A,
                            int64_t lda, const double* B, int64_t ldb,
                            double beta, double* C, int64_t ldc,
                            void* stream) {
  return rt_dgemm_strided_batched_dev(opa, opb, m, n, k, alpha, A, lda, 0, B,
                                      ldb, 0, beta, C, ldc, 0, 1, stream);
}

// Sub-block semantics of general_dgemm_f (restmatr.f90:63-107): the block
// A[start_row_a : start_row_a+len_row_a, start_col_a : ...] of a rows_a x
// cols_a column-major matrix is op'ed and multiplied; leading dimensions are
// those of the full matrices (no section copies needed on the device).
extern "C" int rt_general_dgemm_dev(
    const double* A, int64_t rows_a, int64_t cols_a, int64_t start_row_a,
    int64_t len_row_a, int64_t start_col_a, int64_t len_col_a, char opa,
    const double* B, int64_t rows_b, int64_t cols_b, int64_t start_row_b,
    int64_t len_row_b, int64_t start_col_b, int64_t len_col_b, char opb,
    double* C, int64_t rows_c, int64_t cols_c, int64_t start_row_c,
    int64_t len_row_c, int64_t start_col_c, int64_t len_col_c, double alpha,
    double beta, void* stream) {
  bool ta, tb;
  RTA_TRY(parse_op(opa, &ta));
  RTA_TRY(parse_op(opb, &tb));
  RTA_CHECK(len_row_a >= 0 && len_col_a >= 0 && len_row_b >= 0 &&
                len_col_b >= 0 && len_row_c >= 0 && len_col_c >= 0 &&
                start_row_a >= 0 && start_col_a >= 0 && start_row_b >= 0 &&
                start_col_b >= 0 && start_row_c >= 0 && start_col_c >= 0,
            "negative block start or length");
  RTA_CHECK(start_row_a + len_row_a <= rows_a &&
                start_col_a + len_col_a <= cols_a &&
                start_row_b + len_row_b <= rows_b &&
                start_col_b + len_col_b <= cols_b &&
                start_row_c + len_row_c <= rows_c &&
                start_col_c + len_col_c <= cols_c,
            "sub-block outside its matrix (A %lldx%lld, B %lldx%lld, C "
            "%lldx%lld)",
            (long long)rows_a, (long long)cols_a, (long long)rows_b,
            (long long)cols_b, (long long)rows_c, (long long)cols_c);
  // shape rules of _dgemm (matrix_blas_lapack.rs:132-150)
  int64_t am = ta ? len_col_a : len_row_a, ak = ta ? len_row_a : len_col_a;
  int64_t bk = tb ? len_col_b : len_row_b, bn = tb ? len_row_b : len_col_b;
  RTA_CHECK(ak == bk && am == len_row_c && bn == len_col_c,
            "block shapes do not match: op(A) %lldx%lld, op(B) %lldx%lld, C "
            "%lldx%lld",
            (long long)am, (long long)ak, (long long)bk, (long long)bn,
            (long long)len_row_c, (long long)len_col_c);
  GemmDesc d;
  d.ta = ta;
  d.tb = tb;
  d.M = len_row_c;
  d.N = len_col_c;
  d.K = ak;
  d.A = A + start_row_a + start_col_a * rows_a;
  d.lda = std::max<int64_t>(rows_a, 1);
  d.B = B + start_row_b + start_col_b * rows_b;
  d.ldb = std::max<int64_t>(rows_b, 1);
  d.C = C + start_row_c + start_col_c * rows_c;
  d.c_rs = 1;
  d.c_cs = std::max<int64_t>(rows_c, 1);
  d.alpha = alpha;
  d.beta = beta;
  d.batch = 1;
  return gemm_batched(d, (hipStream_t)stream, "dgemm");
}
