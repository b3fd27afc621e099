// Packed-upper MatrixUpper pack/unpack (A8/A9) and the device block movers
// (copy_mm / copy_mr / copy_rm / copy_rr on device buffers).
//
// Packed order (index.rs:209-233, matrix_trait.rs:170-217): element (i,j),
// i <= j, sits at j*(j+1)/2 + i, i.e. the upper triangle column by column.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "common.hpp"
#include "layout.hpp"
#include "timer.hpp"

namespace rta {

// Exact integer square root (floor) for 0 <= x < 2^62.
static int64_t isqrt64(int64_t x) {
  if (x <= 0) return 0;
  int64_t r = (int64_t)__builtin_sqrt((double)x);
  while (r * r > x) --r;
  while ((r + 1) * (r + 1) <= x) ++r;
  return r;
}

// full[i + n*j] = packed[min(i,j) + max(i,j)*(max(i,j)+1)/2]; 32x32 tiles:
// the upper half of each tile reads packed columns contiguously, the lower
// half is the LDS transpose of the mirrored tile.
__global__ __launch_bounds__(256) void unpack_kernel(
    const double* __restrict__ packed, int64_t n, double* __restrict__ full) {
  __shared__ double T[32][33];
  const int64_t tm = blockIdx.x, tn = blockIdx.y;  // output tile (rows, cols)
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  if (tm <= tn) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int64_t i = tm * 32 + tx, j = tn * 32 + ty + 8 * q;
      if (i < n && j < n) {
        int64_t lo = i < j ? i : j, hi = i < j ? j : i;
        full[i + n * j] = packed[hi * (hi + 1) / 2 + lo];
      }
    }
  } else {
    // read tile (tn, tm) of the upper triangle, write it transposed
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int64_t i = tn * 32 + tx, j = tm * 32 + ty + 8 * q;  // i <= j here
      double v = 0.0;
      if (i < n && j < n) v = packed[j * (j + 1) / 2 + i];
      T[ty + 8 * q][tx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int64_t i = tm * 32 + tx, j = tn * 32 + ty + 8 * q;
      if (i < n && j < n) full[i + n * j] = T[tx][ty + 8 * q];
    }
  }
}

// packed[j(j+1)/2 + i] = src[i + n*j + slice*P] for i <= j, per slice P.
__global__ __launch_bounds__(256) void pack_kernel(const double* __restrict__ src,
                                                   int64_t n, int64_t nslice,
                                                   double* __restrict__ packed) {
  const int64_t npair = n * (n + 1) / 2;
  const int64_t j = blockIdx.x;
  const int64_t P = blockIdx.y;
  if (P >= nslice) return;
  const double* col = src + P * n * n + j * n;
  double* dst = packed + P * npair + j * (j + 1) / 2;
  for (int64_t i = threadIdx.x; i <= j; i += 256) dst[i] = col[i];
}

// dst[ox + x*dx + y*dy + z*dz] = src[ix + x*sx + y*sy + z*sz]
__global__ __launch_bounds__(256) void copy3d_kernel(
    int64_t X, int64_t Y, int64_t Z, const double* __restrict__ src,
    int64_t sx, int64_t sy, int64_t sz, double* __restrict__ dst, int64_t dx,
    int64_t dy, int64_t dz) {
  const int64_t total = X * Y * Z;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t x = e % X, y = (e / X) % Y, z = e / (X * Y);
    dst[x * dx + y * dy + z * dz] = src[x * sx + y * sy + z * sz];
  }
}

int copy3d(int64_t X, int64_t Y, int64_t Z, const double* src, int64_t sx,
           int64_t sy, int64_t sz, double* dst, int64_t dx, int64_t dy,
           int64_t dz, hipStream_t s) {
  int64_t total = X * Y * Z;
  if (total <= 0) return RT_OK;
  KernelTimer kt("copy3d", s);
  unsigned grid =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
  hipLaunchKernelGGL(copy3d_kernel, dim3(grid), dim3(256), 0, s, X, Y, Z, src,
                     sx, sy, sz, dst, dx, dy, dz);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// Batched 2D transpose through LDS: out[b*ob + p*op + q] = in[b*ib + p + q*iq]
// for p < P, q < Q, b < NB (the input runs along p, the output along q).  32 x 32
// tiles, a grid-stride loop over (tile, batch); both sides move 256-byte rows.
__global__ __launch_bounds__(256) void transpose_tile_kernel(
    const double* __restrict__ in, int64_t P, int64_t Q, int64_t NB, int64_t iq, int64_t ib,
    double* __restrict__ out, int64_t op, int64_t ob) {
  __shared__ double T[32][33];
  const int64_t tp = (P + 31) / 32, tq = (Q + 31) / 32;
  const int64_t ntile = tp * tq * NB;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int64_t t = blockIdx.x; t < ntile; t += gridDim.x) {
    const int64_t b = t / (tp * tq), rem = t % (tp * tq);
    const int64_t p0 = (rem % tp) * 32, q0 = (rem / tp) * 32;
    const double* src = in + b * ib;
    double* dst = out + b * ob;
    __syncthreads();  // the previous tile's reads of T are done
#pragma unroll
    for (int r = ty; r < 32; r += 8) {
      const int64_t p = p0 + tx, q = q0 + r;
      if (p < P && q < Q) T[r][tx] = src[p + q * iq];
    }
    __syncthreads();
#pragma unroll
    for (int r = ty; r < 32; r += 8) {
      const int64_t q = q0 + tx, p = p0 + r;
      if (p < P && q < Q) dst[p * op + q] = T[tx][r];
    }
  }
}

static int transpose_tiles(const double* in, int64_t P, int64_t Q, int64_t NB, int64_t iq,
                           int64_t ib, double* out, int64_t op, int64_t ob, hipStream_t s) {
  const int64_t ntile = ((P + 31) / 32) * ((Q + 31) / 32) * NB;
  if (ntile <= 0) return RT_OK;
  const unsigned grid = (unsigned)std::min<int64_t>(ntile, 262144);
  hipLaunchKernelGGL(transpose_tile_kernel, dim3(grid), dim3(256), 0, s, in, P, Q, NB, iq, ib,
                     out, op, ob);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

__global__ __launch_bounds__(256) void axpby3d_kernel(
    int64_t X, int64_t Y, int64_t Z, const double* __restrict__ src,
    int64_t sx, int64_t sy, int64_t sz, double* __restrict__ dst, int64_t dx,
    int64_t dy, int64_t dz, double beta) {
  const int64_t total = X * Y * Z;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int64_t x = e % X, y = (e / X) % Y, z = e / (X * Y);
    double* p = dst + x * dx + y * dy + z * dz;
    double v = src[x * sx + y * sy + z * sz];
    *p = (beta == 0.0) ? v : v + beta * (*p);
  }
}

int axpby3d(int64_t X, int64_t Y, int64_t Z, const double* src, int64_t sx,
            int64_t sy, int64_t sz, double* dst, int64_t dx, int64_t dy,
            int64_t dz, double beta, hipStream_t s) {
  int64_t total = X * Y * Z;
  if (total <= 0) return RT_OK;
  unsigned grid =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>((total + 255) / 256, 65536));
  hipLaunchKernelGGL(axpby3d_kernel, dim3(grid), dim3(256), 0, s, X, Y, Z, src,
                     sx, sy, sz, dst, dx, dy, dz, beta);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

}  // namespace rta

using namespace rta;

extern "C" int64_t rt_upper_dim(int64_t packed_len) {
  if (packed_len < 0) return -1;
  int64_t n = (isqrt64(8 * packed_len + 1) - 1) / 2;
  return (n * (n + 1) / 2 == packed_len) ? n : -1;
}

extern "C" int rt_upper_unpack_dev(const double* packed, int64_t packed_len,
                                   double* full, void* stream) {
  int64_t n = rt_upper_dim(packed_len);
  RTA_CHECK(n >= 0, "packed length %lld is not n(n+1)/2",
            (long long)packed_len);
  if (n == 0) return RT_OK;
  RTA_CHECK(packed && full, "NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  KernelTimer kt("upper_unpack", s);
  unsigned t = (unsigned)((n + 31) / 32);
  hipLaunchKernelGGL(unpack_kernel, dim3(t, t), dim3(256), 0, s, packed, n,
                     full);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

extern "C" int rt_upper_pack_dev(const double* full, int64_t n, double* packed,
                                 void* stream) {
  RTA_CHECK(n >= 0, "negative dimension");
  if (n == 0) return RT_OK;
  RTA_CHECK(full && packed, "NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  KernelTimer kt("upper_pack", s);
  hipLaunchKernelGGL(pack_kernel, dim3((unsigned)n, 1), dim3(256), 0, s, full,
                     n, (int64_t)1, packed);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

extern "C" int rt_rifull_to_matfull_symm_dev(const double* ri, int64_t nao,
                                             int64_t naux, double* out,
                                             void* stream) {
  RTA_CHECK(nao >= 0 && naux >= 0, "negative dimension");
  RTA_CHECK(naux < 65536 * 65536LL, "naux too large");
  if (nao == 0 || naux == 0) return RT_OK;
  RTA_CHECK(ri && out, "NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  KernelTimer kt("rifull_pack", s);
  for (int64_t p0 = 0; p0 < naux; p0 += 65535) {
    int64_t np = std::min<int64_t>(65535, naux - p0);
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)nao, (unsigned)np),
                       dim3(256), 0, s, ri + p0 * nao * nao, nao, np,
                       out + p0 * (nao * (nao + 1) / 2));
    RTA_HIP(hipGetLastError());
  }
  return RT_OK;
}

// ---- block movers ----------------------------------------------------------
// Every block must lie inside its declared array (the Fortran does not check;
// here a violation is RT_ERR_ARG).
static bool in_range(int64_t start, int64_t len, int64_t dim) {
  return start >= 0 && len >= 0 && start + len <= dim;
}

extern "C" int rt_copy_mm_dev(int64_t x_len, int64_t y_len, const double* f,
                              int64_t f_x, int64_t f_y, int64_t f_xs,
                              int64_t f_ys, double* t, int64_t t_x,
                              int64_t t_y, int64_t t_xs, int64_t t_ys,
                              void* stream) {
  RTA_CHECK(in_range(f_xs, x_len, f_x) && in_range(f_ys, y_len, f_y) &&
                in_range(t_xs, x_len, t_x) && in_range(t_ys, y_len, t_y),
            "copy_mm block outside its matrix");
  if (x_len == 0 || y_len == 0) return RT_OK;
  return copy3d(x_len, y_len, 1, f + f_xs + f_ys * f_x, 1, f_x, 0,
                t + t_xs + t_ys * t_x, 1, t_x, 0, (hipStream_t)stream);
}

// Offset and strides of the (a, b) block axes of a rank-3 tensor
// [n0, n1, n2] when mod fixes axis 2 (mod 0), axis 1 (mod 1) or axis 0 (mod 2)
// at index x3 (restmatr.f90:227-236).
static int ri_block(int mod, int64_t n0, int64_t n1, int64_t n2, int64_t as,
                   int64_t alen, int64_t bs, int64_t blen, int64_t x3,
                   int64_t* off, int64_t* sa, int64_t* sb) {
  switch (mod) {
    case 0:
      RTA_CHECK(in_range(as, alen, n0) && in_range(bs, blen, n1) &&
                    in_range(x3, 1, n2),
                "rank-3 block outside its tensor (mod 0)");
      *off = as + bs * n0 + x3 * n0 * n1;
      *sa = 1;
      *sb = n0;
      return RT_OK;
    case 1:
      RTA_CHECK(in_range(as, alen, n0) && in_range(x3, 1, n1) &&
                    in_range(bs, blen, n2),
                "rank-3 block outside its tensor (mod 1)");
      *off = as + x3 * n0 + bs * n0 * n1;
      *sa = 1;
      *sb = n0 * n1;
      return RT_OK;
    case 2:
      RTA_CHECK(in_range(x3, 1, n0) && in_range(as, alen, n1) &&
                    in_range(bs, blen, n2),
                "rank-3 block outside its tensor (mod 2)");
      *off = x3 + as * n0 + bs * n0 * n1;
      *sa = n0;
      *sb = n0 * n1;
      return RT_OK;
    default:
      return fail(RT_ERR_ARG, "copy mod must be 0, 1 or 2 (got %d)", mod);
  }
}

extern "C" int rt_copy_mr_dev(int64_t x_len, int64_t y_len, const double* f,
                              int64_t f_x, int64_t f_y, int64_t f_xs,
                              int64_t f_ys, double* t, int64_t t_x,
                              int64_t t_y, int64_t t_z, int64_t t_xs,
                              int64_t t_ys, int64_t t_x3, int mod,
                              void* stream) {
  RTA_CHECK(in_range(f_xs, x_len, f_x) && in_range(f_ys, y_len, f_y),
            "copy_mr source block outside its matrix");
  int64_t off, sa, sb;
  RTA_TRY(ri_block(mod, t_x, t_y, t_z, t_xs, x_len, t_ys, y_len, t_x3, &off,
                   &sa, &sb));
  if (x_len == 0 || y_len == 0) return RT_OK;
  return copy3d(x_len, y_len, 1, f + f_xs + f_ys * f_x, 1, f_x, 0, t + off, sa,
                sb, 0, (hipStream_t)stream);
}

extern "C" int rt_copy_rm_dev(int64_t x_len, int64_t y_len, const double* f,
                              int64_t f_x, int64_t f_y, int64_t f_z,
                              int64_t f_xs, int64_t f_ys, int64_t f_x3, int mod,
                              double* t, int64_t t_x, int64_t t_y,
                              int64_t t_xs, int64_t t_ys, void* stream) {
  RTA_CHECK(in_range(t_xs, x_len, t_x) && in_range(t_ys, y_len, t_y),
            "copy_rm target block outside its matrix");
  int64_t off, sa, sb;
  RTA_TRY(ri_block(mod, f_x, f_y, f_z, f_xs, x_len, f_ys, y_len, f_x3, &off,
                   &sa, &sb));
  if (x_len == 0 || y_len == 0) return RT_OK;
  return copy3d(x_len, y_len, 1, f + off, sa, sb, 0, t + t_xs + t_ys * t_x, 1,
                t_x, 0, (hipStream_t)stream);
}

extern "C" int rt_copy_rr_dev(int64_t x_len, int64_t y_len, int64_t z_len,
                              const double* f, int64_t f_x, int64_t f_y,
                              int64_t f_z, int64_t f_xs, int64_t f_ys,
                              int64_t f_zs, double* t, int64_t t_x,
                              int64_t t_y, int64_t t_z, int64_t t_xs,
                              int64_t t_ys, int64_t t_zs, void* stream) {
  RTA_CHECK(in_range(f_xs, x_len, f_x) && in_range(f_ys, y_len, f_y) &&
                in_range(f_zs, z_len, f_z) && in_range(t_xs, x_len, t_x) &&
                in_range(t_ys, y_len, t_y) && in_range(t_zs, z_len, t_z),
            "copy_rr block outside its tensor");
  if (x_len == 0 || y_len == 0 || z_len == 0) return RT_OK;
  return copy3d(x_len, y_len, z_len, f + f_xs + f_ys * f_x + f_zs * f_x * f_y,
                1, f_x, f_x * f_y, t + t_xs + t_ys * t_x + t_zs * t_x * t_y, 1,
                t_x, t_x * t_y, (hipStream_t)stream);
}

// RIFull::transpose_jik / transpose_jki / transpose_kji / transpose_ikj
// (ri.rs:225-290): every one is an index permutation of the column-major
// [n0, n1, n2] tensor (the reference builds them from chunked matrix
// transposes and index filters; the element order they produce is this one).
extern "C" int rt_rifull_transpose_dev(const double* src, int64_t n0, int64_t n1, int64_t n2,
                                       const char* order, double* out, void* stream) {
  RTA_CHECK(order, "order is NULL");
  RTA_CHECK(n0 >= 0 && n1 >= 0 && n2 >= 0, "negative dimension");
  const std::string o(order);
  RTA_CHECK(o == "jik" || o == "jki" || o == "kji" || o == "ikj",
            "order must be \"jik\", \"jki\", \"kji\" or \"ikj\", got \"%s\"", order);
  if (n0 * n1 * n2 == 0) return RT_OK;
  RTA_CHECK(src && out, "NULL argument");
  RTA_CHECK(src + n0 * n1 * n2 <= out || out + n0 * n1 * n2 <= src,
            "src and out overlap (the transpose is not in place)");
  hipStream_t s = (hipStream_t)stream;
  KernelTimer kt("rifull_transpose", s);
  if (o == "jik")  // [n1, n0, n2]: per k, an n0 x n1 transpose
    return transpose_tiles(src, n0, n1, n2, n0, n0 * n1, out, n1, n0 * n1, s);
  if (o == "jki")  // [n1, n2, n0]: one n0 x (n1 n2) transpose
    return transpose_tiles(src, n0, n1 * n2, 1, n0, 0, out, n1 * n2, 0, s);
  if (o == "kji")  // [n2, n1, n0]: per j, an n0 x n2 transpose
    return transpose_tiles(src, n0, n2, n1, n0 * n1, n0, out, n2 * n1, n2, s);
  // ikj: [n0, n2, n1], columns of length n0 move as units
  return copy3d(n0, n1, n2, src, 1, n0, n0 * n1, out, 1, n0 * n2, n0, s);
}

