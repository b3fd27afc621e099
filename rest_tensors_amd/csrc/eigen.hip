// Symmetric eigensolver on the device (A11):
//   _dsyev        src/matrix/matrix_blas_lapack.rs:319-348 (LAPACK DSYEV, uplo 'L')
//   lapack_power  src/matrix/matrix_blas_lapack.rs:1004-1064 (and _power, :599-653)
//
// DSYEV is scale, reduce to tridiagonal form, iterate on the tridiagonal
// (implicit QL/QR, DSTEQR), back-transform, sort.  The parts run where they
// fit:
//  * the reduction A = Q T Q^T (Householder reflectors of the lower variant:
//    reflector i annihilates A(i+2:n-1, i)) is O(n^3) bandwidth-bound work on
//    the device, one step per column: reflector, symmetric matrix-vector
//    product, rank-2 update of the trailing block (kept full and symmetric);
//  * the QL/QR iteration on (d, e) is an O(n^2) chain of dependent scalar
//    operations (one bulge chased at a time) and runs on the host, restating
//    LAPACK 3.12's DSTEQR (with DLAEV2, DLARTG, DLAPY2); every plane rotation
//    it makes is recorded instead of being applied;
//  * the eigenvectors of T are the recorded rotations applied to I, on the
//    device: one thread per row, every row sees the same rotation stream, in
//    the order DSTEQR applies them (DLASR, side 'R', pivot 'V');
//  * the back-transformation Z = Q Z_T applies the reflectors in blocks of 64
//    in compact WY form (I - V T V^T, DLARFT forward/columnwise) on the FP64
//    GEMM;
//  * the ascending sort is DSTEQR's selection sort, replayed on the host as a
//    column permutation.
// The rotation stream is uploaded in chunks while the host iteration goes on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "metric.hpp"
#include "timer.hpp"

namespace rta {

static unsigned ev_grid(int64_t len) { return (unsigned)((len + 255) / 256); }

// ---- device: tridiagonal reduction ----------------------------------------

__device__ __forceinline__ double ev_block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// W: full symmetric n x n (ld n).  Step i: x = W(i+1:n-1, i), m = n-1-i.
// Householder reflector H = I - tau v v^T with v(0) = 1 and H x = (beta, 0..)
// (DLARFG: beta = -sign(alpha) |x|, tau = (beta - alpha)/beta,
// v(1:) = x(1:)/(alpha - beta); tau = 0 when x(1:) = 0).  Writes d(i), e(i),
// tau(i), v into vbuf and into column i of V (rows i+1..n-1; V starts zero).
__global__ __launch_bounds__(256) void ev_house_kernel(const double* __restrict__ W, int64_t n,
                                                       int64_t i, double* __restrict__ V,
                                                       double* __restrict__ vbuf,
                                                       double* __restrict__ d,
                                                       double* __restrict__ e,
                                                       double* __restrict__ tau) {
  __shared__ double red[4];
  const int64_t m = n - 1 - i;
  const double* x = W + (i + 1) + i * n;
  double ss = 0.0;
  for (int64_t k = 1 + threadIdx.x; k < m; k += 256) ss += x[k] * x[k];
  ss = ev_block_sum(ss, red);
  const double alpha = x[0];
  double t = 0.0, beta = alpha, scal = 1.0;
  if (m > 1 && ss > 0.0) {
    beta = -copysign(sqrt(alpha * alpha + ss), alpha);
    t = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
  for (int64_t k = threadIdx.x; k < m; k += 256) {
    const double v = (k == 0) ? 1.0 : ((t == 0.0) ? x[k] : x[k] * scal);
    vbuf[k] = v;
    if (V) V[(i + 1 + k) + i * n] = v;
  }
  if (threadIdx.x == 0) {
    d[i] = W[i + i * n];
    e[i] = beta;
    tau[i] = t;
    if (m == 1) d[n - 1] = W[(n - 1) + (n - 1) * n];  // last step: no update follows
  }
}

// y(c) = tau * sum_r W22(r, c) v(r), one wave per column (W22 symmetric, so
// the column dot is the row dot).  W22 = W(i+1:, i+1:).
__global__ __launch_bounds__(256) void ev_symv_kernel(const double* __restrict__ W, int64_t n,
                                                      int64_t i, const double* __restrict__ v,
                                                      const double* __restrict__ tau,
                                                      double* __restrict__ y) {
  const double t = tau[i];
  if (t == 0.0) return;
  const int64_t m = n - 1 - i;
  const int64_t c = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= m) return;
  const int lane = threadIdx.x & 63;
  const double* col = W + (i + 1) + (i + 1 + c) * n;
  double acc = 0.0;
  for (int64_t r = lane; r < m; r += 64) acc += col[r] * v[r];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) y[c] = t * acc;
}

// w = y - (tau/2)(y . v) v
__global__ __launch_bounds__(256) void ev_w_kernel(int64_t n, int64_t i,
                                                   const double* __restrict__ v,
                                                   const double* __restrict__ tau,
                                                   const double* __restrict__ y,
                                                   double* __restrict__ w) {
  __shared__ double red[4];
  const double t = tau[i];
  if (t == 0.0) return;
  const int64_t m = n - 1 - i;
  double s = 0.0;
  for (int64_t k = threadIdx.x; k < m; k += 256) s += y[k] * v[k];
  s = ev_block_sum(s, red);
  const double a = -0.5 * t * s;
  for (int64_t k = threadIdx.x; k < m; k += 256) w[k] = y[k] + a * v[k];
}

// W22 -= v w^T + w v^T over the full trailing block; both products rounded
// on their own so that W22(r,c) and W22(c,r) stay bit-identical.
__global__ __launch_bounds__(256) void ev_syr2_kernel(double* __restrict__ W, int64_t n,
                                                      int64_t i, const double* __restrict__ v,
                                                      const double* __restrict__ w,
                                                      const double* __restrict__ tau) {
  if (tau[i] == 0.0) return;
  const int64_t m = n - 1 - i;
  const int64_t r = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (r >= m) return;
  const double vr = v[r], wr = w[r];
  const int64_t c0 = (int64_t)blockIdx.y * 8;
  const int64_t c1 = (c0 + 8 < m) ? c0 + 8 : m;
  double* base = W + (i + 1) + (i + 1) * n;
  for (int64_t c = c0; c < c1; ++c) {
    const double p1 = __dmul_rn(vr, w[c]);
    const double p2 = __dmul_rn(wr, v[c]);
    base[r + c * n] = __dsub_rn(base[r + c * n], __dadd_rn(p1, p2));
  }
}

// full symmetric copy of A's lower triangle, times `scale`
__global__ void ev_sym_lower_kernel(const double* __restrict__ A, int64_t lda, int64_t n,
                                    double scale, double* __restrict__ W) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= n * n) return;
  const int64_t r = e % n, c = e / n;
  const double v = (r >= c) ? A[r + c * lda] : A[c + r * lda];
  W[e] = v * scale;
}

// per-block max |x| over the lower triangle (partials, one per block)
__global__ __launch_bounds__(256) void ev_maxabs_kernel(const double* __restrict__ A,
                                                        int64_t lda, int64_t n,
                                                        double* __restrict__ part) {
  __shared__ double red[4];
  double mx = 0.0;
  for (int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x; e < n * n;
       e += (int64_t)gridDim.x * 256) {
    const int64_t r = e % n, c = e / n;
    if (r >= c) mx = fmax(mx, fabs(A[r + c * lda]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__global__ void ev_eye_kernel(double* __restrict__ Z, int64_t n) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e < n * n) Z[e] = (e % n == e / n) ? 1.0 : 0.0;
}

__global__ void ev_fill_kernel(double* __restrict__ X, int64_t len, double v) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e < len) X[e] = v;
}

// ---- device: recorded rotations applied to Z_T -----------------------------

// A sweep of `len` rotations on columns l..l+len of Z: rotation j mixes
// columns l+j and l+j+1 with (c, s) = cs[off + j] as DLASR (side 'R', pivot
// 'V') does:  z_{j+1} <- c z_{j+1} - s z_j,  z_j <- s z_{j+1} + c z_j;
// backward sweeps run j = len-1 .. 0, forward sweeps j = 0 .. len-1.
struct EvSweep {
  int32_t l, len, fwd, pad;
  int64_t off;
};

// One workgroup (one wave) per 64 rows of Z, one thread per row.  The column
// a rotation leaves for the next one is carried in a register; the column it
// takes in is loaded ROT_P - 1 chunks of ROT_K rotations ahead (a register
// ring), so several loads are in flight while the chain advances.  A sweep's
// (c, s) pairs are staged in LDS, ROT_SEG at a time, by the whole wave.
constexpr int ROT_K = 8, ROT_P = 4, ROT_SEG = 2048;

// rotations t in [t0, t1) of one sweep; t counts in the sweep's own order
// (backward sweeps: t = 0 is j = len - 1).  qs holds the pairs of this
// segment, qs[t - t0] in that order.
template <bool FWD>
__device__ __forceinline__ void ev_rot_segment(double* __restrict__ z, int64_t ldz, int64_t l,
                                               int len, int t0, int t1,
                                               const double2* __restrict__ qs,
                                               double& carry) {
  // load index of rotation t: the column it takes in
  const int64_t base = FWD ? l + 1 : l + len - 1;
  auto col = [&](int t) { return FWD ? base + t : base - t; };
  const int cnt = t1 - t0;
  auto rot = [&](int t, double a) {  // rotation t of the segment, column value a
    const double2 q = qs[t];
    const int64_t c = col(t0 + t);
    if (FWD) {
      z[(c - 1) * ldz] = q.y * a + q.x * carry;
      carry = q.x * a - q.y * carry;
    } else {
      z[(c + 1) * ldz] = q.x * carry - q.y * a;
      carry = q.y * carry + q.x * a;
    }
  };
  double buf[ROT_P][ROT_K];
  auto load = [&](int ch, double (&b)[ROT_K]) {
#pragma unroll
    for (int u = 0; u < ROT_K; ++u) b[u] = z[col(t0 + ch * ROT_K + u) * ldz];
  };
  auto comp = [&](int ch, const double (&b)[ROT_K]) {
#pragma unroll
    for (int u = 0; u < ROT_K; ++u) rot(ch * ROT_K + u, b[u]);
  };
  // full chunks: a branch-free steady state keeps chunks c .. c+P-2 in flight
  // (buf[p] holds chunk c + p at the top of each pass; <= 48 memory operations
  // outstanding, within the 6-bit vmcnt)
  const int nfull = cnt / ROT_K;
  int c = 0;
  if (nfull >= ROT_P - 1) {
#pragma unroll
    for (int p = 0; p < ROT_P - 1; ++p) load(p, buf[p]);
    for (; c + 2 * ROT_P - 2 < nfull; c += ROT_P) {
#pragma unroll
      for (int p = 0; p < ROT_P; ++p) {
        load(c + p + ROT_P - 1, buf[(p + ROT_P - 1) % ROT_P]);
        comp(c + p, buf[p]);
      }
    }
    // the loaded chunks c .. c+P-2 (those that exist), then the rest
#pragma unroll
    for (int p = 0; p < ROT_P - 1; ++p)
      if (c + p < nfull) comp(c + p, buf[p]);
    c += ROT_P - 1;
  }
  for (int t = c * ROT_K; t < cnt; ++t) rot(t, z[col(t0 + t) * ldz]);
}

__global__ __launch_bounds__(64) void ev_rot_kernel(double* __restrict__ Z, int64_t ldz,
                                                    int64_t nrow,
                                                    const EvSweep* __restrict__ sw, int nsw,
                                                    const double2* __restrict__ cs) {
  __shared__ double2 qs[ROT_SEG];
  const int lane = threadIdx.x;
  const int64_t r = blockIdx.x * (int64_t)64 + lane;
  const bool active = r < nrow;
  double* z = Z + (active ? r : 0);
  for (int k = 0; k < nsw; ++k) {
    const EvSweep S = sw[k];
    const int64_t l = S.l;
    const int len = S.len;
    double carry = 0.0;
    if (active) carry = S.fwd ? z[l * ldz] : z[(l + len) * ldz];
    for (int t0 = 0; t0 < len; t0 += ROT_SEG) {
      const int t1 = (t0 + ROT_SEG < len) ? t0 + ROT_SEG : len;
      __syncthreads();  // the previous segment's pairs are consumed
      for (int t = t0 + lane; t < t1; t += 64)
        qs[t - t0] = S.fwd ? cs[S.off + t] : cs[S.off + (len - 1 - t)];
      __syncthreads();
      if (!active) continue;
      if (S.fwd) ev_rot_segment<true>(z, ldz, l, len, t0, t1, qs, carry);
      else ev_rot_segment<false>(z, ldz, l, len, t0, t1, qs, carry);
    }
    if (active) {
      if (S.fwd) z[(l + len) * ldz] = carry;
      else z[l * ldz] = carry;
    }
  }
}

// out(:, j) = Z(:, perm(j))
__global__ void ev_gather_cols_kernel(const double* __restrict__ Z, int64_t n,
                                      const int32_t* __restrict__ perm, double* __restrict__ out,
                                      int64_t ldo) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e >= n * n) return;
  const int64_t r = e % n, c = e / n;
  out[r + c * ldo] = Z[r + (int64_t)perm[c] * n];
}

// ---- device: back-transformation -------------------------------------------

// T (kb x kb, upper, ld kb) of the compact WY form of H(i0)..H(i0+kb-1)
// (DLARFT forward/columnwise): T(j,j) = tau_j, T(0:j-1, j) =
// -tau_j T(0:j-1, 0:j-1) G(0:j-1, j) with G = V^T V.
__global__ __launch_bounds__(64) void ev_larft_kernel(const double* __restrict__ G, int kb,
                                                      const double* __restrict__ tau,
                                                      double* __restrict__ T) {
  __shared__ double Ts[64][65];
  __shared__ double col[64];
  const int a = threadIdx.x;
  for (int j = 0; j < kb; ++j) {
    const double tj = tau[j];
    double s = 0.0;
    if (a < j)
      for (int b = a; b < j; ++b) s += Ts[a][b] * G[b + j * kb];
    col[a] = -tj * s;
    __syncthreads();
    if (a < kb) Ts[a][j] = (a < j) ? col[a] : (a == j ? tj : 0.0);
    __syncthreads();
  }
  for (int j = 0; j < kb; ++j)
    if (a < kb) T[a + j * kb] = Ts[a][j];
}

// column scale: X(:, j) *= f(j)
__global__ void ev_scale_cols_kernel(double* __restrict__ X, int64_t n,
                                     const double* __restrict__ f) {
  const int64_t e = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (e < n * n) X[e] *= f[e / n];
}

// F = H + J - K/2 over the upper triangle, packed (j(j+1)/2 + i): the RHF
// Fock matrix from the core Hamiltonian and the J and K builds.
__global__ void ev_fock_packed_kernel(const double* __restrict__ H, const double* __restrict__ J,
                                      const double* __restrict__ K, int64_t n,
                                      double* __restrict__ F) {
  const int64_t j = blockIdx.y;
  const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (j >= n || i > j) return;
  const int64_t e = i + j * n;
  F[j * (j + 1) / 2 + i] = (H[e] + J[e]) - 0.5 * K[e];
}

// ---- host: DSTEQR restatement ----------------------------------------------

struct RotSink {
  // host mode: rotations applied to a host Z (LAPACK's own dlasr order)
  double* hz = nullptr;
  int64_t ldz = 0, nrow = 0;
  // device mode: rotations recorded and streamed to the device
  bool device = false;
  double* dZ = nullptr;
  int64_t dldz = 0;
  hipStream_t s = nullptr;
  // two host buffers: the one uploaded last is refilled only after the next
  // flush has synchronized the stream (its asynchronous copy is then done)
  std::vector<EvSweep> sweeps_buf[2];
  std::vector<double> cs_buf[2];
  int cur = 0;
  EvSweep* d_sw = nullptr;
  double* d_cs = nullptr;
  size_t cap_sw = 0, cap_cs = 0;
  int err = RT_OK;
  bool pending = false;

  ~RotSink() {
    if (d_sw) (void)hipFree(d_sw);
    if (d_cs) (void)hipFree(d_cs);
  }

  int flush() {
    std::vector<EvSweep>& sweeps = sweeps_buf[cur];
    std::vector<double>& cs = cs_buf[cur];
    if (!device || sweeps.empty() || err != RT_OK) return err;
    if (pending) {  // the device buffers still feed the previous launch
      if (hipStreamSynchronize(s) != hipSuccess) return err = fail(RT_ERR_HIP, "eigensolver: sync");
      pending = false;
    }
    if (sweeps.size() > cap_sw) {
      if (d_sw) (void)hipFree(d_sw);
      cap_sw = sweeps.size() * 2;
      if (hipMalloc(&d_sw, cap_sw * sizeof(EvSweep)) != hipSuccess)
        return err = fail(RT_ERR_NOMEM, "eigensolver: rotation buffer");
    }
    if (cs.size() > cap_cs) {
      if (d_cs) (void)hipFree(d_cs);
      cap_cs = cs.size() * 2;
      if (hipMalloc(&d_cs, cap_cs * sizeof(double)) != hipSuccess)
        return err = fail(RT_ERR_NOMEM, "eigensolver: rotation buffer");
    }
    if (hipMemcpyAsync(d_sw, sweeps.data(), sweeps.size() * sizeof(EvSweep),
                       hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(d_cs, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice,
                       s) != hipSuccess)
      return err = fail(RT_ERR_HIP, "eigensolver: rotation upload");
    hipLaunchKernelGGL(ev_rot_kernel, dim3((unsigned)((nrow + 63) / 64)), dim3(64), 0, s, dZ,
                       dldz, nrow, d_sw, (int)sweeps.size(), (const double2*)d_cs);
    if (hipGetLastError() != hipSuccess) return err = fail(RT_ERR_HIP, "eigensolver: rotations");
    // the device buffers and these host vectors are reused after the next
    // synchronization; recording goes on in the other host pair
    pending = true;
    cur ^= 1;
    sweeps_buf[cur].clear();
    cs_buf[cur].clear();
    return RT_OK;
  }

  // rotations c[0..len-1], s[0..len-1] on columns l0..l0+len (0-based)
  void sweep(int64_t l0, int len, bool fwd, const double* c, const double* sn) {
    if (len <= 0) return;
    if (!device) {
      if (!hz) return;
      auto rot = [&](int j) {
        const double ct = c[j], st = sn[j];
        if (ct == 1.0 && st == 0.0) return;
        double* z0 = hz + (l0 + j) * ldz;
        double* z1 = z0 + ldz;
        for (int64_t r = 0; r < nrow; ++r) {
          const double temp = z1[r];
          z1[r] = ct * temp - st * z0[r];
          z0[r] = st * temp + ct * z0[r];
        }
      };
      if (fwd)
        for (int j = 0; j < len; ++j) rot(j);
      else
        for (int j = len - 1; j >= 0; --j) rot(j);
      return;
    }
    std::vector<EvSweep>& sweeps = sweeps_buf[cur];
    std::vector<double>& cs = cs_buf[cur];
    EvSweep S;
    S.l = (int32_t)l0;
    S.len = len;
    S.fwd = fwd ? 1 : 0;
    S.pad = 0;
    S.off = (int64_t)cs.size() / 2;
    for (int j = 0; j < len; ++j) {
      cs.push_back(c[j]);
      cs.push_back(sn[j]);
    }
    sweeps.push_back(S);
    if (cs.size() >= ((size_t)1 << 23)) flush();  // 4M rotations (64 MB) per upload
  }
};

static double ev_lapy2(double x, double y) {
  if (std::isnan(x)) return x;
  if (std::isnan(y)) return y;
  const double xa = std::fabs(x), ya = std::fabs(y);
  const double w = std::max(xa, ya), z = std::min(xa, ya);
  if (z == 0.0 || w > DBL_MAX) return w;
  return w * std::sqrt(1.0 + (z / w) * (z / w));
}

// DLARTG (LAPACK 3.10+): c >= 0, r carries the sign of f
static void ev_lartg(double f, double g, double& c, double& s, double& r) {
  const double safmin = DBL_MIN, safmax = 1.0 / DBL_MIN;
  const double rtmin = std::sqrt(safmin), rtmax = std::sqrt(safmax / 2);
  const double f1 = std::fabs(f), g1 = std::fabs(g);
  if (g == 0.0) {
    c = 1.0;
    s = 0.0;
    r = f;
  } else if (f == 0.0) {
    c = 0.0;
    s = std::copysign(1.0, g);
    r = g1;
  } else if (f1 > rtmin && f1 < rtmax && g1 > rtmin && g1 < rtmax) {
    const double d = std::sqrt(f * f + g * g);
    c = f1 / d;
    r = std::copysign(d, f);
    s = g / r;
  } else {
    const double u = std::min(safmax, std::max(safmin, std::max(f1, g1)));
    const double fs = f / u, gs = g / u;
    const double d = std::sqrt(fs * fs + gs * gs);
    c = std::fabs(fs) / d;
    r = std::copysign(d, f);
    s = gs / r;
    r *= u;
  }
}

// DLAEV2: eigen-decomposition of [[a, b], [b, c]]; rt1 the eigenvalue of
// larger magnitude, (cs1, sn1) its unit eigenvector.  want_vec false = DLAE2.
static void ev_laev2(double a, double b, double c, double& rt1, double& rt2, double& cs1,
                     double& sn1, bool want_vec) {
  const double sm = a + c, df = a - c, adf = std::fabs(df), tb = b + b, ab = std::fabs(tb);
  double acmx, acmn;
  if (std::fabs(a) > std::fabs(c)) {
    acmx = a;
    acmn = c;
  } else {
    acmx = c;
    acmn = a;
  }
  double rt;
  if (adf > ab) rt = adf * std::sqrt(1.0 + (ab / adf) * (ab / adf));
  else if (adf < ab) rt = ab * std::sqrt(1.0 + (adf / ab) * (adf / ab));
  else rt = ab * std::sqrt(2.0);
  int sgn1;
  if (sm < 0.0) {
    rt1 = 0.5 * (sm - rt);
    sgn1 = -1;
    rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
  } else if (sm > 0.0) {
    rt1 = 0.5 * (sm + rt);
    sgn1 = 1;
    rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
  } else {
    rt1 = 0.5 * rt;
    rt2 = -0.5 * rt;
    sgn1 = 1;
  }
  if (!want_vec) return;
  double csv;
  int sgn2;
  if (df >= 0.0) {
    csv = df + rt;
    sgn2 = 1;
  } else {
    csv = df - rt;
    sgn2 = -1;
  }
  if (std::fabs(csv) > ab) {
    const double ct = -tb / csv;
    sn1 = 1.0 / std::sqrt(1.0 + ct * ct);
    cs1 = ct * sn1;
  } else if (ab == 0.0) {
    cs1 = 1.0;
    sn1 = 0.0;
  } else {
    const double tn = -csv / tb;
    cs1 = 1.0 / std::sqrt(1.0 + tn * tn);
    sn1 = tn * cs1;
  }
  if (sgn1 == sgn2) {
    const double tn = cs1;
    cs1 = -sn1;
    sn1 = tn;
  }
}

// DSTEQR on d[0..n-1], e[0..n-2] (e needs n entries).  With a sink the
// rotations go to it and `perm` receives the sorting permutation (column j
// of the sorted eigenvectors = column perm[j] of the unsorted ones); without
// one, only eigenvalues (sorted ascending).  Returns the LAPACK info (number
// of off-diagonal entries that did not converge in 30n iterations).
static int ev_steqr(int64_t n, double* d0, double* e0, RotSink* sink, std::vector<int32_t>* perm) {
  if (perm) {
    perm->resize(n);
    for (int64_t j = 0; j < n; ++j) (*perm)[j] = (int32_t)j;
  }
  if (n <= 1) return 0;
  // 1-based views as in the published algorithm
  double* d = d0 - 1;
  double* e = e0 - 1;
  const bool vec = sink != nullptr;
  const double eps = std::ldexp(1.0, -53), eps2 = eps * eps, safmin = DBL_MIN;
  const double safmax = 1.0 / safmin;
  const double ssfmax = std::sqrt(safmax) / 3.0, ssfmin = std::sqrt(safmin) / eps2;
  const int64_t nmaxit = n * 30;
  int64_t jtot = 0;
  std::vector<double> wc(n + 1), ws(n + 1);
  int64_t l1 = 1;
  const int64_t nm1 = n - 1;
  int info = 0;
  for (;;) {
    if (l1 > n) break;
    if (l1 > 1) e[l1 - 1] = 0.0;
    int64_t m = n;
    for (int64_t mm = l1; mm <= nm1; ++mm) {
      const double tst = std::fabs(e[mm]);
      if (tst == 0.0) {
        m = mm;
        break;
      }
      if (tst <= (std::sqrt(std::fabs(d[mm])) * std::sqrt(std::fabs(d[mm + 1]))) * eps) {
        e[mm] = 0.0;
        m = mm;
        break;
      }
    }
    int64_t l = l1;
    const int64_t lsv = l;
    int64_t lend = m;
    const int64_t lendsv = lend;
    l1 = m + 1;
    if (lend == l) continue;
    // scale the block l..lend
    double anorm = 0.0;
    for (int64_t k = l; k <= lend; ++k) anorm = std::max(anorm, std::fabs(d[k]));
    for (int64_t k = l; k < lend; ++k) anorm = std::max(anorm, std::fabs(e[k]));
    int iscale = 0;
    if (anorm == 0.0) continue;
    if (anorm > ssfmax) {
      iscale = 1;
      for (int64_t k = l; k <= lend; ++k) d[k] *= ssfmax / anorm;
      for (int64_t k = l; k < lend; ++k) e[k] *= ssfmax / anorm;
    }
    if (anorm < ssfmin) {
      iscale = 2;
      for (int64_t k = l; k <= lend; ++k) d[k] *= ssfmin / anorm;
      for (int64_t k = l; k < lend; ++k) e[k] *= ssfmin / anorm;
    }
    // QL when the bottom end is the smaller, else QR
    if (std::fabs(d[lend]) < std::fabs(d[l])) {
      lend = lsv;
      l = lendsv;
    }
    if (lend > l) {
      // ---- QL iteration
      for (;;) {
        m = lend;
        if (l != lend) {
          for (int64_t mm = l; mm <= lend - 1; ++mm) {
            const double tst = std::fabs(e[mm]) * std::fabs(e[mm]);
            if (tst <= (eps2 * std::fabs(d[mm])) * std::fabs(d[mm + 1]) + safmin) {
              m = mm;
              break;
            }
          }
        }
        if (m < lend) e[m] = 0.0;
        double p = d[l];
        if (m == l) {  // eigenvalue found
          d[l] = p;
          ++l;
          if (l <= lend) continue;
          break;
        }
        if (m == l + 1) {  // 2 x 2 block
          double rt1, rt2, c = 1.0, s = 0.0;
          ev_laev2(d[l], e[l], d[l + 1], rt1, rt2, c, s, vec);
          if (vec) sink->sweep(l - 1, 1, false, &c, &s);
          d[l] = rt1;
          d[l + 1] = rt2;
          e[l] = 0.0;
          l += 2;
          if (l <= lend) continue;
          break;
        }
        if (jtot == nmaxit) break;
        ++jtot;
        double g = (d[l + 1] - p) / (2.0 * e[l]);
        double r = ev_lapy2(g, 1.0);
        g = d[m] - p + (e[l] / (g + std::copysign(r, g)));
        double s = 1.0, c = 1.0;
        p = 0.0;
        for (int64_t i = m - 1; i >= l; --i) {
          const double f = s * e[i], b = c * e[i];
          ev_lartg(g, f, c, s, r);
          if (i != m - 1) e[i + 1] = r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          if (vec) {
            wc[i] = c;
            ws[i] = -s;
          }
        }
        if (vec) sink->sweep(l - 1, (int)(m - l), false, &wc[l], &ws[l]);
        d[l] -= p;
        e[l] = g;
      }
    } else {
      // ---- QR iteration
      for (;;) {
        m = lend;
        if (l != lend) {
          for (int64_t mm = l; mm >= lend + 1; --mm) {
            const double tst = std::fabs(e[mm - 1]) * std::fabs(e[mm - 1]);
            if (tst <= (eps2 * std::fabs(d[mm])) * std::fabs(d[mm - 1]) + safmin) {
              m = mm;
              break;
            }
          }
        }
        if (m > lend) e[m - 1] = 0.0;
        double p = d[l];
        if (m == l) {
          d[l] = p;
          --l;
          if (l >= lend) continue;
          break;
        }
        if (m == l - 1) {
          double rt1, rt2, c = 1.0, s = 0.0;
          ev_laev2(d[l - 1], e[l - 1], d[l], rt1, rt2, c, s, vec);
          if (vec) sink->sweep(m - 1, 1, true, &c, &s);
          d[l - 1] = rt1;
          d[l] = rt2;
          e[l - 1] = 0.0;
          l -= 2;
          if (l >= lend) continue;
          break;
        }
        if (jtot == nmaxit) break;
        ++jtot;
        double g = (d[l - 1] - p) / (2.0 * e[l - 1]);
        double r = ev_lapy2(g, 1.0);
        g = d[m] - p + (e[l - 1] / (g + std::copysign(r, g)));
        double s = 1.0, c = 1.0;
        p = 0.0;
        for (int64_t i = m; i <= l - 1; ++i) {
          const double f = s * e[i], b = c * e[i];
          ev_lartg(g, f, c, s, r);
          if (i != m) e[i - 1] = r;
          g = d[i] - p;
          r = (d[i + 1] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i] = g + p;
          g = c * r - b;
          if (vec) {
            wc[i] = c;
            ws[i] = s;
          }
        }
        if (vec) sink->sweep(m - 1, (int)(l - m), true, &wc[m], &ws[m]);
        d[l] -= p;
        e[l - 1] = g;
      }
    }
    // undo the scaling of the block
    if (iscale == 1) {
      for (int64_t k = lsv; k <= lendsv; ++k) d[k] *= anorm / ssfmax;
      for (int64_t k = lsv; k < lendsv; ++k) e[k] *= anorm / ssfmax;
    } else if (iscale == 2) {
      for (int64_t k = lsv; k <= lendsv; ++k) d[k] *= anorm / ssfmin;
      for (int64_t k = lsv; k < lendsv; ++k) e[k] *= anorm / ssfmin;
    }
    if (jtot < nmaxit) continue;
    for (int64_t k = 1; k <= n - 1; ++k)
      if (e[k] != 0.0) ++info;
    return info;
  }
  // ascending order: DSTEQR's selection sort (plain sort without vectors)
  if (!vec) {
    std::sort(d0, d0 + n);
    return 0;
  }
  for (int64_t ii = 2; ii <= n; ++ii) {
    const int64_t i = ii - 1;
    int64_t k = i;
    double p = d[i];
    for (int64_t j = ii; j <= n; ++j)
      if (d[j] < p) {
        k = j;
        p = d[j];
      }
    if (k != i) {
      d[k] = d[i];
      d[i] = p;
      if (perm) std::swap((*perm)[i - 1], (*perm)[k - 1]);
      if (!sink->device && sink->hz) {
        for (int64_t r = 0; r < sink->nrow; ++r)
          std::swap(sink->hz[r + (i - 1) * sink->ldz], sink->hz[r + (k - 1) * sink->ldz]);
      }
    }
  }
  return 0;
}

struct EvBufs {
  std::vector<void*> p;
  ~EvBufs() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  int take(size_t count, T** out) {
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(T) * std::max<size_t>(count, 1)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(RT_ERR_NOMEM, "eigensolver: allocation of %zu elements failed", count);
    }
    p.push_back(q);
    *out = (T*)q;
    return RT_OK;
  }
};

// Eigen-decomposition of the symmetric matrix whose lower triangle is
// scale * A (lda).  w_host (n) receives ascending eigenvalues; with Z (n x n,
// ld ldz) non-null, the eigenvectors.  Returns RT status; *info as LAPACK.
int syev_device(int64_t n, const double* A, int64_t lda, double scale, double* w_host,
                double* Z, int64_t ldz, int* info, hipStream_t s) {
  *info = 0;
  if (n == 0) return RT_OK;
  EvBufs b;
  double *W, *vbuf, *y, *wv, *dd, *ee, *tau, *part;
  RTA_TRY(b.take((size_t)n * n, &W));
  RTA_TRY(b.take((size_t)n, &vbuf));
  RTA_TRY(b.take((size_t)n, &y));
  RTA_TRY(b.take((size_t)n, &wv));
  RTA_TRY(b.take((size_t)n, &dd));
  RTA_TRY(b.take((size_t)n, &ee));
  RTA_TRY(b.take((size_t)n, &tau));
  const unsigned nmax_blocks = (unsigned)std::min<int64_t>(1024, ((n * n) + 255) / 256);
  RTA_TRY(b.take((size_t)nmax_blocks, &part));
  // DSYEV's scaling into [sqrt(smlnum), sqrt(bignum)]
  hipLaunchKernelGGL(ev_maxabs_kernel, dim3(nmax_blocks), dim3(256), 0, s, A, lda, n, part);
  RTA_HIP(hipGetLastError());
  std::vector<double> ph(nmax_blocks);
  RTA_HIP(hipMemcpyAsync(ph.data(), part, sizeof(double) * nmax_blocks, hipMemcpyDeviceToHost,
                         s));
  RTA_HIP(hipStreamSynchronize(s));
  double anrm = 0.0;
  for (double v : ph) anrm = std::max(anrm, v);
  anrm *= std::fabs(scale);
  const double smlnum = DBL_MIN / std::ldexp(1.0, -52), bignum = 1.0 / smlnum;
  const double rmin = std::sqrt(smlnum), rmax = std::sqrt(bignum);
  double sigma = 1.0;
  if (anrm > 0.0 && anrm < rmin) sigma = rmin / anrm;
  else if (anrm > rmax) sigma = rmax / anrm;
  hipLaunchKernelGGL(ev_sym_lower_kernel, dim3(ev_grid(n * n)), dim3(256), 0, s, A, lda, n,
                     scale * sigma, W);
  RTA_HIP(hipGetLastError());

  const bool vec = Z != nullptr;
  double* V = nullptr;
  if (vec && n > 1) {
    RTA_TRY(b.take((size_t)n * n, &V));
    hipLaunchKernelGGL(ev_fill_kernel, dim3(ev_grid(n * n)), dim3(256), 0, s, V, n * n, 0.0);
  }
  // ---- tridiagonal reduction
  {
    KernelTimer kt("syev_tridiag", s);
    if (n == 1) {
      hipLaunchKernelGGL(ev_fill_kernel, dim3(1), dim3(256), 0, s, tau, (int64_t)1, 0.0);
      RTA_HIP(hipMemcpyAsync(dd, W, sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    for (int64_t i = 0; i + 1 < n; ++i) {
      const int64_t m = n - 1 - i;
      hipLaunchKernelGGL(ev_house_kernel, dim3(1), dim3(256), 0, s, W, n, i, V, vbuf, dd, ee,
                         tau);
      if (m > 1) {
        hipLaunchKernelGGL(ev_symv_kernel, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, s, W, n,
                           i, vbuf, tau, y);
        hipLaunchKernelGGL(ev_w_kernel, dim3(1), dim3(256), 0, s, n, i, vbuf, tau, y, wv);
        hipLaunchKernelGGL(ev_syr2_kernel, dim3((unsigned)((m + 255) / 256), (unsigned)((m + 7) / 8)),
                           dim3(256), 0, s, W, n, i, vbuf, wv, tau);
      }
    }
    RTA_HIP(hipGetLastError());
  }
  std::vector<double> dh(n), eh(n, 0.0);
  RTA_HIP(hipMemcpyAsync(dh.data(), dd, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  if (n > 1)
    RTA_HIP(hipMemcpyAsync(eh.data(), ee, sizeof(double) * (n - 1), hipMemcpyDeviceToHost, s));
  RTA_HIP(hipStreamSynchronize(s));

  // ---- QL/QR on the tridiagonal; rotations streamed to Z_T = I
  double* ZT = nullptr;
  std::vector<int32_t> perm;
  if (vec) {
    RTA_TRY(b.take((size_t)n * n, &ZT));
    hipLaunchKernelGGL(ev_eye_kernel, dim3(ev_grid(n * n)), dim3(256), 0, s, ZT, n);
    RTA_HIP(hipGetLastError());
    RotSink sink;
    sink.device = true;
    sink.dZ = ZT;
    sink.dldz = n;
    sink.nrow = n;
    sink.s = s;
    {
      KernelTimer kt("syev_rotations", s);
      *info = ev_steqr(n, dh.data(), eh.data(), &sink, &perm);
      RTA_TRY(sink.err);
      RTA_TRY(sink.flush());
      RTA_HIP(hipStreamSynchronize(s));  // before the sink frees its buffers
    }
  } else {
    *info = ev_steqr(n, dh.data(), eh.data(), nullptr, nullptr);
  }
  for (int64_t j = 0; j < n; ++j) w_host[j] = dh[j] / sigma;
  if (!vec) return RT_OK;

  // ---- sort the columns of Z_T, back-transform with the reflectors
  int32_t* dperm;
  RTA_TRY(b.take((size_t)n, &dperm));
  RTA_HIP(hipMemcpyAsync(dperm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(ev_gather_cols_kernel, dim3(ev_grid(n * n)), dim3(256), 0, s, ZT, n, dperm,
                     W, n);  // W is free again: sorted Z_T
  RTA_HIP(hipGetLastError());
  {
    KernelTimer kt("syev_backtransform", s);
    constexpr int NB = 64;
    double *G, *T, *P, *P2;
    RTA_TRY(b.take((size_t)NB * NB, &G));
    RTA_TRY(b.take((size_t)NB * NB, &T));
    RTA_TRY(b.take((size_t)NB * n, &P));
    RTA_TRY(b.take((size_t)NB * n, &P2));
    const int64_t nref = n - 1;  // reflectors 0..n-2
    for (int64_t i0 = ((nref - 1) / NB) * NB; nref > 0 && i0 >= 0; i0 -= NB) {
      const int kb = (int)std::min<int64_t>(NB, nref - i0);
      const int64_t r0 = i0 + 1, mr = n - r0;
      const double* Vb = V + r0 + i0 * n;
      RTA_TRY(rt_dgemm_dev('T', 'N', kb, kb, mr, 1.0, Vb, n, Vb, n, 0.0, G, kb, s));
      hipLaunchKernelGGL(ev_larft_kernel, dim3(1), dim3(64), 0, s, G, kb, tau + i0, T);
      RTA_TRY(rt_dgemm_dev('T', 'N', kb, n, mr, 1.0, Vb, n, W + r0, n, 0.0, P, kb, s));
      RTA_TRY(rt_dgemm_dev('N', 'N', kb, n, kb, 1.0, T, kb, P, kb, 0.0, P2, kb, s));
      RTA_TRY(rt_dgemm_dev('N', 'N', mr, n, kb, -1.0, Vb, n, P2, kb, 1.0, W + r0, n, s));
    }
    RTA_HIP(hipGetLastError());
  }
  RTA_HIP(hipMemcpy2DAsync(Z, sizeof(double) * ldz, W, sizeof(double) * n, sizeof(double) * n,
                           n, hipMemcpyDeviceToDevice, s));
  RTA_HIP(hipStreamSynchronize(s));
  return RT_OK;
}

}  // namespace rta

using namespace rta;

extern "C" int rt_dsteqr_host(char compz, int64_t n, double* d, double* e, double* z,
                              int64_t ldz, int* info) {
  RTA_CHECK(compz == 'N' || compz == 'n' || compz == 'I' || compz == 'i' || compz == 'V' ||
                compz == 'v',
            "compz must be 'N', 'I' or 'V'");
  RTA_CHECK(n >= 0 && info, "bad arguments");
  *info = 0;
  if (n == 0) return RT_OK;
  RTA_CHECK(d && (n == 1 || e), "NULL argument");
  const bool vec = !(compz == 'N' || compz == 'n');
  std::vector<double> eh(n, 0.0);
  for (int64_t k = 0; k + 1 < n; ++k) eh[k] = e[k];
  if (!vec) {
    *info = ev_steqr(n, d, eh.data(), nullptr, nullptr);
  } else {
    RTA_CHECK(z && ldz >= n, "bad z");
    if (compz == 'I' || compz == 'i')
      for (int64_t c = 0; c < n; ++c)
        for (int64_t r = 0; r < n; ++r) z[r + c * ldz] = (r == c) ? 1.0 : 0.0;
    RotSink sink;
    sink.hz = z;
    sink.ldz = ldz;
    sink.nrow = n;
    std::vector<int32_t> perm;
    *info = ev_steqr(n, d, eh.data(), &sink, &perm);
  }
  for (int64_t k = 0; k + 1 < n; ++k) e[k] = eh[k];
  return RT_OK;
}

extern "C" int rt_dsyev_dev(char jobz, int64_t n, double* A, int64_t lda, double* w, int* info,
                            void* stream) {
  RTA_CHECK(jobz == 'N' || jobz == 'n' || jobz == 'V' || jobz == 'v', "jobz must be 'N' or 'V'");
  RTA_CHECK(n >= 0 && lda >= std::max<int64_t>(1, n), "bad dimensions");
  RTA_CHECK(info, "info is NULL");
  *info = 0;
  if (n == 0) return RT_OK;
  RTA_CHECK(A && w, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  const bool vec = (jobz == 'V' || jobz == 'v');
  std::vector<double> wh(n);
  RTA_TRY(syev_device(n, A, lda, 1.0, wh.data(), vec ? A : nullptr, lda, info, s));
  RTA_HIP(hipMemcpyAsync(w, wh.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  RTA_HIP(hipStreamSynchronize(s));
  return RT_OK;
}

extern "C" int rt_matrix_power_dev(int64_t n, const double* A, int64_t lda, double p,
                                   double threshold, double* out, int64_t ldo, void* stream) {
  RTA_CHECK(n >= 0 && lda >= std::max<int64_t>(1, n) && ldo >= std::max<int64_t>(1, n),
            "bad dimensions");
  if (n == 0) return RT_OK;
  RTA_CHECK(A && out, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  KernelTimer kt("matrix_power", s);
  EvBufs b;
  double *E, *f;
  RTA_TRY(b.take((size_t)n * n, &E));
  RTA_TRY(b.take((size_t)n, &f));
  // eigen-decomposition of -A: ascending there = descending for A
  std::vector<double> w(n);
  int info = 0;
  RTA_TRY(syev_device(n, A, lda, -1.0, w.data(), E, n, &info, s));
  if (info != 0)
    return fail(RT_ERR_NUMERIC, "matrix power: eigensolver did not converge (%d)", info);
  // keep eigenvalues >= threshold, scale their vectors by sqrt(lambda)^p
  std::vector<double> fh(n, 0.0);
  for (int64_t j = 0; j < n; ++j) {
    const double lam = -w[j];
    if (lam >= threshold) fh[j] = std::pow(std::sqrt(lam), p);
  }
  RTA_HIP(hipMemcpyAsync(f, fh.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(ev_scale_cols_kernel, dim3(ev_grid(n * n)), dim3(256), 0, s, E, n, f);
  RTA_HIP(hipGetLastError());
  RTA_TRY(rt_dgemm_dev('N', 'T', n, n, n, 1.0, E, n, E, n, 0.0, out, ldo, s));
  RTA_HIP(hipStreamSynchronize(s));
  return RT_OK;
}

// A x = lambda B x for packed-upper A, B (itype 1): the lowest num_orb
// eigenpairs with Z^T B Z = I (_dspgvx / lapack_dspgvx,
// src/matrix/matrix_blas_lapack.rs:2123-2186, 1096-1147, which call LAPACK
// DSPGVX with range 'I', il = 1, iu = num_orb).  B = L L^T, C = L^-1 A L^-T,
// C Y = Y Lambda, Z = L^-T Y.  w[0..num_orb) receives the eigenvalues, the rest
// of w[n] is zero.  info: LAPACK's (n + k when B's leading minor of order k is
// not positive definite).  Synchronizes.
extern "C" int rt_dspgvx_dev(int64_t n, const double* Ap, const double* Bp, int64_t num_orb,
                             double* Z, double* w, int* info, void* stream) {
  RTA_CHECK(info, "info is NULL");
  *info = 0;
  RTA_CHECK(n >= 0 && num_orb >= 0 && num_orb <= n, "need 0 <= num_orb <= n (got %lld, %lld)",
            (long long)num_orb, (long long)n);
  if (n == 0) return RT_OK;
  RTA_CHECK(Ap && Bp && Z && w, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  KernelTimer kt("dspgvx", s);
  EvBufs b;
  double *Af, *Lf, *X, *T1, *C, *Y;
  int* dinfo;
  RTA_TRY(b.take((size_t)n * n, &Af));
  RTA_TRY(b.take((size_t)n * n, &Lf));
  RTA_TRY(b.take((size_t)n * n, &X));
  RTA_TRY(b.take((size_t)n * n, &T1));
  RTA_TRY(b.take((size_t)n * n, &C));
  RTA_TRY(b.take((size_t)n * n, &Y));
  RTA_TRY(b.take((size_t)1, &dinfo));
  const int64_t plen = n * (n + 1) / 2;
  RTA_TRY(rt_upper_unpack_dev(Ap, plen, Af, s));
  RTA_TRY(rt_upper_unpack_dev(Bp, plen, Lf, s));
  int cinfo = 0;
  RTA_TRY(chol_lower(Lf, n, &cinfo, dinfo, s));
  if (cinfo != 0) {
    *info = (int)n + cinfo;
    return RT_OK;
  }
  RTA_TRY(tri_inverse_lower(Lf, n, X, s));                                          // L^-1
  RTA_TRY(rt_dgemm_dev('N', 'N', n, n, n, 1.0, X, n, Af, n, 0.0, T1, n, s));        // L^-1 A
  RTA_TRY(rt_dgemm_dev('N', 'T', n, n, n, 1.0, T1, n, X, n, 0.0, C, n, s));         // .. L^-T
  std::vector<double> wh(n);
  RTA_TRY(syev_device(n, C, n, 1.0, wh.data(), Y, n, info, s));
  if (*info != 0) return RT_OK;
  if (num_orb > 0)
    RTA_TRY(rt_dgemm_dev('T', 'N', n, num_orb, n, 1.0, X, n, Y, n, 0.0, Z, n, s));  // L^-T Y
  for (int64_t j = num_orb; j < n; ++j) wh[j] = 0.0;
  RTA_HIP(hipMemcpyAsync(w, wh.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  RTA_HIP(hipStreamSynchronize(s));
  return RT_OK;
}

// All eigenpairs of a packed-upper symmetric matrix (lapack_dspevx,
// src/matrix/matrix_blas_lapack.rs:1075-1095: DSPEVX with range 'A'):
// Z (n x n) and ascending w[n]; *n_found = n.  Synchronizes.
extern "C" int rt_dspevx_dev(int64_t n, const double* Ap, double* Z, double* w, int* n_found,
                             void* stream) {
  RTA_CHECK(n_found, "n_found is NULL");
  *n_found = 0;
  RTA_CHECK(n >= 0, "negative dimension");
  if (n == 0) return RT_OK;
  RTA_CHECK(Ap && Z && w, "NULL argument");
  hipStream_t s = (hipStream_t)stream;
  EvBufs b;
  double* Af;
  RTA_TRY(b.take((size_t)n * n, &Af));
  RTA_TRY(rt_upper_unpack_dev(Ap, n * (n + 1) / 2, Af, s));
  std::vector<double> wh(n);
  int info = 0;
  RTA_TRY(syev_device(n, Af, n, 1.0, wh.data(), Z, n, &info, s));
  if (info != 0) return fail(RT_ERR_NOTCONV, "dspevx: %d eigenvalues did not converge", info);
  RTA_HIP(hipMemcpyAsync(w, wh.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
  RTA_HIP(hipStreamSynchronize(s));
  *n_found = (int)n;
  return RT_OK;
}

// Packed RHF Fock matrix F = H + J - K/2 (upper triangle, j(j+1)/2 + i) from
// full n x n H, J, K: the input the generalized eigensolve takes.
extern "C" int rt_fock_rhf_packed_dev(int64_t n, const double* H, const double* J,
                                      const double* K, double* F, void* stream) {
  RTA_CHECK(n >= 0, "negative dimension");
  if (n == 0) return RT_OK;
  RTA_CHECK(H && J && K && F, "NULL argument");
  RTA_CHECK(n <= 65535, "n = %lld exceeds the grid's y extent", (long long)n);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(ev_fock_packed_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)n),
                     dim3(256), 0, s, H, J, K, n, F);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

