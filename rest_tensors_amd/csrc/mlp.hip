// Learned Davidson preconditioner (A15): inference, loss/gradient and Adam
// training kernels of the 2-layer FP32 MLP described in mlp.hpp.
//
// Training signal: samples recorded from plain Davidson runs (residual r and
// Ritz vector x0 of root k at some cycle, and the subspace Q that cycle's
// corrections are orthogonalized against) paired with that run's converged
// eigenvector x*; the target direction is tau = x* - Q Q^T x* (the part of
// the answer the subspace lacks), normalized.  With p_s = (I - Q Q^T) t_s,
// the correction as Davidson uses it, the loss is
//   L = sum_s w_s (1 - (p_s.tau_s)^2 / (|p_s|^2 |tau_s|^2)) / sum_s w_s,
// the squared-cosine misalignment of the projected correction with tau_s
// (sign-free, scale-free, as Davidson only uses the direction of p).  Its
// gradient with respect to t_s is (I - Q Q^T) applied to the gradient with
// respect to p_s, which equals -2a tau/(b c) + 2a^2 p/(b^2 c) since tau and p
// already lie in the complement (a = p.tau, b = p.p, c = tau.tau).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "dvkernels.hpp"
#include "mlp.hpp"
#include "rng.hpp"
#include "timer.hpp"

namespace rta {

struct MlpSample {
  const double* R;      // residuals, column s at R + s*n
  const double* X;      // Ritz vectors, column xcol[s] at X + xcol[s]*n
  const double* A;      // auxiliary corrections, column s at A + s*n (or NULL)
  const int* xcol;      // device
  const double* ek;     // device, per sample
  const double* e0;     // device, per sample
  const double* stats;  // device, 4 per sample: sum r^2, gk^2, g0^2, a^2
  const double* hdiag;
  int64_t n;
};
constexpr int MLP_NSTAT = 4;

__device__ __forceinline__ double mlp_safe(double h) {
  return fabs(h) < 1e-8 ? 1e-8 : h;
}

__device__ __forceinline__ double mlp_wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// basis b[5] and features f[7] of component x of sample s
__device__ __forceinline__ void mlp_inputs(const MlpSample& S, int s, int64_t x,
                                           double b[MLP_O], float f[MLP_F]) {
  const double r = S.R[x + S.n * s];
  const double x0 = S.X[x + S.n * S.xcol[s]];
  const double a = S.A ? S.A[x + S.n * s] : 0.0;
  const double hd = S.hdiag[x];
  const double hk = hd - S.ek[s], h0 = hd - S.e0[s];
  const double gk = r / mlp_safe(hk), g0 = r / mlp_safe(h0);
  const double nn = (double)S.n;
  const double* st = S.stats + MLP_NSTAT * s;
  const double rr = sqrt(st[0] / nn), rk = sqrt(st[1] / nn), r0 = sqrt(st[2] / nn),
               ra = sqrt(st[3] / nn);
  b[0] = rk > 0 ? gk / rk : 0.0;
  b[1] = r0 > 0 ? g0 / r0 : 0.0;
  b[2] = x0 * sqrt(nn);
  b[3] = rr > 0 ? r / rr : 0.0;
  b[4] = ra > 0 ? a / ra : 0.0;
#pragma unroll
  for (int q = 0; q < 5; ++q) f[q] = asinhf((float)b[q]);
  f[5] = tanhf((float)hk / MLP_SHIFT_SCALE);
  f[6] = tanhf((float)h0 / MLP_SHIFT_SCALE);
}

// hidden activations z and mixing weights w from features (params in LDS)
__device__ __forceinline__ void mlp_core(const float* __restrict__ P,
                                         const float f[MLP_F], float z[MLP_H],
                                         float w[MLP_O]) {
#pragma unroll
  for (int q = 0; q < MLP_O; ++q) w[q] = P[MLP_C2 + q];
  for (int j = 0; j < MLP_H; ++j) {
    float u = P[MLP_C1 + j];
#pragma unroll
    for (int i = 0; i < MLP_F; ++i) u += P[MLP_W1 + j * MLP_F + i] * f[i];
    const float zj = tanhf(u);
    z[j] = zj;
#pragma unroll
    for (int q = 0; q < MLP_O; ++q) w[q] += P[MLP_W2 + q * MLP_H + j] * zj;
  }
}

// stats: sum r^2, gk^2, g0^2, a^2 per sample (partials over x slices)
__global__ __launch_bounds__(256) void mlp_stats_kernel(MlpSample S,
                                                        int64_t slice,
                                                        double* __restrict__ part) {
  __shared__ double red[MLP_NSTAT][4];
  const int s = blockIdx.y;
  const int64_t x0 = (int64_t)blockIdx.x * slice;
  const int64_t x1 = (x0 + slice < S.n) ? x0 + slice : S.n;
  double acc[MLP_NSTAT] = {0, 0, 0, 0};
  for (int64_t x = x0 + threadIdx.x; x < x1; x += 256) {
    const double r = S.R[x + S.n * s];
    const double hd = S.hdiag[x];
    const double gk = r / mlp_safe(hd - S.ek[s]), g0 = r / mlp_safe(hd - S.e0[s]);
    const double a = S.A ? S.A[x + S.n * s] : 0.0;
    acc[0] += r * r;
    acc[1] += gk * gk;
    acc[2] += g0 * g0;
    acc[3] += a * a;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < MLP_NSTAT; ++q) {
    const double v = mlp_wsum(acc[q]);
    if (lane == 0) red[q][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < MLP_NSTAT) {
    const int q = threadIdx.x;
    part[((int64_t)s * gridDim.x + blockIdx.x) * MLP_NSTAT + q] =
        (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
  }
}

__global__ void mlp_stats_reduce(const double* __restrict__ part, int nsplit,
                                 int ns, double* __restrict__ stats) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ns * MLP_NSTAT) return;
  const int s = e / MLP_NSTAT, q = e % MLP_NSTAT;
  double acc = 0.0;
  for (int i = 0; i < nsplit; ++i)
    acc += part[((int64_t)s * nsplit + i) * MLP_NSTAT + q];
  stats[e] = acc;
}

__global__ __launch_bounds__(256) void mlp_forward_kernel(
    MlpSample S, const float* __restrict__ params, double* __restrict__ T) {
  __shared__ float P[MLP_NPARAM];
  for (int i = threadIdx.x; i < MLP_NPARAM; i += 256) P[i] = params[i];
  __syncthreads();
  const int s = blockIdx.y;
  const int64_t x = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (x >= S.n) return;
  double b[MLP_O];
  float f[MLP_F], z[MLP_H], w[MLP_O];
  mlp_inputs(S, s, x, b, f);
  mlp_core(P, f, z, w);
  double t = 0.0;
#pragma unroll
  for (int q = 0; q < MLP_O; ++q) t += (double)w[q] * b[q];
  T[x + S.n * s] = t;
}

// partial dots per sample: t.tau, t.t, tau.tau
__global__ __launch_bounds__(256) void mlp_dots_kernel(
    const double* __restrict__ T, const double* __restrict__ tau, int64_t n,
    int64_t slice, double* __restrict__ part) {
  __shared__ double red[3][4];
  const int s = blockIdx.y;
  const int64_t x0 = (int64_t)blockIdx.x * slice;
  const int64_t x1 = (x0 + slice < n) ? x0 + slice : n;
  double a = 0, b = 0, c = 0;
  for (int64_t x = x0 + threadIdx.x; x < x1; x += 256) {
    const double t = T[x + n * s], u = tau[x + n * s];
    a += t * u;
    b += t * t;
    c += u * u;
  }
  a = mlp_wsum(a);
  b = mlp_wsum(b);
  c = mlp_wsum(c);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = a;
    red[1][wave] = b;
    red[2][wave] = c;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int q = threadIdx.x;
    part[((int64_t)s * gridDim.x + blockIdx.x) * 3 + q] =
        (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
  }
}

// Per-block partial gradients.  coef[s*2 + {0,1}] = (-2 a w/(W b c), 2 a^2 w/(W b^2 c))
// so that dL/dt = coef0 * tau + coef1 * p, p = TP (the projected correction)
// or, without TP, the network output t.  128 components per block.
constexpr int MLP_BX = 128;
__global__ __launch_bounds__(MLP_BX) void mlp_backward_kernel(
    MlpSample S, const float* __restrict__ params,
    const double* __restrict__ tau, const double* __restrict__ coef,
    const double* __restrict__ TP, float* __restrict__ gpart) {
  __shared__ float P[MLP_NPARAM];
  __shared__ float sf[MLP_BX][MLP_F];
  __shared__ float sz[MLP_BX][MLP_H + 1];
  __shared__ float sdw[MLP_BX][MLP_O];
  __shared__ float sdu[MLP_BX][MLP_H + 1];
  for (int i = threadIdx.x; i < MLP_NPARAM; i += MLP_BX) P[i] = params[i];
  __syncthreads();
  const int s = blockIdx.y;
  const int64_t x = blockIdx.x * (int64_t)MLP_BX + threadIdx.x;
  const int tx = threadIdx.x;
  if (x < S.n) {
    double b[MLP_O];
    float f[MLP_F], z[MLP_H], w[MLP_O];
    mlp_inputs(S, s, x, b, f);
    mlp_core(P, f, z, w);
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < MLP_O; ++q) t += (double)w[q] * b[q];
    if (TP) t = TP[x + S.n * s];  // the projected correction
    const double gt = coef[2 * s] * tau[x + S.n * s] + coef[2 * s + 1] * t;
    float dw[MLP_O];
#pragma unroll
    for (int q = 0; q < MLP_O; ++q) dw[q] = (float)(gt * b[q]);
#pragma unroll
    for (int i = 0; i < MLP_F; ++i) sf[tx][i] = f[i];
#pragma unroll
    for (int q = 0; q < MLP_O; ++q) sdw[tx][q] = dw[q];
    for (int j = 0; j < MLP_H; ++j) {
      float dz = 0.f;
#pragma unroll
      for (int q = 0; q < MLP_O; ++q) dz += P[MLP_W2 + q * MLP_H + j] * dw[q];
      sz[tx][j] = z[j];
      sdu[tx][j] = dz * (1.f - z[j] * z[j]);
    }
  } else {
#pragma unroll
    for (int i = 0; i < MLP_F; ++i) sf[tx][i] = 0.f;
#pragma unroll
    for (int q = 0; q < MLP_O; ++q) sdw[tx][q] = 0.f;
    for (int j = 0; j < MLP_H; ++j) {
      sz[tx][j] = 0.f;
      sdu[tx][j] = 0.f;
    }
  }
  __syncthreads();
  float* out = gpart + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * MLP_NPARAM;
  for (int p = tx; p < MLP_NPARAM; p += MLP_BX) {
    float acc = 0.f;
    if (p < MLP_C1) {  // W1[j][i] += du_j f_i
      const int j = p / MLP_F, i = p % MLP_F;
      for (int c = 0; c < MLP_BX; ++c) acc += sdu[c][j] * sf[c][i];
    } else if (p < MLP_W2) {  // c1[j] += du_j
      const int j = p - MLP_C1;
      for (int c = 0; c < MLP_BX; ++c) acc += sdu[c][j];
    } else if (p < MLP_C2) {  // W2[q][j] += dw_q z_j
      const int q = (p - MLP_W2) / MLP_H, j = (p - MLP_W2) % MLP_H;
      for (int c = 0; c < MLP_BX; ++c) acc += sdw[c][q] * sz[c][j];
    } else {  // c2[q] += dw_q
      const int q = p - MLP_C2;
      for (int c = 0; c < MLP_BX; ++c) acc += sdw[c][q];
    }
    out[p] = acc;
  }
}

__global__ void mlp_grad_reduce(const float* __restrict__ gpart, int64_t nblk,
                                double* __restrict__ grad) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  double acc = 0.0;
  for (int64_t b = 0; b < nblk; ++b) acc += gpart[b * MLP_NPARAM + p];
  grad[p] = acc;
}

__global__ void mlp_adam_kernel(float* __restrict__ params, float* __restrict__ m,
                                float* __restrict__ v,
                                const double* __restrict__ grad, float lr,
                                float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  const float g = (float)grad[p];
  const float mp = b1 * m[p] + (1.f - b1) * g;
  const float vp = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mp;
  v[p] = vp;
  params[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);
}

// tau[:, s] = normalize(x* - (x*.x0) x0), x* = Xfinal[:, root[s]];
// part receives per-sample (x*.x0, |tau|^2) through two passes
__global__ __launch_bounds__(256) void mlp_overlap_kernel(
    const double* __restrict__ X, const int* __restrict__ root,
    const double* __restrict__ Xf, int64_t n, double* __restrict__ ov) {
  __shared__ double red[4];
  const int s = blockIdx.x;
  const double* x0 = X + n * s;
  const double* xf = Xf + n * root[s];
  double a = 0;
  for (int64_t x = threadIdx.x; x < n; x += 256) a += x0[x] * xf[x];
  a = mlp_wsum(a);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) ov[s] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void mlp_target_kernel(
    const double* __restrict__ X, const int* __restrict__ root,
    const double* __restrict__ Xf, const double* __restrict__ ov, int64_t n,
    double* __restrict__ tau, double* __restrict__ sq) {
  __shared__ double red[4];
  const int s = blockIdx.x;
  const double* x0 = X + n * s;
  const double* xf = Xf + n * root[s];
  const double o = ov[s];
  double a = 0;
  for (int64_t x = threadIdx.x; x < n; x += 256) {
    const double t = xf[x] - o * x0[x];
    tau[x + n * s] = t;
    a += t * t;
  }
  a = mlp_wsum(a);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) sq[s] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void mlp_scale_kernel(double* __restrict__ tau, const double* __restrict__ sc,
                                 int64_t n) {
  const int64_t x = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (x < n) tau[x + n * blockIdx.y] *= sc[blockIdx.y];
}

// ---------------------------------------------------------------------------
static int nsplit_for(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(64, (n + 4095) / 4096));
}

// Device workspace for ns samples: xcol (ns ints), ek, e0 (ns), stats
// (MLP_NSTAT ns), partials.  Lives in scratch slots (grow on demand).
// t (n) -= Q (Q^T t), twice, Q the qc orthonormal columns of a recorded
// subspace (the block projection of the Davidson solve)
static int project_one(const double* Q, int qc, double* t, int64_t n, hipStream_t s) {
  double* G = nullptr;
  RTA_TRY(scratch_get(SCRATCH_DV_PROJ, sizeof(double) * (qc + dv_gram_scratch(n, qc, 1)),
                      (void**)&G));
  return dv_project_out_block(Q, n, qc, t, 1, n, G, G + qc, s);
}

struct MlpWork {
  int* xcol;
  double* ek;
  double* e0;
  double* stats;
  double* part;
};

static int mlp_work(int ns, int64_t n, MlpWork* w) {
  const int nsp = nsplit_for(n);
  const size_t bytes = sizeof(int) * (ns + 2) + sizeof(double) * (2 * ns + MLP_NSTAT * ns) +
                       sizeof(double) * MLP_NSTAT * (size_t)ns * nsp + 256;
  char* p = nullptr;
  RTA_TRY(scratch_get(SCRATCH_MLP_A, bytes, (void**)&p));
  size_t off = 0;
  auto take = [&](size_t b) {
    char* q = p + off;
    off += (b + 15) & ~size_t(15);
    return q;
  };
  w->xcol = (int*)take(sizeof(int) * ns);
  w->ek = (double*)take(sizeof(double) * ns);
  w->e0 = (double*)take(sizeof(double) * ns);
  w->stats = (double*)take(sizeof(double) * MLP_NSTAT * ns);
  w->part = (double*)take(sizeof(double) * MLP_NSTAT * (size_t)ns * nsp);
  return RT_OK;
}

static int mlp_prepare(const MlpWork& w, MlpSample* S, const double* R,
                       const double* X, const double* A, const double* hdiag,
                       int64_t n, int ns, const int* xcol_h, const double* ek_h,
                       const double* e0_h, hipStream_t s) {
  RTA_HIP(hipMemcpyAsync(w.xcol, xcol_h, sizeof(int) * ns, hipMemcpyHostToDevice, s));
  RTA_HIP(hipMemcpyAsync(w.ek, ek_h, sizeof(double) * ns, hipMemcpyHostToDevice, s));
  RTA_HIP(hipMemcpyAsync(w.e0, e0_h, sizeof(double) * ns, hipMemcpyHostToDevice, s));
  *S = MlpSample{R, X, A, w.xcol, w.ek, w.e0, w.stats, hdiag, n};
  const int nsp = nsplit_for(n);
  const int64_t slice = (n + nsp - 1) / nsp;
  hipLaunchKernelGGL(mlp_stats_kernel, dim3(nsp, ns), dim3(256), 0, s, *S, slice,
                     w.part);
  hipLaunchKernelGGL(mlp_stats_reduce, dim3((MLP_NSTAT * ns + 255) / 256), dim3(256), 0,
                     s, w.part, nsp, ns, w.stats);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

int mlp_precondition(const MlpNet* net, const double* R, const double* X0,
                     const double* Saux, const int* keep, int nt, const double* hdiag,
                     const double* e_host, int64_t n, double* T,
                     double* /*scratch*/, hipStream_t s, int* /*fallbacks*/) {
  if (nt <= 0) return RT_OK;
  MlpWork w;
  RTA_TRY(mlp_work(nt, n, &w));
  std::vector<double> ek(nt), e0(nt, e_host[0]);
  for (int j = 0; j < nt; ++j) ek[j] = e_host[keep[j]];
  MlpSample S;
  RTA_TRY(mlp_prepare(w, &S, R, X0, Saux, hdiag, n, nt, keep, ek.data(), e0.data(), s));
  KernelTimer kt("mlp_forward", s);
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + 255) / 256), nt),
                     dim3(256), 0, s, S, net->params, T);
  RTA_HIP(hipGetLastError());
  // the host copies above read pageable memory: finish before it goes away
  RTA_HIP(hipStreamSynchronize(s));
  return RT_OK;
}

// Loss and gradient over samples [0, ns) (columns of R/X/A/tau), FP64
// result.  Qs[j], qc[j]: sample j's subspace (NULL / 0: no projection).
static int mlp_loss_grad(const MlpNet* net, const double* R, const double* X,
                         const double* A, const double* hdiag, const double* ek_h,
                         const double* e0_h, const double* tau, const double* wts_h,
                         int64_t n, int ns, const double* const* Qs, const int* qc,
                         double* loss, double* grad_dev, double* T, hipStream_t s) {
  MlpWork w;
  RTA_TRY(mlp_work(ns, n, &w));
  std::vector<int> xcol(ns);
  for (int j = 0; j < ns; ++j) xcol[j] = j;
  MlpSample S;
  RTA_TRY(mlp_prepare(w, &S, R, X, A, hdiag, n, ns, xcol.data(), ek_h, e0_h, s));
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + 255) / 256), ns),
                     dim3(256), 0, s, S, net->params, T);
  bool projected = false;
  if (Qs)
    for (int j = 0; j < ns; ++j)
      if (Qs[j] && qc[j] > 0) {
        RTA_TRY(project_one(Qs[j], qc[j], T + (size_t)j * n, n, s));
        projected = true;
      }
  const int nsp = nsplit_for(n);
  const int64_t slice = (n + nsp - 1) / nsp;
  hipLaunchKernelGGL(mlp_dots_kernel, dim3(nsp, ns), dim3(256), 0, s, T, tau, n,
                     slice, w.part);
  RTA_HIP(hipGetLastError());
  std::vector<double> part((size_t)3 * ns * nsp);
  RTA_HIP(hipMemcpyAsync(part.data(), w.part, sizeof(double) * part.size(),
                         hipMemcpyDeviceToHost, s));
  RTA_HIP(hipStreamSynchronize(s));
  double W = 0.0;
  for (int j = 0; j < ns; ++j) W += wts_h[j];
  std::vector<double> coef(2 * (size_t)ns, 0.0);
  double L = 0.0;
  for (int j = 0; j < ns; ++j) {
    double a = 0, b = 0, c = 0;
    for (int i = 0; i < nsp; ++i) {
      a += part[((size_t)j * nsp + i) * 3 + 0];
      b += part[((size_t)j * nsp + i) * 3 + 1];
      c += part[((size_t)j * nsp + i) * 3 + 2];
    }
    if (wts_h[j] <= 0.0 || W <= 0.0 || b <= 0.0 || c <= 0.0) continue;
    const double wj = wts_h[j] / W;
    L += wj * (1.0 - a * a / (b * c));
    coef[2 * j] = -2.0 * a * wj / (b * c);
    coef[2 * j + 1] = 2.0 * a * a * wj / (b * b * c);
  }
  *loss = L;
  double* dcoef = nullptr;
  const int64_t nbx = (n + MLP_BX - 1) / MLP_BX;
  const size_t gbytes = sizeof(float) * (size_t)MLP_NPARAM * nbx * ns;
  float* gpart = nullptr;
  RTA_TRY(scratch_get(SCRATCH_MLP_B, gbytes + sizeof(double) * 2 * ns + 64,
                      (void**)&gpart));
  dcoef = (double*)((char*)gpart + ((gbytes + 15) & ~size_t(15)));
  RTA_HIP(hipMemcpyAsync(dcoef, coef.data(), sizeof(double) * 2 * ns,
                         hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(mlp_backward_kernel, dim3((unsigned)nbx, ns), dim3(MLP_BX), 0, s,
                     S, net->params, tau, dcoef, projected ? T : nullptr, gpart);
  hipLaunchKernelGGL(mlp_grad_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0,
                     s, gpart, nbx * ns, grad_dev);
  RTA_HIP(hipGetLastError());
  RTA_HIP(hipStreamSynchronize(s));  // host-side coef must outlive the copy
  return RT_OK;
}

// per-sample subspace pointers from (Q, q_off, q_cnt)
static void mlp_subspaces(const double* Q, const int64_t* q_off, const int* q_cnt,
                          int ns, std::vector<const double*>& Qs, std::vector<int>& qc) {
  Qs.assign(ns, nullptr);
  qc.assign(ns, 0);
  if (!Q || !q_off || !q_cnt) return;
  for (int j = 0; j < ns; ++j)
    if (q_off[j] >= 0 && q_cnt[j] > 0) {
      Qs[j] = Q + q_off[j];
      qc[j] = q_cnt[j];
    }
}

}  // namespace rta

using namespace rta;

extern "C" int rt_mlp_num_params(void) { return MLP_NPARAM; }

extern "C" int rt_mlp_create(uint64_t seed, void** out) {
  RTA_CHECK(out, "NULL output");
  std::vector<float> p(MLP_NPARAM, 0.f);
  const uint64_t key = rng_key(seed);
  for (int i = 0; i < MLP_H * MLP_F; ++i) {
    // uniform(-a, a), a = sqrt(3/F): unit-variance pre-activations for unit inputs
    const uint64_t h = splitmix64(key + (uint64_t)i);
    const double u = (double)(h >> 11) * 0x1.0p-53;
    p[MLP_W1 + i] = (float)((2.0 * u - 1.0) * std::sqrt(3.0 / MLP_F));
  }
  p[MLP_C2 + 1] = 1.0f;  // start as the reference preconditioner
  MlpNet* net = new MlpNet{};
  hipError_t e1 = hipMalloc(&net->params, sizeof(float) * MLP_NPARAM);
  hipError_t e2 = hipMalloc(&net->adam_m, sizeof(float) * MLP_NPARAM);
  hipError_t e3 = hipMalloc(&net->adam_v, sizeof(float) * MLP_NPARAM);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
    (void)hipGetLastError();
    delete net;
    return fail(RT_ERR_NOMEM, "MLP allocation failed");
  }
  RTA_HIP(hipMemcpy(net->params, p.data(), sizeof(float) * MLP_NPARAM,
                    hipMemcpyHostToDevice));
  RTA_HIP(hipMemset(net->adam_m, 0, sizeof(float) * MLP_NPARAM));
  RTA_HIP(hipMemset(net->adam_v, 0, sizeof(float) * MLP_NPARAM));
  net->step = 0;
  *out = net;
  return RT_OK;
}

extern "C" int rt_mlp_destroy(void* net) {
  if (!net) return RT_OK;
  MlpNet* m = (MlpNet*)net;
  (void)hipFree(m->params);
  (void)hipFree(m->adam_m);
  (void)hipFree(m->adam_v);
  delete m;
  return RT_OK;
}

extern "C" int rt_mlp_get_params(void* net, float* out) {
  RTA_CHECK(net && out, "NULL argument");
  RTA_HIP(hipMemcpy(out, ((MlpNet*)net)->params, sizeof(float) * MLP_NPARAM,
                    hipMemcpyDeviceToHost));
  return RT_OK;
}

extern "C" int rt_mlp_set_params(void* net, const float* in) {
  RTA_CHECK(net && in, "NULL argument");
  RTA_HIP(hipMemcpy(((MlpNet*)net)->params, in, sizeof(float) * MLP_NPARAM,
                    hipMemcpyHostToDevice));
  return RT_OK;
}

extern "C" int rt_mlp_apply_dev(void* net, const double* R, const double* X,
                                const double* A, const double* hdiag, const double* ek,
                                const double* e0, int64_t n, int ns, double* T,
                                void* stream) {
  RTA_CHECK(net && R && X && hdiag && ek && e0 && T, "NULL argument");
  RTA_CHECK(n > 0 && ns > 0, "empty input");
  hipStream_t s = (hipStream_t)stream;
  MlpWork w;
  RTA_TRY(mlp_work(ns, n, &w));
  std::vector<int> xcol(ns);
  for (int j = 0; j < ns; ++j) xcol[j] = j;
  MlpSample S;
  RTA_TRY(mlp_prepare(w, &S, R, X, A, hdiag, n, ns, xcol.data(), ek, e0, s));
  hipLaunchKernelGGL(mlp_forward_kernel, dim3((unsigned)((n + 255) / 256), ns),
                     dim3(256), 0, s, S, ((MlpNet*)net)->params, T);
  RTA_HIP(hipGetLastError());
  RTA_HIP(hipStreamSynchronize(s));
  return RT_OK;
}

extern "C" int rt_mlp_loss_grad_dev(void* net, const double* R, const double* X,
                                    const double* A, const double* hdiag,
                                    const double* ek, const double* e0,
                                    const double* tau, const double* wts, int64_t n,
                                    int ns, const double* Q, const int64_t* q_off,
                                    const int* q_cnt, double* loss, double* grad,
                                    void* stream) {
  RTA_CHECK(net && R && X && hdiag && ek && e0 && tau && wts && loss && grad,
            "NULL argument");
  RTA_CHECK(n > 0 && ns > 0, "empty input");
  hipStream_t s = (hipStream_t)stream;
  std::vector<const double*> Qs;
  std::vector<int> qc;
  mlp_subspaces(Q, q_off, q_cnt, ns, Qs, qc);
  double *T = nullptr, *g = nullptr;
  RTA_HIP(hipMalloc(&T, sizeof(double) * n * ns));
  RTA_HIP(hipMalloc(&g, sizeof(double) * MLP_NPARAM));
  int st = mlp_loss_grad((MlpNet*)net, R, X, A, hdiag, ek, e0, tau, wts, n, ns,
                         Qs.data(), qc.data(), loss, g, T, s);
  if (st == RT_OK) {
    hipError_t e = hipMemcpy(grad, g, sizeof(double) * MLP_NPARAM,
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) st = fail(RT_ERR_HIP, "gradient copy failed");
  }
  (void)hipFree(T);
  (void)hipFree(g);
  return st;
}

extern "C" int rt_mlp_train_dev(void* net_, const double* R, const double* X,
                                const double* A, const double* hdiag,
                                const double* ek, const double* e0,
                                const double* tau, const double* wts, int64_t n,
                                int ns, const double* Q, const int64_t* q_off,
                                const int* q_cnt, int steps, int batch, double lr,
                                uint64_t seed, double* loss_hist, void* stream) {
  RTA_CHECK(net_ && R && X && hdiag && ek && e0 && tau && wts, "NULL argument");
  RTA_CHECK(n > 0 && ns > 0 && steps >= 0 && batch > 0, "bad training sizes");
  MlpNet* net = (MlpNet*)net_;
  hipStream_t s = (hipStream_t)stream;
  batch = std::min(batch, ns);
  std::vector<const double*> Qs_all, Qs(batch);
  std::vector<int> qc_all, qc(batch);
  mlp_subspaces(Q, q_off, q_cnt, ns, Qs_all, qc_all);
  double *Rb = nullptr, *Xb = nullptr, *Ab = nullptr, *Tb = nullptr, *taub = nullptr,
         *g = nullptr;
  auto dmalloc = [](double** p, size_t cnt) {
    if (hipMalloc(p, sizeof(double) * cnt) == hipSuccess) return true;
    (void)hipGetLastError();
    *p = nullptr;
    return false;
  };
  const size_t nb = (size_t)n * batch;
  int st = RT_OK;
  if (!dmalloc(&Rb, nb) || !dmalloc(&Xb, nb) || (A && !dmalloc(&Ab, nb)) || !dmalloc(&Tb, nb) ||
      !dmalloc(&taub, nb) || !dmalloc(&g, MLP_NPARAM))
    st = fail(RT_ERR_NOMEM, "MLP training: device allocation failed");
  std::vector<double> ekb(batch), e0b(batch), wb(batch);
  const uint64_t key = rng_key(seed);
  const float b1 = 0.9f, b2 = 0.999f;
  for (int it = 0; it < steps && st == RT_OK; ++it) {
    for (int j = 0; j < batch && st == RT_OK; ++j) {
      const int q = (int)(splitmix64(key + (uint64_t)it * batch + j) % (uint64_t)ns);
      const size_t bytes = sizeof(double) * n;
      hipError_t e = hipMemcpyAsync(Rb + n * j, R + n * q, bytes, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess)
        e = hipMemcpyAsync(Xb + n * j, X + n * q, bytes, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess && A)
        e = hipMemcpyAsync(Ab + n * j, A + n * q, bytes, hipMemcpyDeviceToDevice, s);
      if (e == hipSuccess)
        e = hipMemcpyAsync(taub + n * j, tau + n * q, bytes, hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess) st = fail(RT_ERR_HIP, "MLP training: batch gather failed");
      ekb[j] = ek[q];
      e0b[j] = e0[q];
      wb[j] = wts[q];
      Qs[j] = Qs_all[q];
      qc[j] = qc_all[q];
    }
    if (st != RT_OK) break;
    double L = 0.0;
    st = mlp_loss_grad(net, Rb, Xb, Ab, hdiag, ekb.data(), e0b.data(), taub, wb.data(),
                       n, batch, Qs.data(), qc.data(), &L, g, Tb, s);
    if (st != RT_OK) break;
    if (loss_hist) loss_hist[it] = L;
    net->step += 1;
    const float c1 = 1.f - std::pow(b1, (float)net->step);
    const float c2 = 1.f - std::pow(b2, (float)net->step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0,
                       s, net->params, net->adam_m, net->adam_v, g, (float)lr, b1, b2,
                       1e-8f, c1, c2);
    if (hipGetLastError() != hipSuccess) st = fail(RT_ERR_HIP, "adam launch failed");
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(Rb);
  (void)hipFree(Xb);
  if (Ab) (void)hipFree(Ab);
  (void)hipFree(Tb);
  (void)hipFree(taub);
  (void)hipFree(g);
  return st;
}

extern "C" int rt_mlp_targets_dev(const double* X, const int* root, int ns,
                                  const double* Xfinal, int64_t n,
                                  double min_overlap, const double* Q,
                                  const int64_t* q_off, const int* q_cnt, double* tau,
                                  double* wts, void* stream) {
  RTA_CHECK(X && root && Xfinal && tau && wts, "NULL argument");
  RTA_CHECK(n > 0 && ns > 0, "empty input");
  hipStream_t s = (hipStream_t)stream;
  std::vector<const double*> Qs;
  std::vector<int> qc;
  mlp_subspaces(Q, q_off, q_cnt, ns, Qs, qc);
  int* droot;
  double *ov, *sq;
  RTA_HIP(hipMalloc(&droot, sizeof(int) * ns));
  RTA_HIP(hipMalloc(&ov, sizeof(double) * ns));
  RTA_HIP(hipMalloc(&sq, sizeof(double) * ns * 257));  // norms + dv_sqnorms partials
  RTA_HIP(hipMemcpy(droot, root, sizeof(int) * ns, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(mlp_overlap_kernel, dim3(ns), dim3(256), 0, s, X, droot, Xfinal,
                     n, ov);
  hipLaunchKernelGGL(mlp_target_kernel, dim3(ns), dim3(256), 0, s, X, droot, Xfinal,
                     ov, n, tau, sq);
  std::vector<double> hov(ns), hsq(ns), sc(ns);
  RTA_HIP(hipMemcpy(hov.data(), ov, sizeof(double) * ns, hipMemcpyDeviceToHost));
  if (Q) {
    // tau = x* - Q Q^T x* (x* copied in, then projected; norms recomputed)
    for (int j = 0; j < ns; ++j)
      RTA_HIP(hipMemcpyAsync(tau + (size_t)j * n, Xfinal + (size_t)root[j] * n,
                             sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    for (int j = 0; j < ns; ++j)
      if (Qs[j]) RTA_TRY(project_one(Qs[j], qc[j], tau + (size_t)j * n, n, s));
    RTA_TRY(dv_sqnorms(tau, ns, n, sq, sq + ns, s));
  }
  RTA_HIP(hipMemcpy(hsq.data(), sq, sizeof(double) * ns, hipMemcpyDeviceToHost));
  for (int j = 0; j < ns; ++j) {
    const bool ok = std::fabs(hov[j]) >= min_overlap && hsq[j] > 1e-24 &&
                    (!Q || Qs[j] != nullptr);
    wts[j] = ok ? 1.0 : 0.0;
    sc[j] = ok ? 1.0 / std::sqrt(hsq[j]) : 0.0;
  }
  RTA_HIP(hipMemcpy(ov, sc.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(mlp_scale_kernel, dim3((unsigned)((n + 255) / 256), ns),
                     dim3(256), 0, s, tau, ov, n);
  RTA_HIP(hipGetLastError());
  RTA_HIP(hipStreamSynchronize(s));
  (void)hipFree(droot);
  (void)hipFree(ov);
  (void)hipFree(sq);
  return RT_OK;
}
