// Learned Davidson preconditioner: a 2-layer FP32 MLP applied per vector
// component.  For residual r, Ritz vector x0, auxiliary correction a (a
// structured approximation of (A - sigma)^-1 r, or zero), diagonal h and
// shifts e_k (the sample's root) and e_0 (lowest root), with safe(v) = v or
// 1e-8 when |v| < 1e-8 (the reference's clamp, davidson.rs:66):
//   basis  b0 = gk/rms(gk), gk = r/safe(h - e_k)
//          b1 = g0/rms(g0), g0 = r/safe(h - e_0)     (the reference's correction)
//          b2 = x0*sqrt(n),  b3 = r/rms(r),  b4 = a/rms(a)
//   input  f = [asinh b0..b4, tanh((h-e_k)/s), tanh((h-e_0)/s)],  s = 0.25
//   hidden z = tanh(W1 f + c1)          (64 units)
//   output w = W2 z + c2                (5 mixing weights)
//   t      = sum_q w_q b_q
// A fresh network has W2 = 0, c2 = (0,1,0,0,0): t is exactly the reference's
// diagonal correction (up to scale, which Davidson normalizes away).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rta {

constexpr int MLP_F = 7, MLP_H = 64, MLP_O = 5;
constexpr int MLP_NPARAM = MLP_H * MLP_F + MLP_H + MLP_O * MLP_H + MLP_O;  // 837
// parameter offsets
constexpr int MLP_W1 = 0, MLP_C1 = MLP_H * MLP_F, MLP_W2 = MLP_C1 + MLP_H,
              MLP_C2 = MLP_W2 + MLP_O * MLP_H;
constexpr float MLP_SHIFT_SCALE = 0.25f;

struct MlpNet {
  float* params;  // device, MLP_NPARAM
  float* adam_m;  // device
  float* adam_v;  // device
  int64_t step;
};

// T[:, j] = MLP correction for the residual in R[:, j] of root keep[j]
// (x0 = X0[:, keep[j]], auxiliary correction Saux[:, j] or none, shifts
// e[keep[j]] and e[0]).
int mlp_precondition(const MlpNet* net, const double* R, const double* X0,
                     const double* Saux, const int* keep, int nt, const double* hdiag,
                     const double* e_host, int64_t n, double* T,
                     double* scratch, hipStream_t s, int* fallbacks);

}  // namespace rta
