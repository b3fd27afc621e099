// Device building blocks of the Davidson solver and the RI linear-response
// operator (see dvkernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rta {

// A = diag(d) + 2 scale^2 B B^T, B an n x naux column block (leading dim ldb).
struct CasidaOp {
  const double* B;
  int64_t ldb;
  int64_t n;
  int64_t naux;
  double scale;
  const double* d;  // device, length n
};

int casida_apply(const CasidaOp* op, const double* X, double* Y, int64_t n,
                 int64_t k, hipStream_t s);
int casida_hdiag(const CasidaOp* op, double* hdiag, hipStream_t s);

// Shift-invert preconditioner of the Casida operator: T = (A - sigma)^-1 R by
// the Woodbury identity with A = D + c B B^T, c = 2 scale^2, Dm = 1/(d - sigma)
// (sigma below min d):
//   (A - sigma)^-1 = Dm - Dm B (I/c + B^T Dm B)^-1 B^T Dm.
// Ci = (I/c + B^T Dm B)^-1 (naux x naux) is formed once: the FP64 GEMM for
// B^T Dm B, then the inverse through its Cholesky factor.
struct CasidaSI {
  CasidaOp op;
  double sigma;
  double* Dm;   // device, n
  double* Ci;   // device, naux x naux
  double inv_err; // |I - Cap Ci|_F
};

int casida_si_setup(CasidaSI* si, double shift, hipStream_t s);
int casida_si_apply(const CasidaSI* si, const double* R, double* T, int64_t n,
                    int64_t k, hipStream_t s);

// G[i + ldg*j] = U[:,i] . W[:,j], i < nu, j < nw; scratch >= dv_gram_scratch
int dv_gram(const double* U, int64_t ldu, int nu, const double* W, int64_t ldw,
            int nw, int64_t n, double* G, int64_t ldg, double* scratch,
            hipStream_t s);
int64_t dv_gram_scratch(int64_t n, int nu, int nw);
// Y[:,c] = sum_s V[:,s] coef[s + ldc*c]
int dv_combine(const double* V, int64_t ldv, int ns, const double* coef,
               int ldc, int nc, int64_t n, double* Y, int64_t ldy,
               hipStream_t s);
// X0[:,k] = XS C[:,k]; R[:,k] = AX C[:,k] - e[k] X0[:,k]; sq_out[k] = |R[:,k]|^2
// (device), in one pass over the subspace blocks; scratch >= 4 nk doubles per
// 1024-row slice (dv_gram_scratch(n, 4, nk))
int dv_ritz_residuals(const double* XS, const double* AX, int64_t ld, int ns,
                      const double* coef, int ldc, int nk, const double* e_dev, int64_t n,
                      double* X0, double* R, double* sq_out, double* scratch, hipStream_t s);
int dv_sqnorms(const double* X, int nk, int64_t n, double* sq_out,
               double* scratch, hipStream_t s);
int dv_precond_diag(const double* R, const double* hdiag, const double* shift,
                    int nk, int64_t n, double* T, hipStream_t s);
// X[:, k] /= |X[:, k]| for k < nk, the norms computed and applied on the
// device (sq: nk doubles of device output; scratch as dv_sqnorms)
int dv_normalize_cols(double* X, int nk, int64_t n, double* sq, double* scratch,
                      hipStream_t s);
// T -= Q (Q^T T), twice (ns + nt columns; G >= ns*nt doubles, scratch >=
// dv_gram_scratch(n, ns, nt)); classical Gram-Schmidt with reorthogonalization
int dv_project_out_block(const double* Q, int64_t ldq, int ns, double* T, int nt, int64_t n,
                         double* G, double* scratch, hipStream_t s);
// _qr (modified Gram-Schmidt with lindep drops) as Gram-matrix Gram-Schmidt,
// repeated once when a column kept less than 1 % of its squared norm (heavy
// cancellation): the accepted columns of X (ld n) go to Q (ld n), their count to
// *nout (host).  X is overwritten.  G, coef: device, >= nx^2 doubles;
// scratch >= dv_gram_scratch(n, nx, nx); host: >= 2 nx^2 doubles (the Gram
// matrix and the coefficients; pinned memory keeps the two transfers short)
// that stay valid until the stream has passed this call.  prescale: the
// columns of X are not normalized yet; one with squared norm <= lindep is
// dropped and the others are normalized before the Gram-Schmidt (all in the
// Gram matrix).
int dv_qr(double* X, int nx, int64_t n, double lindep, bool prescale, double* Q, int* nout,
          double* G, double* coef, double* scratch, double* host, hipStream_t s);

}  // namespace rta
