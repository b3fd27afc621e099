// Davidson eigensolver (A12), restated from davidson_solve (src/davidson.rs:48-266)
// with every vector resident in HBM.  The host keeps only the small
// subspace matrix heff and the control flow; dots, norms, Gram-Schmidt and
// Ritz-vector assembly run as device kernels (dvkernels.hip) on one stream.
//
// Kept from the reference: one initial guess vector, maxspace grown by
// 3*(nroots-1) (:74), fresh restarts when space + nroots > maxspace (:260),
// at most 40 new vectors (:115-117), the convergence test |de| < tol and
// |r| < sqrt(tol) (:189), the diagonal preconditioner with the shift e[0]
// for every root (:224), projection against the subspace (:235-243) and the
// lindep drops (:268-325).  The "large |r|" restore (:210-217) compares
// against a value that is never updated (:259 shadows it); the same constant
// 1e9 is used here.  Where the reference keeps using a pre-QR vector count
// after _qr dropped vectors (:112-136, an out-of-range index there), the
// post-QR count is used.
//
// Changed in form, equal in exact arithmetic: the projection of the new
// vectors against the orthonormal subspace is t -= Q (Q^T t) done twice
// (classical Gram-Schmidt with one reorthogonalization) instead of one
// subspace vector at a time, and _qr runs its column-by-column loop with the
// same drops on the small Gram matrix X^T X (host), then forms Q = X C on the
// device (a second time when the first pass cancelled heavily).  Both differ
// from the sequential order by rounding only and
// use the whole GPU instead of one workgroup per vector.  The subspace
// eigenproblem is solved by Householder tridiagonalization and implicit QL
// (the reference calls LAPACK's dspevx).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "dvkernels.hpp"
#include "mlp.hpp"

namespace rta {

// Eigenpairs of the small symmetric subspace matrix (column-major n x n):
// Householder reduction to tridiagonal form with Q accumulated (for k = 0..n-3
// the reflector v = x - alpha e1 zeroes A[k+2:, k], alpha = -sign(x0) |x|),
// then the implicit QL/QR iteration of rt_dsteqr_host applied to Q.
// Eigenvalues ascending in w, eigenvectors in the columns of V.
static int small_eigh(const std::vector<double>& A_in, int n, std::vector<double>& w,
                      std::vector<double>& V) {
  std::vector<double> A(A_in), v(n), p(n);
  auto a = [&](int i, int j) -> double& { return A[i + (size_t)n * j]; };
  V.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) V[i + (size_t)n * i] = 1.0;
  for (int k = 0; k + 2 < n; ++k) {
    double xn = 0.0;
    for (int i = k + 1; i < n; ++i) xn += a(i, k) * a(i, k);
    xn = std::sqrt(xn);
    if (xn == 0.0) continue;
    const double x0 = a(k + 1, k);
    const double alpha = x0 >= 0.0 ? -xn : xn;
    // v = x - alpha e1, scaled to |v| = 1 (|v|^2 = 2 xn (xn + |x0|))
    const double vn = std::sqrt(2.0 * xn * (xn + std::fabs(x0)));
    for (int i = k + 1; i < n; ++i) v[i] = a(i, k) / vn;
    v[k + 1] = (x0 - alpha) / vn;
    // trailing block B = A[k+1:, k+1:] -> H B H, H = I - 2 v v^T:
    // p = B v, q = p - (v.p) v, B -= 2 (v q^T + q v^T)
    double vp = 0.0;
    for (int i = k + 1; i < n; ++i) {
      double acc = 0.0;
      for (int j = k + 1; j < n; ++j) acc += a(i, j) * v[j];
      p[i] = acc;
    }
    for (int i = k + 1; i < n; ++i) vp += v[i] * p[i];
    for (int i = k + 1; i < n; ++i) p[i] -= vp * v[i];
    for (int j = k + 1; j < n; ++j)
      for (int i = k + 1; i < n; ++i) a(i, j) -= 2.0 * (v[i] * p[j] + p[i] * v[j]);
    a(k + 1, k) = a(k, k + 1) = alpha;
    for (int i = k + 2; i < n; ++i) a(i, k) = a(k, i) = 0.0;
    // Q = Q H on columns k+1..n-1
    for (int r = 0; r < n; ++r) {
      double acc = 0.0;
      for (int j = k + 1; j < n; ++j) acc += V[r + (size_t)n * j] * v[j];
      for (int j = k + 1; j < n; ++j) V[r + (size_t)n * j] -= 2.0 * acc * v[j];
    }
  }
  std::vector<double> e(n > 1 ? n - 1 : 1, 0.0);
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = a(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = a(i + 1, i);
  int info = 0;
  RTA_TRY(rt_dsteqr_host('V', n, w.data(), e.data(), V.data(), n, &info));
  if (info != 0)
    return fail(RT_ERR_NUMERIC, "davidson: subspace eigensolver did not converge (info %d)",
                info);
  return RT_OK;
}

struct DevBuf {
  double* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t count) {
    hipError_t e = hipMalloc(&p, sizeof(double) * std::max<size_t>(count, 1));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return fail(RT_ERR_NOMEM, "davidson: device allocation of %zu doubles failed",
                  count);
    }
    return RT_OK;
  }
};

// Pinned host memory for the per-iteration transfers (subspace matrices,
// coefficients, norms): pageable transfers go through a staging copy and
// cost tens of microseconds each, several times per iteration.
struct HostBuf {
  double* p = nullptr;
  ~HostBuf() {
    if (p) (void)hipHostFree(p);
  }
  int alloc(size_t count) {
    if (hipHostMalloc((void**)&p, sizeof(double) * std::max<size_t>(count, 1),
                      hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return fail(RT_ERR_NOMEM, "davidson: pinned host allocation of %zu doubles failed",
                  count);
    }
    return RT_OK;
  }
};

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace rta

using namespace rta;

extern "C" int rt_davidson_solve_dev(rt_matvec_fn a_x, void* ctx,
                                     const double* x0_dev,
                                     const double* hdiag_dev, int64_t n,
                                     const rt_davidson_params* prm, int* conv_out,
                                     double* e_out, double* x_out,
                                     rt_davidson_stats* stats, void* stream) {
  RTA_CHECK(a_x && prm && x0_dev && hdiag_dev && conv_out && e_out,
            "NULL argument");
  const int nroots = prm->nroots;
  RTA_CHECK(nroots >= 1 && nroots <= 64, "nroots must be in [1, 64]");
  RTA_CHECK(n >= 1, "empty problem");
  RTA_CHECK(n - nroots >= 3, "Too much nroots required (n - nroots < 3)");
  RTA_CHECK(prm->precond != RT_PRECOND_MLP || prm->mlp,
            "MLP preconditioner requested without a network");
  RTA_CHECK(prm->precond != RT_PRECOND_AUX || prm->aux_op,
            "auxiliary preconditioner requested without an operator");
  hipStream_t s = (hipStream_t)stream;
  const double t_start = now_s();
  const double tol = prm->tol, lindep = prm->lindep;
  const int maxspace = prm->maxspace + (nroots - 1) * 3;
  const int cap = maxspace + 2 * nroots + 48;
  const int tmax = std::max(nroots, 40);

  DevBuf XS, AX, X0, R, T, Sa, scr, small;
  RTA_TRY(XS.alloc((size_t)n * cap));
  RTA_TRY(AX.alloc((size_t)n * cap));
  RTA_TRY(X0.alloc((size_t)n * cap));
  RTA_TRY(R.alloc((size_t)n * tmax));
  RTA_TRY(T.alloc((size_t)n * tmax));
  if (prm->precond == RT_PRECOND_MLP && prm->aux_op) RTA_TRY(Sa.alloc((size_t)n * tmax));
  const int64_t scr_n =
      std::max<int64_t>(dv_gram_scratch(n, cap, tmax), dv_gram_scratch(n, 1, 1) * tmax);
  RTA_TRY(scr.alloc((size_t)scr_n));
  // small device area: e (cap), shift (cap), spare (cap), sq (cap), coef (cap*cap,
  // also the _qr coefficients), gram out (cap*tmax, also the _qr and projection Gram
  // matrices)
  const size_t o_e = 0, o_shift = cap,
               o_sq = 3 * (size_t)cap, o_coef = 4 * (size_t)cap,
               o_gram = o_coef + (size_t)cap * cap,
               o_end = o_gram + (size_t)cap * tmax;
  RTA_TRY(small.alloc(o_end));
  double* d_e = small.p + o_e;
  double* d_shift = small.p + o_shift;
  double* d_sq = small.p + o_sq;
  double* d_coef = small.p + o_coef;
  double* d_gram = small.p + o_gram;

  RTA_HIP(hipMemcpyAsync(X0.p, x0_dev, sizeof(double) * n,
                         hipMemcpyDeviceToDevice, s));
  int nx0 = 1;           // vectors in X0
  int nt = 0;            // new vectors in T (non-fresh)
  int space = 0;
  bool fresh = true;
  std::vector<double> heff, w, V, e, elast;
  std::vector<int> conv(nroots, 0);
  std::vector<double> dx_norm, de;
  // pinned host area, laid out like the small device area: e, shift, spare,
  // sq (cap each), coef (cap*cap), gram (cap*tmax), then the _qr Gram matrix
  // and coefficients (2*tmax*tmax)
  HostBuf hbuf;
  RTA_TRY(hbuf.alloc(o_end + 2 * (size_t)tmax * tmax));
  double* h_e = hbuf.p + o_e;
  double* h_shift = hbuf.p + o_shift;
  double* h_sq = hbuf.p + o_sq;
  double* h_coef = hbuf.p + o_coef;
  double* h_gram = hbuf.p + o_gram;
  double* h_qr = hbuf.p + o_end;
  int nroots_current = 0;
  rt_davidson_stats st;
  memset(&st, 0, sizeof st);
  int status = RT_ERR_NOTCONV;

  for (int icyc = 0; icyc < prm->maxcyc; ++icyc) {
    int xtlen = 0;
    if (fresh) {
      space = 0;
      RTA_TRY(dv_qr(X0.p, nx0, n, lindep, false, XS.p, &xtlen, d_gram, d_coef, scr.p, h_qr,
                    s));
      if (xtlen == 0)
        return fail(RT_ERR_NUMERIC, "No more linear independent basis found");
      ++st.restarts;
    } else {
      // the projected corrections, not yet normalized: the lindep drop and
      // normalization of davidson.rs:245-249 happen in the Gram matrix of
      // _qr (for nt = 1, where the reference skips _qr, the same drop and
      // normalization and nothing else)
      RTA_TRY(dv_qr(T.p, nt, n, lindep, true, XS.p + (size_t)space * n, &xtlen, d_gram,
                    d_coef, scr.p, h_qr, s));
      xtlen = std::min(xtlen, 40);
      if (xtlen == 0) {
        // linear dependency in the trial space (davidson.rs:250-257)
        bool ok = true;
        for (int k = 0; k < nroots; ++k) {
          if (k < (int)dx_norm.size())
            conv[k] = conv[k] || dx_norm[k] < std::sqrt(tol);
          ok = ok && conv[k];
        }
        status = ok ? RT_OK : RT_ERR_NOTCONV;
        break;
      }
    }
    if (space + xtlen > cap)
      return fail(RT_ERR_NUMERIC, "subspace overflow (%d + %d > %d)", space, xtlen, cap);
    // A x for the new vectors, written straight into the AX block
    RTA_TRY(a_x(ctx, XS.p + (size_t)space * n, AX.p + (size_t)space * n, n,
                xtlen, stream));
    st.matvecs += xtlen;
    const int row0 = space;
    space += xtlen;
    // heff: H[i,j] = H[j,i] = ax_i . x_j for i <= j, j new (davidson.rs:356-387)
    RTA_TRY(dv_gram(AX.p, n, space, XS.p + (size_t)row0 * n, n, xtlen, n, d_gram,
                    space, scr.p, s));
    RTA_HIP(hipMemcpyAsync(h_gram, d_gram, sizeof(double) * space * xtlen,
                           hipMemcpyDeviceToHost, s));
    RTA_HIP(hipStreamSynchronize(s));
    {
      std::vector<double> hn((size_t)space * space, 0.0);
      if (!fresh) {
        for (int j = 0; j < row0; ++j)
          for (int i = 0; i < row0; ++i)
            hn[i + (size_t)space * j] = heff[i + (size_t)row0 * j];
      }
      for (int jj = 0; jj < xtlen; ++jj) {
        const int j = row0 + jj;
        for (int i = 0; i <= j; ++i) {
          const double v = h_gram[i + (size_t)space * jj];
          hn[i + (size_t)space * j] = v;
          hn[j + (size_t)space * i] = v;
        }
      }
      heff.swap(hn);
    }
    elast = e;
    RTA_TRY(small_eigh(heff, space, w, V));
    const int nkeep = std::min(space, nroots);
    e.assign(w.begin(), w.begin() + nkeep);
    nroots_current = nkeep;
    nx0 = nkeep;
    for (int c = 0; c < nx0; ++c)
      for (int i = 0; i < space; ++i)
        h_coef[i + (size_t)space * c] = V[i + (size_t)space * c];
    for (int k = 0; k < nkeep; ++k) h_e[k] = e[k];
    RTA_HIP(hipMemcpyAsync(d_coef, h_coef, sizeof(double) * space * nx0,
                           hipMemcpyHostToDevice, s));
    RTA_HIP(hipMemcpyAsync(d_e, h_e, sizeof(double) * nkeep,
                           hipMemcpyHostToDevice, s));
    RTA_TRY(dv_ritz_residuals(XS.p, AX.p, n, space, d_coef, space, nx0, d_e, n, X0.p, R.p,
                              d_sq, scr.p, s));
    RTA_HIP(hipMemcpyAsync(h_sq, d_sq, sizeof(double) * nroots_current,
                           hipMemcpyDeviceToHost, s));
    RTA_HIP(hipStreamSynchronize(s));
    dx_norm.assign(nroots_current, 0.0);
    de.assign(nroots_current, 0.0);
    bool all_conv = true;
    double max_dx = 0.0;
    for (int k = 0; k < nroots_current; ++k) {
      de[k] = (k < (int)elast.size()) ? e[k] - elast[k] : e[k];
      dx_norm[k] = std::sqrt(h_sq[k]);
      conv[k] = (std::fabs(de[k]) < tol && dx_norm[k] < std::sqrt(tol)) ? 1 : 0;
      max_dx = std::max(max_dx, dx_norm[k]);
    }
    for (int k = 0; k < nroots; ++k) all_conv = all_conv && conv[k];
    if (prm->record) {
      rt_davidson_record* rec = prm->record;
      // the subspace this cycle's corrections are projected against
      int64_t xs_off = -1;
      if (rec->xs && rec->xs_off && rec->xs_cnt && rec->count < rec->cap &&
          rec->xs_used + (int64_t)space * n <= rec->xs_cap) {
        xs_off = rec->xs_used;
        RTA_HIP(hipMemcpyAsync(rec->xs + xs_off, XS.p, sizeof(double) * n * space,
                               hipMemcpyDeviceToDevice, s));
        rec->xs_used += (int64_t)space * n;
      }
      for (int k = 0; k < nroots_current && rec->count < rec->cap; ++k) {
        const int64_t q = rec->count++;
        if (rec->xs_off && rec->xs_cnt) {
          rec->xs_off[q] = xs_off;
          rec->xs_cnt[q] = xs_off >= 0 ? space : 0;
        }
        RTA_HIP(hipMemcpyAsync(rec->r + q * n, R.p + (size_t)k * n,
                               sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        RTA_HIP(hipMemcpyAsync(rec->x + q * n, X0.p + (size_t)k * n,
                               sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        rec->e_k[q] = e[k];
        rec->e_0[q] = e[0];
        rec->root[q] = k;
        rec->iter[q] = icyc;
      }
    }
    st.iterations = icyc + 1;
    st.max_residual = max_dx;
    if (prm->verbose)
      fprintf(stderr, "davidson cyc %d space %d |r|max %.3e e0 %.12f\n", icyc, space,
              max_dx, e[0]);
    if (all_conv) {
      status = RT_OK;
      break;
    }
    if (max_dx > 1.0 && max_dx / 1e9 > 3.0 && space > nroots + 2) {
      fresh = true;  // davidson.rs:210-217 (restart from the current Ritz vectors)
      continue;
    }
    // corrections for roots with |r|^2 > lindep (davidson.rs:220-233)
    std::vector<int> keep;
    for (int k = 0; k < nroots_current; ++k)
      if (dx_norm[k] * dx_norm[k] > lindep) keep.push_back(k);
    nt = (int)keep.size();
    // compact the kept residuals / Ritz vectors to the front
    for (int j = 0; j < nt; ++j)
      if (keep[j] != j) {
        RTA_HIP(hipMemcpyAsync(R.p + (size_t)j * n, R.p + (size_t)keep[j] * n,
                               sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      }
    if (nt > 0) {
      if (prm->precond == RT_PRECOND_MLP) {
        if (prm->aux_op)
          RTA_TRY(prm->aux_op(prm->aux_ctx, R.p, Sa.p, n, nt, stream));
        RTA_TRY(mlp_precondition((const MlpNet*)prm->mlp, R.p, X0.p,
                                 prm->aux_op ? Sa.p : nullptr, keep.data(), nt,
                                 hdiag_dev, e.data(), n, T.p, scr.p, s,
                                 &st.mlp_fallbacks));
      } else if (prm->precond == RT_PRECOND_AUX) {
        RTA_TRY(prm->aux_op(prm->aux_ctx, R.p, T.p, n, nt, stream));
      } else {
        for (int j = 0; j < nt; ++j)
          h_shift[j] = (prm->precond == RT_PRECOND_DIAG_ROOT) ? e[keep[j]] : e[0];
        RTA_HIP(hipMemcpyAsync(d_shift, h_shift, sizeof(double) * nt,
                               hipMemcpyHostToDevice, s));
        RTA_TRY(dv_precond_diag(R.p, hdiag_dev, d_shift, nt, n, T.p, s));
      }
      // normalize (on the device), project out the subspace
      RTA_TRY(dv_normalize_cols(T.p, nt, n, d_sq, scr.p, s));
      RTA_TRY(dv_project_out_block(XS.p, n, space, T.p, nt, n, d_gram, scr.p, s));
    }
    const bool fresh_next = space + nroots > maxspace;
    if (nt > 0 && fresh_next) {
      // a fresh restart does not use the corrections; only the lindep test
      // on them (davidson.rs:245-257) is needed here.  Otherwise the next
      // _qr drops and normalizes them without a synchronization of its own.
      RTA_TRY(dv_sqnorms(T.p, nt, n, d_sq, scr.p, s));
      RTA_HIP(hipMemcpyAsync(h_sq, d_sq, sizeof(double) * nt,
                             hipMemcpyDeviceToHost, s));
      RTA_HIP(hipStreamSynchronize(s));
      int m = 0;
      for (int j = 0; j < nt; ++j) m += h_sq[j] > lindep ? 1 : 0;
      nt = m;
    }
    if (nt == 0) {
      // linear dependency in the trial space (davidson.rs:250-257)
      bool ok = true;
      for (int k = 0; k < nroots; ++k) {
        if (k < (int)dx_norm.size())
          conv[k] = conv[k] || dx_norm[k] < std::sqrt(tol);
        ok = ok && conv[k];
      }
      status = ok ? RT_OK : RT_ERR_NOTCONV;
      break;
    }
    fresh = fresh_next;
  }
  RTA_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < nroots; ++k) conv_out[k] = conv[k];
  for (int k = 0; k < nroots; ++k)
    e_out[k] = (k < (int)e.size()) ? e[k] : std::nan("");
  if (x_out && nroots_current > 0)
    RTA_HIP(hipMemcpyAsync(x_out, X0.p, sizeof(double) * n * nroots_current,
                           hipMemcpyDeviceToDevice, s));
  RTA_HIP(hipStreamSynchronize(s));
  st.nroots_found = nroots_current;
  st.seconds = now_s() - t_start;
  if (stats) *stats = st;
  if (status != RT_OK)
    fail(status, "davidson did not converge in %d cycles (max |r| %.3e)",
         prm->maxcyc, st.max_residual);
  return status;
}

// ---- the RI linear-response operator behind the C ABI -----------------------------
extern "C" int rt_casida_op_create(const double* B, int64_t ldb, int64_t n,
                                   int64_t naux, double scale, const double* d,
                                   void** op) {
  RTA_CHECK(op && B && d, "NULL argument");
  RTA_CHECK(n > 0 && naux >= 0 && ldb >= n, "bad operator shape");
  CasidaOp* o = new CasidaOp{B, ldb, n, naux, scale, d};
  *op = o;
  return RT_OK;
}

extern "C" int rt_casida_op_destroy(void* op) {
  delete (CasidaOp*)op;
  return RT_OK;
}

extern "C" int rt_casida_matvec(void* op, const double* X, double* Y, int64_t n,
                                int64_t k, void* stream) {
  RTA_CHECK(op, "NULL operator");
  return casida_apply((const CasidaOp*)op, X, Y, n, k, (hipStream_t)stream);
}

extern "C" int rt_casida_hdiag(void* op, double* hdiag, void* stream) {
  RTA_CHECK(op && hdiag, "NULL argument");
  return casida_hdiag((const CasidaOp*)op, hdiag, (hipStream_t)stream);
}

extern "C" int rt_casida_si_create(void* op, double shift, void** si, void* stream) {
  RTA_CHECK(op && si, "NULL argument");
  CasidaSI* p = new CasidaSI{};
  p->op = *(const CasidaOp*)op;
  const int st = casida_si_setup(p, shift, (hipStream_t)stream);
  if (st != RT_OK) {
    if (p->Dm) (void)hipFree(p->Dm);
    if (p->Ci) (void)hipFree(p->Ci);
    delete p;
    return st;
  }
  *si = p;
  return RT_OK;
}

extern "C" int rt_casida_si_destroy(void* si) {
  if (!si) return RT_OK;
  CasidaSI* p = (CasidaSI*)si;
  (void)hipFree(p->Dm);
  (void)hipFree(p->Ci);
  delete p;
  return RT_OK;
}

extern "C" int rt_casida_si_apply(void* si, const double* R, double* T, int64_t n,
                                  int64_t k, void* stream) {
  RTA_CHECK(si && R && T, "NULL argument");
  return casida_si_apply((const CasidaSI*)si, R, T, n, k, (hipStream_t)stream);
}

extern "C" int rt_casida_si_info(void* si, double* sigma, double* err) {
  RTA_CHECK(si, "NULL argument");
  const CasidaSI* p = (const CasidaSI*)si;
  if (sigma) *sigma = p->sigma;
  if (err) *err = p->inv_err;
  return RT_OK;
}
