// Drop-in replacements of the reference's seven librestmatr symbols
// (src/external_libs/ffi_restmatr.rs:4-61), host buffers in, host buffers
// out, arithmetic on the GPU.
//
// Each symbol uploads the smallest contiguous range of every host array that
// contains its block (the "box": first to last element touched, column-major),
// runs the device kernel on box-relative pointers and downloads the output
// box.  Rows of the output box outside the block round-trip unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/restmatr.h"
#include "common.hpp"
#include "gemm.hpp"
#include "layout.hpp"

namespace rta {

// A block of a column-major array: lengths (X,Y,Z), offset of its first
// element, strides of its three axes (all >= 0).
struct BlockView {
  int64_t X = 0, Y = 0, Z = 0;
  int64_t off = 0, sx = 1, sy = 0, sz = 0;
  bool empty() const { return X <= 0 || Y <= 0 || Z <= 0; }
  int64_t last() const {
    return off + (X - 1) * sx + (Y - 1) * sy + (Z - 1) * sz;
  }
  int64_t box_len() const { return empty() ? 0 : last() - off + 1; }
};

static int upload_box(const double* host, const BlockView& v, int slot,
                      double** dev) {
  *dev = nullptr;
  int64_t n = v.box_len();
  if (n == 0) return RT_OK;
  RTA_TRY(scratch_get(slot, sizeof(double) * n, (void**)dev));
  RTA_HIP(hipMemcpy(*dev, host + v.off, sizeof(double) * n,
                    hipMemcpyHostToDevice));
  return RT_OK;
}

static int download_box(double* host, const BlockView& v, const double* dev) {
  int64_t n = v.box_len();
  if (n == 0) return RT_OK;
  RTA_HIP(hipMemcpy(host + v.off, dev, sizeof(double) * n,
                    hipMemcpyDeviceToHost));
  return RT_OK;
}

static bool blk_ok(int64_t start, int64_t len, int64_t dim) {
  return start >= 0 && len >= 0 && start + len <= dim;
}

static int op_is_trans(char c, bool* t) {
  if (c == 'N' || c == 'n') {
    *t = false;
    return RT_OK;
  }
  if (c == 'T' || c == 't' || c == 'C' || c == 'c') {
    *t = true;
    return RT_OK;
  }
  return fail(RT_ERR_ARG, "op must be 'N' or 'T', got 0x%02x", (unsigned char)c);
}

static int ri_ao2mo_host(const double* C, const double* ri3fn, double* ri3mo,
                         int64_t ns, int64_t nbas, int64_t naux) {
  RTA_CHECK(ns >= 0 && nbas >= 0 && naux >= 0, "negative dimension");
  RTA_CHECK(ns <= nbas, "num_states (%lld) > num_basis (%lld) is undefined",
            (long long)ns, (long long)nbas);
  if (ns == 0 || naux == 0) return RT_OK;
  RTA_CHECK(C && ri3fn && ri3mo, "NULL buffer");
  if (nbas == 0) {  // empty contraction: the reference zero-fills ri3mo
    memset(ri3mo, 0, sizeof(double) * naux * ns * ns);
    return RT_OK;
  }
  const int64_t slice = nbas * nbas;
  const size_t budget = size_t(4) << 30;
  int64_t pc = std::max<int64_t>(
      1, std::min<int64_t>(naux, (int64_t)(budget / (sizeof(double) * slice))));
  double *dC, *dA, *dR;
  RTA_TRY(scratch_get(SCRATCH_DROPIN_B, sizeof(double) * nbas * ns,
                      (void**)&dC));
  RTA_TRY(scratch_get(SCRATCH_DROPIN_A, sizeof(double) * slice * pc,
                      (void**)&dA));
  RTA_TRY(scratch_get(SCRATCH_DROPIN_C, sizeof(double) * ns * ns * pc,
                      (void**)&dR));
  RTA_HIP(hipMemcpy(dC, C, sizeof(double) * nbas * ns, hipMemcpyHostToDevice));
  for (int64_t p0 = 0; p0 < naux; p0 += pc) {
    int64_t np = std::min(pc, naux - p0);
    RTA_HIP(hipMemcpy(dA, ri3fn + p0 * slice, sizeof(double) * slice * np,
                      hipMemcpyHostToDevice));
    RTA_TRY(rt_ao2mo_dev(dC, dA, dR, ns, nbas, np, nullptr));
    // chunk layout [np, ns, ns] -> host rows of length np at pitch naux
    RTA_HIP(hipMemcpy2D(ri3mo + p0, sizeof(double) * naux, dR,
                        sizeof(double) * np, sizeof(double) * np,
                        (size_t)(ns * ns), hipMemcpyDeviceToHost));
  }
  return RT_OK;
}

static int general_dgemm_host(const double* A, int64_t ra, int64_t ca,
                              int64_t sra, int64_t lra, int64_t sca,
                              int64_t lca, char opa, const double* B,
                              int64_t rb, int64_t cb, int64_t srb, int64_t lrb,
                              int64_t scb, int64_t lcb, char opb, double* C,
                              int64_t rc, int64_t cc, int64_t src, int64_t lrc,
                              int64_t scc, int64_t lcc, double alpha,
                              double beta) {
  bool ta, tb;
  RTA_TRY(op_is_trans(opa, &ta));
  RTA_TRY(op_is_trans(opb, &tb));
  RTA_CHECK(blk_ok(sra, lra, ra) && blk_ok(sca, lca, ca) &&
                blk_ok(srb, lrb, rb) && blk_ok(scb, lcb, cb) &&
                blk_ok(src, lrc, rc) && blk_ok(scc, lcc, cc),
            "sub-block outside its matrix");
  RTA_CHECK(A && B && C, "NULL buffer");
  // boxes: whole columns of each block
  BlockView va{ra, lca, 1, sca * ra, 1, ra, 0};
  BlockView vb{rb, lcb, 1, scb * rb, 1, rb, 0};
  BlockView vc{rc, lcc, 1, scc * rc, 1, rc, 0};
  if (lra == 0) va.X = 0;
  if (lrb == 0) vb.X = 0;
  if (lrc == 0 || lcc == 0) return RT_OK;  // nothing to write
  double *dA, *dB, *dC;
  RTA_TRY(upload_box(A, va, SCRATCH_DROPIN_A, &dA));
  RTA_TRY(upload_box(B, vb, SCRATCH_DROPIN_B, &dB));
  RTA_TRY(upload_box(C, vc, SCRATCH_DROPIN_C, &dC));
  // device blocks are box-relative: same rows, columns start at 0
  RTA_TRY(rt_general_dgemm_dev(dA, ra, lca, sra, lra, 0, lca, ta ? 'T' : 'N',
                               dB, rb, lcb, srb, lrb, 0, lcb, tb ? 'T' : 'N',
                               dC, rc, lcc, src, lrc, 0, lcc, alpha, beta,
                               nullptr));
  return download_box(C, vc, dC);
}

// special_dgemm_f_01 on device buffers: Tz points at the z-slab start of the
// xa x ya x za tensor (its element (x, y, z) at x + xa*y + xa*ya*(z - sza)), Bc
// at column scb of the rb-row matrix B.  Per y slice the product is formed
// out of place (R) and then combined: T = R + beta T.
static int special_dgemm_core(double* Tz, int64_t xa, int64_t ya, int64_t sxa, int64_t lxa,
                              int64_t lza, const double* Bc, int64_t rb, int64_t srb,
                              double alpha, double beta, hipStream_t s) {
  double* dR;
  RTA_TRY(scratch_get(SCRATCH_DROPIN_C, sizeof(double) * lxa * lza * ya, (void**)&dR));
  // R_y = alpha * T[sxa:, y, z-slab] * B[srb:, cols]
  GemmDesc d;
  d.ta = false;
  d.tb = false;
  d.M = lxa;
  d.N = lza;
  d.K = lza;
  d.A = Tz + sxa;
  d.lda = xa * ya;
  d.sA = xa;
  d.B = Bc + srb;
  d.ldb = rb;
  d.sB = 0;
  d.C = dR;
  d.c_rs = 1;
  d.c_cs = lxa;
  d.sC = lxa * lza;
  d.alpha = alpha;
  d.beta = 0.0;
  d.batch = ya;
  RTA_TRY(gemm_batched(d, s, "special_dgemm"));
  // T[x, y, z] = R_y[x, z] + beta * T[x, y, z]
  return axpby3d(lxa, lza, ya, dR, 1, lxa, lxa * lza, Tz + sxa, 1, xa * ya, xa, beta, s);
}

static int special_dgemm_check(int64_t xa, int64_t ya, int64_t za, int64_t sxa, int64_t lxa,
                               int64_t sza, int64_t lza, int64_t rb, int64_t cb, int64_t srb,
                               int64_t lrb, int64_t scb, int64_t lcb) {
  RTA_CHECK(xa >= 0 && ya >= 0 && za >= 0, "negative dimension");
  RTA_CHECK(blk_ok(sxa, lxa, xa) && blk_ok(sza, lza, za) &&
                blk_ok(srb, lrb, rb) && blk_ok(scb, lcb, cb),
            "block outside its array");
  RTA_CHECK(lrb == lza && lcb == lza,
            "B block must be len_z x len_z (got %lldx%lld, len_z %lld)",
            (long long)lrb, (long long)lcb, (long long)lza);
  return RT_OK;
}

static int special_dgemm_host(double* T, int64_t xa, int64_t ya, int64_t za,
                              int64_t sxa, int64_t lxa, int64_t sza,
                              int64_t lza, const double* Bm, int64_t rb,
                              int64_t cb, int64_t srb, int64_t lrb,
                              int64_t scb, int64_t lcb, double alpha,
                              double beta) {
  RTA_TRY(special_dgemm_check(xa, ya, za, sxa, lxa, sza, lza, rb, cb, srb, lrb, scb, lcb));
  if (lxa == 0 || ya == 0 || lza == 0) return RT_OK;
  RTA_CHECK(T && Bm, "NULL buffer");
  // tensor box: the z-slab [sza, sza+lza) (contiguous), B box: its columns
  BlockView vt{xa * ya, lza, 1, sza * xa * ya, 1, xa * ya, 0};
  BlockView vb{rb, lcb, 1, scb * rb, 1, rb, 0};
  double *dT, *dB;
  RTA_TRY(upload_box(T, vt, SCRATCH_DROPIN_A, &dT));
  RTA_TRY(upload_box(Bm, vb, SCRATCH_DROPIN_B, &dB));
  RTA_TRY(special_dgemm_core(dT, xa, ya, sxa, lxa, lza, dB, rb, srb, alpha, beta, nullptr));
  return download_box(T, vt, dT);
}

// copy between two block views of host arrays through the device
static int copy_host(const double* f, const BlockView& vf, double* t,
                     const BlockView& vt) {
  if (vf.empty()) return RT_OK;
  RTA_CHECK(f && t, "NULL buffer");
  double *df, *dt;
  RTA_TRY(upload_box(f, vf, SCRATCH_DROPIN_A, &df));
  RTA_TRY(upload_box(t, vt, SCRATCH_DROPIN_B, &dt));
  RTA_TRY(copy3d(vf.X, vf.Y, vf.Z, df, vf.sx, vf.sy, vf.sz, dt, vt.sx, vt.sy,
                 vt.sz, nullptr));
  return download_box(t, vt, dt);
}

// (a,b) block of a rank-3 tensor [n0,n1,n2] with `mod` fixing one axis at x3
static int ri_view(int mod, int64_t n0, int64_t n1, int64_t n2, int64_t as,
                   int64_t alen, int64_t bs, int64_t blen, int64_t x3,
                   BlockView* v) {
  v->X = alen;
  v->Y = blen;
  v->Z = 1;
  v->sz = 0;
  switch (mod) {
    case 0:
      RTA_CHECK(blk_ok(as, alen, n0) && blk_ok(bs, blen, n1) && blk_ok(x3, 1, n2),
                "rank-3 block outside its tensor (mod 0)");
      v->off = as + bs * n0 + x3 * n0 * n1;
      v->sx = 1;
      v->sy = n0;
      return RT_OK;
    case 1:
      RTA_CHECK(blk_ok(as, alen, n0) && blk_ok(x3, 1, n1) && blk_ok(bs, blen, n2),
                "rank-3 block outside its tensor (mod 1)");
      v->off = as + x3 * n0 + bs * n0 * n1;
      v->sx = 1;
      v->sy = n0 * n1;
      return RT_OK;
    case 2:
      RTA_CHECK(blk_ok(x3, 1, n0) && blk_ok(as, alen, n1) && blk_ok(bs, blen, n2),
                "rank-3 block outside its tensor (mod 2)");
      v->off = x3 + as * n0 + bs * n0 * n1;
      v->sx = n0;
      v->sy = n0 * n1;
      return RT_OK;
    default:
      return fail(RT_ERR_ARG, "copy mod must be 0, 1 or 2 (got %d)", mod);
  }
}

static int mat_view(int64_t rows, int64_t cols, int64_t xs, int64_t xlen,
                    int64_t ys, int64_t ylen, BlockView* v) {
  RTA_CHECK(blk_ok(xs, xlen, rows) && blk_ok(ys, ylen, cols),
            "matrix block outside its matrix");
  *v = BlockView{xlen, ylen, 1, xs + ys * rows, 1, rows, 0};
  return RT_OK;
}

}  // namespace rta

using namespace rta;

#define DROPIN(sym, call)                              \
  do {                                                 \
    if ((call) != RT_OK) report_dropin_error(sym);     \
  } while (0)

extern "C" {

void ri_ao2mo_f_(const double* eigenvector, const double* ri3fn, double* ri3mo,
                 const int* num_states, const int* num_basis,
                 const int* num_auxbas) {
  DROPIN("ri_ao2mo_f_", ri_ao2mo_host(eigenvector, ri3fn, ri3mo, *num_states,
                                      *num_basis, *num_auxbas));
}

void general_dgemm_f_(
    const double* matr_a, const int* rows_a, const int* columns_a,
    const int* start_row_a, const int* len_row_a, const int* start_column_a,
    const int* len_column_a, const char* opa, const double* matr_b,
    const int* rows_b, const int* columns_b, const int* start_row_b,
    const int* len_row_b, const int* start_column_b, const int* len_column_b,
    const char* opb, double* matr_c, const int* rows_c, const int* columns_c,
    const int* start_row_c, const int* len_row_c, const int* start_column_c,
    const int* len_column_c, const double* alpha, const double* beta) {
  DROPIN("general_dgemm_f_",
         general_dgemm_host(matr_a, *rows_a, *columns_a, *start_row_a,
                            *len_row_a, *start_column_a, *len_column_a, *opa,
                            matr_b, *rows_b, *columns_b, *start_row_b,
                            *len_row_b, *start_column_b, *len_column_b, *opb,
                            matr_c, *rows_c, *columns_c, *start_row_c,
                            *len_row_c, *start_column_c, *len_column_c, *alpha,
                            *beta));
}

void special_dgemm_f_01_(double* ten3_a, const int* x_a, const int* y_a,
                         const int* z_a, const int* start_x_a,
                         const int* len_x_a, const int* i_y,
                         const int* start_z_a, const int* len_z_a,
                         const double* matr_b, const int* rows_b,
                         const int* columns_b, const int* start_row_b,
                         const int* len_row_b, const int* start_column_b,
                         const int* len_column_b, const double* alpha,
                         const double* beta) {
  (void)i_y;  // the reference walks every y slice (restmatr.f90:130)
  DROPIN("special_dgemm_f_01_",
         special_dgemm_host(ten3_a, *x_a, *y_a, *z_a, *start_x_a, *len_x_a,
                            *start_z_a, *len_z_a, matr_b, *rows_b, *columns_b,
                            *start_row_b, *len_row_b, *start_column_b,
                            *len_column_b, *alpha, *beta));
}

void copy_mm_(const int* x_len, const int* y_len, const double* f_matr,
              const int* f_x_len, const int* f_y_len, const int* f_x_start,
              const int* f_y_start, double* t_matr, const int* t_x_len,
              const int* t_y_len, const int* t_x_start, const int* t_y_start) {
  BlockView vf, vt;
  int st = mat_view(*f_x_len, *f_y_len, *f_x_start, *x_len, *f_y_start, *y_len,
                    &vf);
  if (st == RT_OK)
    st = mat_view(*t_x_len, *t_y_len, *t_x_start, *x_len, *t_y_start, *y_len,
                  &vt);
  if (st == RT_OK) st = copy_host(f_matr, vf, t_matr, vt);
  if (st != RT_OK) report_dropin_error("copy_mm_");
}

void copy_mr_(const int* x_len, const int* y_len, const double* f_matr,
              const int* f_x_len, const int* f_y_len, const int* f_x_start,
              const int* f_y_start, double* t_ri, const int* t_x_len,
              const int* t_y_len, const int* t_z_len, const int* t_x_start,
              const int* t_y_start, const int* t_x3, const int* t_mod) {
  BlockView vf, vt;
  int st = mat_view(*f_x_len, *f_y_len, *f_x_start, *x_len, *f_y_start, *y_len,
                    &vf);
  if (st == RT_OK)
    st = ri_view(*t_mod, *t_x_len, *t_y_len, *t_z_len, *t_x_start, *x_len,
                 *t_y_start, *y_len, *t_x3, &vt);
  if (st == RT_OK) st = copy_host(f_matr, vf, t_ri, vt);
  if (st != RT_OK) report_dropin_error("copy_mr_");
}

void copy_rm_(const int* x_len, const int* y_len, const double* f_ri,
              const int* f_x_len, const int* f_y_len, const int* f_z_len,
              const int* f_x_start, const int* f_y_start, const int* f_x3,
              const int* f_mod, double* t_matr, const int* t_x_len,
              const int* t_y_len, const int* t_x_start, const int* t_y_start) {
  BlockView vf, vt;
  int st = ri_view(*f_mod, *f_x_len, *f_y_len, *f_z_len, *f_x_start, *x_len,
                   *f_y_start, *y_len, *f_x3, &vf);
  if (st == RT_OK)
    st = mat_view(*t_x_len, *t_y_len, *t_x_start, *x_len, *t_y_start, *y_len,
                  &vt);
  if (st == RT_OK) st = copy_host(f_ri, vf, t_matr, vt);
  if (st != RT_OK) report_dropin_error("copy_rm_");
}

void copy_rr_(const int* x_len, const int* y_len, const int* z_len,
              const double* f_ri, const int* f_x_len, const int* f_y_len,
              const int* f_z_len, const int* f_x_start, const int* f_y_start,
              const int* f_z_start, double* t_ri, const int* t_x_len,
              const int* t_y_len, const int* t_z_len, const int* t_x_start,
              const int* t_y_start, const int* t_z_start) {
  int st = RT_OK;
  if (!(blk_ok(*f_x_start, *x_len, *f_x_len) &&
        blk_ok(*f_y_start, *y_len, *f_y_len) &&
        blk_ok(*f_z_start, *z_len, *f_z_len) &&
        blk_ok(*t_x_start, *x_len, *t_x_len) &&
        blk_ok(*t_y_start, *y_len, *t_y_len) &&
        blk_ok(*t_z_start, *z_len, *t_z_len)))
    st = fail(RT_ERR_ARG, "copy_rr block outside its tensor");
  if (st == RT_OK) {
    int64_t fx = *f_x_len, fy = *f_y_len, tx = *t_x_len, ty = *t_y_len;
    BlockView vf{*x_len, *y_len, *z_len,
                 *f_x_start + *f_y_start * fx + (int64_t)*f_z_start * fx * fy,
                 1, fx, fx * fy};
    BlockView vt{*x_len, *y_len, *z_len,
                 *t_x_start + *t_y_start * tx + (int64_t)*t_z_start * tx * ty,
                 1, tx, tx * ty};
    st = copy_host(f_ri, vf, t_ri, vt);
  }
  if (st != RT_OK) report_dropin_error("copy_rr_");
}

}  // extern "C"

// special_dgemm_f_01 (restmatr.f90:111-154) on device buffers, in place.
extern "C" int rt_special_dgemm_dev(double* ten3_a, int64_t x_a, int64_t y_a, int64_t z_a,
                                    int64_t start_x_a, int64_t len_x_a, int64_t start_z_a,
                                    int64_t len_z_a, const double* matr_b, int64_t rows_b,
                                    int64_t columns_b, int64_t start_row_b, int64_t len_row_b,
                                    int64_t start_column_b, int64_t len_column_b, double alpha,
                                    double beta, void* stream) {
  RTA_TRY(special_dgemm_check(x_a, y_a, z_a, start_x_a, len_x_a, start_z_a, len_z_a, rows_b,
                              columns_b, start_row_b, len_row_b, start_column_b, len_column_b));
  if (len_x_a == 0 || y_a == 0 || len_z_a == 0) return RT_OK;
  RTA_CHECK(ten3_a && matr_b, "NULL buffer");
  return special_dgemm_core(ten3_a + start_z_a * x_a * y_a, x_a, y_a, start_x_a, len_x_a,
                            len_z_a, matr_b + start_column_b * rows_b, rows_b, start_row_b,
                            alpha, beta, (hipStream_t)stream);
}

