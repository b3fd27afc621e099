/*
 * restmatr.h — drop-in replacements for the seven symbols of the reference's
 * librestmatr.so, with the exact by-reference (Fortran implicit-interface)
 * calling convention that the Rust crate binds in
 * src/external_libs/ffi_restmatr.rs:4-61.
 *
 * Conventions kept from the reference:
 *   - every argument is a pointer, the return type is void, names carry a
 *     trailing underscore;
 *   - arrays are column-major host buffers owned by the caller;
 *   - block starts are 0-based, lengths are element counts (the Fortran adds 1);
 *   - op characters are read from their first byte only; 'N' means no
 *     transpose, anything else means transpose (restmatr.f90:85-97);
 *   - the callee writes its output block in place.
 *
 * The arithmetic runs on the GPU (hand-written gfx950 kernels).  There is no
 * status channel in this ABI: on a bad argument or a HIP failure the symbol
 * prints the reason to stderr, leaves the output untouched where it can, and
 * records the reason for rt_last_error() (rest_tensors_amd.h).
 */
#ifndef RESTMATR_H
#define RESTMATR_H

#ifdef __cplusplus
extern "C" {
#endif

/* ri3mo(P,i,j) = sum_{mu,nu} C(mu,i) ri3fn(mu,nu,P) C(nu,j)
 * ri3fn: [num_basis, num_basis, num_auxbas], C: [num_basis, num_states],
 * ri3mo: [num_auxbas, num_states, num_states] (aux index fastest).
 * Replaces ri_ao2mo_f_ (ffi_restmatr.rs:5-12; restmatr.f90:158-194).
 * The reference is only defined for num_states == num_basis (it passes
 * ldc = num_basis for an ns x ns section); this version gives the
 * mathematically intended result for every num_states <= num_basis. */
void ri_ao2mo_f_(const double* eigenvector, const double* ri3fn, double* ri3mo,
                 const int* num_states, const int* num_basis,
                 const int* num_auxbas);

/* C[rc] = alpha * op(A[ra]) * op(B[rb]) + beta * C[rc] on sub-blocks.
 * Replaces general_dgemm_f_ (ffi_restmatr.rs:13-24; restmatr.f90:63-107). */
void general_dgemm_f_(
    const double* matr_a, const int* rows_a, const int* columns_a,
    const int* start_row_a, const int* len_row_a, const int* start_column_a,
    const int* len_column_a, const char* opa,
    const double* matr_b, const int* rows_b, const int* columns_b,
    const int* start_row_b, const int* len_row_b, const int* start_column_b,
    const int* len_column_b, const char* opb,
    double* matr_c, const int* rows_c, const int* columns_c,
    const int* start_row_c, const int* len_row_c, const int* start_column_c,
    const int* len_column_c, const double* alpha, const double* beta);

/* For every y-slice of ten3_a: ten3_a[x-range, y, z-range] <-
 *   alpha * ten3_a[x-range, y, z-range] * B[row-range, col-range]
 *   + beta * ten3_a[x-range, y, z-range].
 * Replaces special_dgemm_f_01_ (ffi_restmatr.rs:25-33; restmatr.f90:111-154).
 * Like the reference it walks all y slices and ignores i_y; unlike it, the
 * product is formed out of place (the reference aliases dgemm's A and C). */
void special_dgemm_f_01_(
    double* ten3_a, const int* x_a, const int* y_a, const int* z_a,
    const int* start_x_a, const int* len_x_a, const int* i_y,
    const int* start_z_a, const int* len_z_a,
    const double* matr_b, const int* rows_b, const int* columns_b,
    const int* start_row_b, const int* len_row_b, const int* start_column_b,
    const int* len_column_b, const double* alpha, const double* beta);

/* Block copy matrix -> matrix.  copy_mm_ (ffi_restmatr.rs:35-39;
 * restmatr.f90:197-212). */
void copy_mm_(const int* x_len, const int* y_len,
              const double* f_matr, const int* f_x_len, const int* f_y_len,
              const int* f_x_start, const int* f_y_start,
              double* t_matr, const int* t_x_len, const int* t_y_len,
              const int* t_x_start, const int* t_y_start);

/* Block copy matrix -> rank-3 tensor; mod 0/1/2 chooses which tensor axis is
 * fixed at t_x3.  copy_mr_ (ffi_restmatr.rs:41-46; restmatr.f90:215-238). */
void copy_mr_(const int* x_len, const int* y_len,
              const double* f_matr, const int* f_x_len, const int* f_y_len,
              const int* f_x_start, const int* f_y_start,
              double* t_ri, const int* t_x_len, const int* t_y_len,
              const int* t_z_len, const int* t_x_start, const int* t_y_start,
              const int* t_x3, const int* t_mod);

/* Block copy rank-3 tensor -> matrix.  copy_rm_ (ffi_restmatr.rs:48-53;
 * restmatr.f90:241-264). */
void copy_rm_(const int* x_len, const int* y_len,
              const double* f_ri, const int* f_x_len, const int* f_y_len,
              const int* f_z_len, const int* f_x_start, const int* f_y_start,
              const int* f_x3, const int* f_mod,
              double* t_matr, const int* t_x_len, const int* t_y_len,
              const int* t_x_start, const int* t_y_start);

/* Block copy rank-3 tensor -> rank-3 tensor.  copy_rr_ (ffi_restmatr.rs:55-60;
 * restmatr.f90:266-285). */
void copy_rr_(const int* x_len, const int* y_len, const int* z_len,
              const double* f_ri, const int* f_x_len, const int* f_y_len,
              const int* f_z_len, const int* f_x_start, const int* f_y_start,
              const int* f_z_start,
              double* t_ri, const int* t_x_len, const int* t_y_len,
              const int* t_z_len, const int* t_x_start, const int* t_y_start,
              const int* t_z_start);

#ifdef __cplusplus
}
#endif

#endif /* RESTMATR_H */
