// Dense FP64 factorizations shared by the metric and eigensolver entry points.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rta {

// W (ld n) <- lower Cholesky factor in its lower triangle; *info_h = first
// non-positive pivot (1-based) or 0; info_d: one device int of scratch.
// Synchronizes the stream.
int chol_lower(double* W, int64_t n, int* info_h, int* info_d, hipStream_t s);

// X (ld n) <- L^-1 for the lower triangular factor in Lf (ld n).
int tri_inverse_lower(const double* Lf, int64_t n, double* X, hipStream_t s);

// Eigen-decomposition of the symmetric matrix whose lower triangle is
// scale * A (ld lda): ascending eigenvalues into w_host[n]; eigenvectors into
// Z (ld ldz) when Z is non-null.  *info as LAPACK DSYEV.  Synchronizes.
int syev_device(int64_t n, const double* A, int64_t lda, double scale, double* w_host,
                double* Z, int64_t ldz, int* info, hipStream_t s);

}  // namespace rta
