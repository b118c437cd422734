// FP64 dense subspace solvers on the GPU (see dense_gpu.hip).  Every routine
// works on a batch of matrices; unless stated otherwise matrix kz of a batch
// is m x m, column-major with ld m, at offset kz*m*m (vectors at kz*m).
#pragma once
#include "common.hpp"

namespace derek {

// Doubles of workspace per matrix for heev_gpu(n, nev).
int64_t heev_work_doubles(int n, int nev);
// Lowest nev eigenpairs of `batch` Hermitian n x n matrices (A destroyed;
// eigenvalues ascending in W, eigenvectors in the columns of Z).
void heev_gpu(double2* A, int lda, int64_t strideA, int n, int nev, double* W, int64_t strideW, double2* Z, int ldz,
              int64_t strideZ, double* work, int64_t strideWork, int batch, hipStream_t s);
// A <- (A + A^H)/2
void hermitize_gpu(double2* A, int m, int batch, hipStream_t s);
// dsc = 1/sqrt(diag S) (0 where the diagonal is 0), S <- D S D
void unit_diag_gpu(double2* S, int m, double* dsc, int batch, hipStream_t s);
// Cholesky of the unit-diagonal S skipping columns whose remaining pivot is
// <= tol * S_jj: Lf = L on the kept rows/columns and the identity on the
// dropped ones; keep[m] flags, na = kept count, ok = (na >= nev), per matrix.
// Wm: m x m work copy per matrix.
void chol_drop_gpu(const double2* S, int m, double tol, int nev, double2* Wm, double2* Lf, int* keep, int* na,
                   int* ok, int batch, hipStream_t s);
// Linv = L^-1 for lower-triangular L (m <= 640)
void trinv_gpu(const double2* L, int m, double2* Linv, int batch, hipStream_t s);
// B = A^H
void conj_transpose_gpu(const double2* A, int m, double2* B, int batch, hipStream_t s);
// Hc = D H D on kept rows and columns, 0 elsewhere
void scale_mask_gpu(const double2* H, const double* dsc, const int* keep, int m, double2* Hc, int batch,
                    hipStream_t s);
// dropped diagonal entries of A <- a bound above every eigenvalue of the kept block
void big_diag_gpu(double2* A, const int* keep, int m, int batch, hipStream_t s);
// U (m x nev, ld m, offset kz*m*nev) = D C on kept rows, 0 on dropped rows;
// w (offset kz*nev) = wnew where ok, unchanged (and U = [I; 0]) where not.
void expand_gpu(const double2* C, const double* dsc, const int* keep, const int* ok, int m, int nev,
                const double* wnew, double* w, double2* U, int batch, hipStream_t s);

}  // namespace derek
