// FP64 dense subspace solvers on the GPU (see dense_gpu.hip).
#pragma once
#include "common.hpp"

namespace derek {

// Doubles of workspace per matrix for heev_gpu(n, nev).
int64_t heev_work_doubles(int n, int nev);
// Lowest nev eigenpairs of `batch` Hermitian n x n matrices (A destroyed;
// eigenvalues ascending in W, eigenvectors in the columns of Z).
void heev_gpu(double2* A, int lda, int64_t strideA, int n, int nev, double* W, int64_t strideW, double2* Z, int ldz,
              int64_t strideZ, double* work, int64_t strideWork, int batch, hipStream_t s);
// A <- (A + A^H)/2, m x m
void hermitize_gpu(double2* A, int m, hipStream_t s);
// dsc = 1/sqrt(diag S), S <- D S D (m <= 2048)
void hegv_gpu_prepare(double2* S, int m, double* dsc, hipStream_t s);
// Cholesky of the unit-diagonal S skipping dependent columns: act[0..*na) kept
// columns, Lbuild m x m row-major build buffer, Wm m x m work copy.
void chol_drop_gpu(const double2* S, int m, double tol, double2* Wm, double2* Lbuild, int* act, int* na,
                   hipStream_t s);
// Lc (na x na lower) and Hc = (D H D)[act, act]
void chol_compact_gpu(const double2* Lbuild, int m, const int* act, const int* na, const double2* H, int ldh,
                      const double* dsc, double2* Lc, double2* Hc, hipStream_t s);
void trinv_gpu(const double2* L, const int* na, int m, double2* Linv, hipStream_t s);
void conj_transpose_gpu(const double2* A, const int* na, int m, double2* B, hipStream_t s);
// C (m x nev) = 0, rows act[r] = dsc[act[r]] * Ca[r, :]
void expand_rows_gpu(const double2* Ca, const int* na, const int* act, const double* dsc, int m, int nev, double2* C,
                     hipStream_t s);

}  // namespace derek
