// FP64 dense linear algebra on the host for the small subspace problems
// (n = bands or 2*bands): Cholesky, triangular inverse, Hermitian
// eigensolver (Householder tridiagonalisation + implicit QL with Wilkinson
// shifts), generalised Hermitian-definite eigenproblem.
// Matrices are column-major: A[i + j*lda].
#pragma once
#include <complex>
#include <vector>

namespace derek {
namespace dense {

using cd = std::complex<double>;

// A = L L^H; the lower triangle of A is overwritten by L, the strict upper
// triangle is zeroed.  Returns false if A is not numerically positive definite.
bool cholesky(int n, cd* A, int lda);
// In place inverse of a lower-triangular matrix.
void inv_lower(int n, cd* L, int lda);
// Eigenvalues ascending in w, orthonormal eigenvectors in the columns of Z.
// A is destroyed.  Throws derek::Error if QL does not converge.
void heev(int n, cd* A, int lda, double* w, cd* Z, int ldz);
// Lowest nev eigenpairs only (Z: n x nev).
void heev_lowest(int n, cd* A, int lda, int nev, double* w, cd* Z, int ldz);
// Lowest nev solutions of H c = lambda S c with C^H S C = I.  S and H are
// destroyed.  Returns false if S is not positive definite.
bool hegv_lowest(int n, cd* S, cd* H, int ld, int nev, double* w, cd* C, int ldc);
// Columns kept by a Cholesky of the unit-diagonal overlap S that skips every
// column whose remaining pivot is below tol (dependent on earlier kept columns).
std::vector<int> cholesky_active(int n, const cd* S, int lda, double tol);
// A <- (A + A^H)/2
void hermitize(int n, cd* A, int lda);

}  // namespace dense
}  // namespace derek
