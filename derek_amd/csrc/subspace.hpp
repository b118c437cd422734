// FP64 subspace problems of the block Davidson for a batch of k-points,
// composed from the dense GPU kernels (dense_gpu.hip) and the MFMA rotation.
#pragma once
#include "common.hpp"

namespace derek {

// Scratch for batches of up to `batch` m x m problems with nev wanted pairs.
struct DenseWork {
  int m = 0, nev = 0, batch = 0;
  double2 *Wm = nullptr, *Lf = nullptr, *Linv = nullptr, *LinvH = nullptr, *Hc = nullptr, *T = nullptr,
          *A = nullptr, *C = nullptr, *Z = nullptr;
  double *dsc = nullptr, *wtmp = nullptr, *hwork = nullptr;
  int *keep = nullptr, *na = nullptr, *ok = nullptr;
  int64_t hstride = 0;  // doubles of heev workspace per matrix
  DenseWork() = default;
  DenseWork(const DenseWork&) = delete;
  DenseWork& operator=(const DenseWork&) = delete;
  ~DenseWork() { release(); }
  void ensure(int m, int nev, int batch);
  void release();
};

// Rayleigh-Ritz: eigenpairs of the Hermitian m x m matrices H (destroyed):
// w (batch x m, offset kz*m), U (m x m, offset kz*m*m).
void rr_solve(DenseWork& d, double2* H, int m, int batch, double* w, double2* U, hipStream_t s);
// Lowest nev solutions of H c = w S c (S, H: m x m per matrix, destroyed),
// columns of S dependent on earlier ones (pivot <= tol after scaling to a
// unit diagonal) dropped: U (m x nev, offset kz*m*nev) = coefficients with
// zero rows at dropped columns; w (offset kz*nev) updated where the problem
// kept at least nev columns (d.ok), U = [I; 0] elsewhere.
void gen_solve(DenseWork& d, double2* S, double2* H, int m, int nev, int batch, double tol, double* w, double2* U,
               hipStream_t s);
// Orthonormalising coefficients U = D L^-H of the overlap S = X^H X
// (m x m per matrix, destroyed); d.ok[kz] = 0 where X is rank deficient.
void orth_coeffs(DenseWork& d, double2* S, int m, int batch, double2* U, hipStream_t s);

}  // namespace derek
