// FP64 subspace problems for a batch of matrices (see subspace.hpp).
//
// The generalized problem H c = w S c is reduced to standard form on the
// full m x m index set: the Cholesky factor with dropped columns replaced by
// the identity (Lf), its inverse, A = Lf^-1 (D H D)|kept Lf^-H with the
// dropped diagonal entries pushed above every kept eigenvalue.  The lowest
// nev eigenvectors of A are then those of the kept block, so every launch
// has a size known on the host and nothing waits for the kept count.
#include "dense_gpu.hpp"
#include "launch.hpp"
#include "subspace.hpp"

namespace derek {

static void* dalloc(size_t bytes) {
  void* p = nullptr;
  DEREK_HIP(hipMalloc(&p, std::max<size_t>(bytes, 16)));
  return p;
}

void DenseWork::release() {
  for (void* p : {(void*)Wm, (void*)Lf, (void*)Linv, (void*)LinvH, (void*)Hc, (void*)T, (void*)A, (void*)C, (void*)Z,
                  (void*)dsc, (void*)wtmp, (void*)hwork, (void*)keep, (void*)na, (void*)ok})
    if (p) (void)hipFree(p);
  Wm = Lf = Linv = LinvH = Hc = T = A = C = Z = nullptr;
  dsc = wtmp = hwork = nullptr;
  keep = na = ok = nullptr;
  m = nev = batch = 0;
}

void DenseWork::ensure(int m_, int nev_, int batch_) {
  if (m_ <= m && nev_ <= nev && batch_ <= batch) return;
  const int M = std::max(m_, m), V = std::max(nev_, nev), B = std::max(batch_, batch);
  release();
  m = M;
  nev = V;
  batch = B;
  const size_t mm = (size_t)M * M * B;
  for (double2** q : {&Wm, &Lf, &Linv, &LinvH, &Hc, &T, &A, &C, &Z}) *q = (double2*)dalloc(mm * sizeof(double2));
  dsc = (double*)dalloc((size_t)M * B * sizeof(double));
  wtmp = (double*)dalloc((size_t)M * B * sizeof(double));
  hstride = heev_work_doubles(M, M);
  hwork = (double*)dalloc((size_t)hstride * B * sizeof(double));
  keep = (int*)dalloc((size_t)M * B * sizeof(int));
  na = (int*)dalloc((size_t)B * sizeof(int));
  ok = (int*)dalloc((size_t)B * sizeof(int));
}

void rr_solve(DenseWork& d, double2* H, int m, int batch, double* w, double2* U, hipStream_t s) {
  DEREK_REQUIRE(m <= d.m && batch <= d.batch, "rr_solve: workspace too small");
  const int64_t mm = (int64_t)m * m;
  hermitize_gpu(H, m, batch, s);
  heev_gpu(H, m, mm, m, m, w, m, U, m, mm, d.hwork, d.hstride, batch, s);
}

void gen_solve(DenseWork& d, double2* S, double2* H, int m, int nev, int batch, double tol, double* w, double2* U,
               hipStream_t s) {
  DEREK_REQUIRE(m <= d.m && nev <= m && batch <= d.batch, "gen_solve: workspace too small");
  const int64_t mm = (int64_t)m * m;
  hermitize_gpu(S, m, batch, s);
  hermitize_gpu(H, m, batch, s);
  unit_diag_gpu(S, m, d.dsc, batch, s);
  chol_drop_gpu(S, m, tol, nev, d.Wm, d.Lf, d.keep, d.na, d.ok, batch, s);
  trinv_gpu(d.Lf, m, d.Linv, batch, s);
  scale_mask_gpu(H, d.dsc, d.keep, m, d.Hc, batch, s);
  rotate<double>(d.Linv, m, mm, d.Hc, m, mm, m, m, m, nullptr, d.T, m, mm, batch, s);  // T = L^-1 Hc
  conj_transpose_gpu(d.Linv, m, d.LinvH, batch, s);
  rotate<double>(d.T, m, mm, d.LinvH, m, mm, m, m, m, nullptr, d.A, m, mm, batch, s);  // A = T L^-H
  hermitize_gpu(d.A, m, batch, s);
  big_diag_gpu(d.A, d.keep, m, batch, s);
  heev_gpu(d.A, m, mm, m, nev, d.wtmp, nev, d.Z, m, (int64_t)m * nev, d.hwork, d.hstride, batch, s);
  // C = L^-H Z (m x nev per matrix, stride m*m)
  rotate<double>(d.LinvH, m, mm, d.Z, m, (int64_t)m * nev, m, nev, m, nullptr, d.C, m, mm, batch, s);
  expand_gpu(d.C, d.dsc, d.keep, d.ok, m, nev, d.wtmp, w, U, batch, s);
}

void orth_coeffs(DenseWork& d, double2* S, int m, int batch, double2* U, hipStream_t s) {
  DEREK_REQUIRE(m <= d.m && batch <= d.batch, "orth_coeffs: workspace too small");
  unit_diag_gpu(S, m, d.dsc, batch, s);
  chol_drop_gpu(S, m, 0.0, m, d.Wm, d.Lf, d.keep, d.na, d.ok, batch, s);
  trinv_gpu(d.Lf, m, d.Linv, batch, s);
  conj_transpose_gpu(d.Linv, m, d.LinvH, batch, s);
  expand_gpu(d.LinvH, d.dsc, d.keep, d.ok, m, m, d.wtmp, d.wtmp, U, batch, s);
}

}  // namespace derek
