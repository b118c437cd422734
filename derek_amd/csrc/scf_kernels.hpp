// Device-side pieces of the batched SCF step (see scf_kernels.hip).
#pragma once
#include <cfloat>

#include "common.hpp"
#include "pw_maps.hpp"

namespace derek {

// Band columns of a batch of k-points: column b of k-point k0+kz at
// p + kz*ks + b*ld.
template <typename C> struct BandBlock {
  C* p = nullptr;
  int64_t ld = 0, ks = 0;
};

// Slots of the device-side step report.
enum ReportSlot {
  REP_EBAND = 0,
  REP_EH,
  REP_EXC,
  REP_EDC,
  REP_MTS,
  REP_MU,
  REP_DRHO,
  REP_MAXRES,
  REP_COUNT
};

int band_stats_chunks(int64_t npw);
// ekin[k*nb+b] = sum kin |x|^2, rn[k*nb+b] = |hx - eps x| for k = k0..k0+nkc-1;
// partial: nkc*nb*band_stats_chunks(max_npw)*2 doubles.
template <typename R>
void band_stats(BandBlock<cx_t<R>> X, BandBlock<cx_t<R>> HX, const PwMaps* mt, int k0, int nkc, int64_t max_npw,
                const double* eps, int nb, double* partial, double* ekin, double* rn, hipStream_t s);
// W = TPA-preconditioned residuals, zero columns where rn <= drop_tol.
template <typename R>
void tpa_batch(BandBlock<cx_t<R>> X, BandBlock<cx_t<R>> HX, BandBlock<cx_t<R>> W, const PwMaps* mt, int k0, int nkc,
               int64_t max_npw, const double* eps, const double* ekin, const double* rn, int nb, double drop_tol,
               hipStream_t s);
template <typename R>
void init_psi_batch(BandBlock<cx_t<R>> X, const PwMaps* mt, int k0, int nkc, int64_t max_npw, int nb, uint64_t seed,
                    hipStream_t s);
// Fermi-Dirac (kT > 0) or insulating (kT = 0) occupations of nk*nb
// eigenvalues; writes occ, occw = w_k occ/volume and report slots MU, MTS, EBAND.
void occupations(const double* eps, const double* wk, int nk, int nb, double nelec, double kT, double volume,
                 double* occ, double* occw, double* rep, hipStream_t s);
// rep[slot] = scale * sum(partial[0..n)) (sqrt of it if root)
void finish_sum(const double* partial, int n, double scale, bool root, double* rep, int slot, hipStream_t s);
// rep[slot] = max(x[0..n))
void finish_max(const double* x, int n, double* rep, int slot, hipStream_t s);
// Pulay step on device: push (rho_in, rho_out - rho_in) into the ring, update
// the residual overlaps A (H x H), solve for the coefficients and form
// r_opt = sum c_i rho_in,i and f_opt = sum c_i F_i.  st[0] counts pushes.
// S(X,X) = I, H(X,X) = diag(eps), lower-left blocks mirrored (Y = [X, R]).
void subspace_blocks(double2* S, double2* H, const double* eps, int m2, int nb, int nkc, hipStream_t s);

// partial: H * pulay_blocks(n) doubles.
int pulay_blocks(int64_t n);
void pulay_update(int* st, int H, const double* rin, const double* rout, double* hin, double* hres, double* A,
                  double* coef, double* partial, double* r_opt, double* f_opt, int64_t n, hipStream_t s);

}  // namespace derek
