// Sphere-pruned FFT pipelines (see pw_fft.hip), batched over k-points and bands.
#pragma once
#include <cstdint>
#include <vector>

#include "launch.hpp"
#include "pw_maps.hpp"

namespace derek {

// Wavefunction blocks: item (kz, b) of a launch is column b of k-point
// k0+kz at psi + kz*pks + b*ldp.
template <typename R> struct PwBlock {
  cx_t<R>* p = nullptr;
  int64_t ld = 0, ks = 0;
};

// z pass on sticks, inverse, with the plane-wave scatter: psi -> S.
// zmajor: S[item][z][stick] (for the fused y/x pass) instead of S[item][stick][z].
template <typename R>
void sticks_z_inv(const PwBatch& B, PwBlock<R> psi, const Axis<R>* ax, int nz, cx_t<R>* S, bool zmajor,
                  hipStream_t s);
// y pass on the active planes: sign +1 inverse S -> P, sign -1 forward P -> S;
// bands_per_wg (1, 2, 4, 8) bands of one k-point per workgroup.
template <typename R>
void planes_y(const PwBatch& B, cx_t<R>* S, cx_t<R>* P, const Axis<R>* ax, int ny, int nz, int sign,
              int bands_per_wg, hipStream_t s);
// z pass on sticks, forward, with the gather and the kinetic term:
// hpsi = FFT(S)[plane waves] + kin * psi.
template <typename R>
void sticks_z_fwd(const PwBatch& B, const cx_t<R>* S, const Axis<R>* ax, int nz, PwBlock<R> psi, PwBlock<R> hpsi,
                  bool zmajor, hipStream_t s);
// Fused y and x passes on z-major sticks, in place: y inverse, x inverse,
// * vt, x forward, y forward per z plane.  vt[(z*ny + y)*nx + x] is V / (nx ny nz);
// bands_per_wg bands of one k-point per workgroup.
template <typename R> bool slab_supported(int nx, int ny, int nz, int nax_cap, int nst_cap);
template <typename R>
void slab_vmul(const PwBatch& B, cx_t<R>* S, const Axis<R>* ax, int nz, const R* vt, int bands_per_wg, hipStream_t s);
// Fused density on z-major sticks: rho = keep*rho + sum_items occ[kz*occ_ks + b] |psi(r)|^2.
// The items are split into `groups` band groups per (z plane, k-point), each
// writing an FP64 partial plane into part[nkc*groups][N], summed in order.
template <typename R> int slab_density_groups(const PwBatch& B, const Axis<R>* ax, int nz);
template <typename R>
void slab_density(const PwBatch& B, const cx_t<R>* S, const Axis<R>* ax, int nz, const double* occ, int occ_ks,
                  double* part, int groups, double* rho, int keep, hipStream_t s);
// x pass: inverse, * vloc * vscale, forward, in place on the active rows of P.
template <typename R>
void planes_x_vmul(const PwBatch& B, cx_t<R>* P, const Axis<R>* ax, int64_t nyz, const R* vloc, R vscale,
                   int bands_per_wg, hipStream_t s);
// x pass: rho = keep*rho + sum_items occ[kz*occ_ks + b] |inverse x transform|^2.
template <typename R>
void planes_x_density(const PwBatch& B, const cx_t<R>* P, const Axis<R>* ax, int64_t nyz, const double* occ,
                      int occ_ks, double* rho, int keep, hipStream_t s);

}  // namespace derek
