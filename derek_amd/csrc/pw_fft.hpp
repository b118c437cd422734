// Plane-wave sphere maps for the pruned FFT pipelines (see pw_fft.hip).
#pragma once
#include <cstdint>
#include <vector>

#include "launch.hpp"
#include "pw_maps.hpp"

namespace derek {

// z pass on sticks, inverse, with the plane-wave scatter: psi -> S.
template <typename R>
void sticks_z_inv(const cx_t<R>* psi, int64_t ldp, int nb, const PwMaps& m, const Axis<R>* ax, int nz, cx_t<R>* S,
                  hipStream_t s);
// y pass on the active planes: sign +1 inverse S -> P, sign -1 forward P -> S_out.
template <typename R>
void planes_y(const cx_t<R>* S_in, cx_t<R>* P, int nb, const PwMaps& m, const Axis<R>* ax, int ny, int nz,
              int sign, cx_t<R>* S_out, hipStream_t s);
// z pass on sticks, forward, with the gather and the kinetic term:
// hpsi = FFT(S)[plane waves] + kin * psi.
template <typename R>
void sticks_z_fwd(const cx_t<R>* S, int nb, const PwMaps& m, const Axis<R>* ax, int nz, const cx_t<R>* psi,
                  int64_t ldp, const double* kin, cx_t<R>* hpsi, int64_t ldh, hipStream_t s);
// x pass: inverse, * vloc * vscale, forward, in place on the active rows of P.
template <typename R>
void planes_x_vmul(cx_t<R>* P, int nb, const PwMaps& m, const Axis<R>* ax, int64_t nyz, const R* vloc, R vscale,
                   hipStream_t s);
// x pass: rho = keep*rho + sum_b occ[b] |inverse x transform|^2 on the full grid.
template <typename R>
void planes_x_density(const cx_t<R>* P, int nb, const PwMaps& m, const Axis<R>* ax, int64_t nyz, const double* occ,
                      double* rho, int keep, hipStream_t s);

}  // namespace derek
