// Sphere-pruned FFT pipelines for H|psi> and the density.
//
// The plane waves of one k-point occupy few z columns ("sticks", one per
// distinct (x, y) FFT bin) and few x planes of the grid.  The inverse
// transform runs z on the sticks only (the plane-wave scatter fused into the
// LDS load), y on the active x planes only (sticks scattered into y rows),
// x over every (y, z) line reading/writing only the active x rows; the
// forward transform mirrors it and fuses the gather and the kinetic term
// into the last z pass.  Buffers:
//   S[band][stick][nz]           stick lines
//   P[band][active x][ny][nz]    active planes
#include "launch.hpp"
#include "pw_fft.hpp"

namespace derek {

constexpr int PNT = 256;        // threads per workgroup
constexpr int PL = 16;          // z (or yz) columns per tile
constexpr int BANDS_PER_WG = 4; // bands handled by one x-pass workgroup

// ---------------------------------------------------------------- z inverse on sticks
// psi[b*ldp + perm[p]] -> S[b][s][:], p over the plane waves of sticks s.
template <typename R, int NC>
__global__ __launch_bounds__(PNT) void k_sticks_z_inv(const cx_t<R>* __restrict__ psi, int64_t ldp,
                                                      PwMaps m, int lpb, FftPlan plan,
                                                      const cx_t<R>* __restrict__ tw, cx_t<R>* __restrict__ S) {
  using C = cx_t<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int nz = NC > 0 ? NC : plan.n;
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + nz;
  C* b1 = b0 + (size_t)lpb * nz;
  for (int t = threadIdx.x; t < nz; t += PNT) twl[t] = tw[t];
  const int band = blockIdx.y;
  const int s0 = blockIdx.x * lpb;
  const int lines = min(lpb, m.nst - s0);
  const int ne = lines * nz;
  for (int e = threadIdx.x; e < ne; e += PNT) b0[e] = cmk<C>(0, 0);
  __syncthreads();
  const int p0 = m.st_pw[s0], p1 = m.st_pw[s0 + lines];
  const C* pb = psi + (int64_t)band * ldp;
  for (int p = p0 + threadIdx.x; p < p1; p += PNT) {
    const int line = m.pw_st[p] - s0;
    b0[line * nz + m.pw_z[p]] = pb[m.perm[p]];
  }
  __syncthreads();
  C* res = fft_lds<+1, false, 1, PNT, NC>(b0, b1, twl, plan, lines);
  C* dst = S + ((int64_t)band * m.nst + s0) * nz;
  for (int e = threadIdx.x; e < ne; e += PNT) dst[e] = res[e];
}

// ---------------------------------------------------------------- z forward on sticks + gather
// hpsi[b*ldh + perm[p]] = S-line FFT at pw_z[p] + kin[perm[p]] * psi[...]
template <typename R, int NC>
__global__ __launch_bounds__(PNT) void k_sticks_z_fwd(const cx_t<R>* __restrict__ S, PwMaps m, int lpb,
                                                      FftPlan plan, const cx_t<R>* __restrict__ tw,
                                                      const cx_t<R>* __restrict__ psi, int64_t ldp,
                                                      const double* __restrict__ kin, cx_t<R>* __restrict__ hpsi,
                                                      int64_t ldh) {
  using C = cx_t<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int nz = NC > 0 ? NC : plan.n;
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + nz;
  C* b1 = b0 + (size_t)lpb * nz;
  for (int t = threadIdx.x; t < nz; t += PNT) twl[t] = tw[t];
  const int band = blockIdx.y;
  const int s0 = blockIdx.x * lpb;
  const int lines = min(lpb, m.nst - s0);
  const int ne = lines * nz;
  const C* src = S + ((int64_t)band * m.nst + s0) * nz;
  for (int e = threadIdx.x; e < ne; e += PNT) b0[e] = src[e];
  __syncthreads();
  C* res = fft_lds<-1, false, 1, PNT, NC>(b0, b1, twl, plan, lines);
  const int p0 = m.st_pw[s0], p1 = m.st_pw[s0 + lines];
  const C* pb = psi + (int64_t)band * ldp;
  C* hb = hpsi + (int64_t)band * ldh;
  for (int p = p0 + threadIdx.x; p < p1; p += PNT) {
    const int line = m.pw_st[p] - s0;
    const int g = m.perm[p];
    const C v = res[line * nz + m.pw_z[p]];
    const C x = pb[g];
    const R t = (R)kin[g];
    hb[g] = cmk<C>(v.x + t * x.x, v.y + t * x.y);
  }
}

// ---------------------------------------------------------------- y passes on active planes
// Inverse: sticks of plane a -> full y rows of P[band][a][:, z0:z0+PL].
template <typename R, int NC>
__global__ __launch_bounds__(PNT) void k_planes_y_inv(const cx_t<R>* __restrict__ S, PwMaps m, int nz, int ztiles,
                                                      FftPlan plan, const cx_t<R>* __restrict__ tw,
                                                      cx_t<R>* __restrict__ P) {
  using C = cx_t<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ny = NC > 0 ? NC : plan.n;
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + ny;
  C* b1 = b0 + (size_t)PL * ny;
  for (int t = threadIdx.x; t < ny; t += PNT) twl[t] = tw[t];
  const int band = blockIdx.y;
  const int a = blockIdx.x / ztiles;
  const int z0 = (blockIdx.x - a * ztiles) * PL;
  const int lines = min(PL, nz - z0);
  const int ne = ny * PL;
  for (int e = threadIdx.x; e < ne; e += PNT) b0[e] = cmk<C>(0, 0);
  __syncthreads();
  const int sa = m.pl_st[a], sb = m.pl_st[a + 1];
  const int nput = (sb - sa) * PL;
  for (int e = threadIdx.x; e < nput; e += PNT) {
    const int j = sa + e / PL, l = e % PL;
    if (l < lines) b0[m.st_y[j] * PL + l] = S[((int64_t)band * m.nst + j) * nz + z0 + l];
  }
  __syncthreads();
  C* res = fft_lds<+1, true, PL, PNT, NC>(b0, b1, twl, plan, PL);
  C* dst = P + (((int64_t)band * m.nax + a) * ny) * nz + z0;
  for (int e = threadIdx.x; e < ne; e += PNT) {
    const int y = e / PL, l = e - y * PL;
    if (l < lines) dst[(int64_t)y * nz + l] = res[e];
  }
}

// Forward: full y rows of P -> stick lines of plane a.
template <typename R, int NC>
__global__ __launch_bounds__(PNT) void k_planes_y_fwd(const cx_t<R>* __restrict__ P, PwMaps m, int nz, int ztiles,
                                                      FftPlan plan, const cx_t<R>* __restrict__ tw,
                                                      cx_t<R>* __restrict__ S) {
  using C = cx_t<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int ny = NC > 0 ? NC : plan.n;
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + ny;
  C* b1 = b0 + (size_t)PL * ny;
  for (int t = threadIdx.x; t < ny; t += PNT) twl[t] = tw[t];
  const int band = blockIdx.y;
  const int a = blockIdx.x / ztiles;
  const int z0 = (blockIdx.x - a * ztiles) * PL;
  const int lines = min(PL, nz - z0);
  const int ne = ny * PL;
  const C* src = P + (((int64_t)band * m.nax + a) * ny) * nz + z0;
  for (int e = threadIdx.x; e < ne; e += PNT) {
    const int y = e / PL, l = e - y * PL;
    b0[e] = l < lines ? src[(int64_t)y * nz + l] : cmk<C>(0, 0);
  }
  __syncthreads();
  C* res = fft_lds<-1, true, PL, PNT, NC>(b0, b1, twl, plan, PL);
  const int sa = m.pl_st[a], sb = m.pl_st[a + 1];
  const int nput = (sb - sa) * PL;
  for (int e = threadIdx.x; e < nput; e += PNT) {
    const int j = sa + e / PL, l = e % PL;
    if (l < lines) S[((int64_t)band * m.nst + j) * nz + z0 + l] = res[m.st_y[j] * PL + l];
  }
}

// ---------------------------------------------------------------- x pass, V multiply
// For PL consecutive (y,z) columns: per band, load the active x rows,
// inverse FFT along x, multiply by V*vscale, forward FFT, store active rows.
template <typename R, int NC>
__global__ __launch_bounds__(PNT) void k_planes_x_vmul(cx_t<R>* __restrict__ P, PwMaps m, int64_t nyz, int nbands,
                                                       FftPlan plan, const cx_t<R>* __restrict__ tw,
                                                       const R* __restrict__ vloc, R vscale) {
  using C = cx_t<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int nx = NC > 0 ? NC : plan.n;
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + nx;
  C* b1 = b0 + (size_t)PL * nx;
  R* vt = reinterpret_cast<R*>(b1 + (size_t)PL * nx);
  for (int t = threadIdx.x; t < nx; t += PNT) twl[t] = tw[t];
  const int64_t c0 = (int64_t)blockIdx.x * PL;
  const int lines = (int)min((int64_t)PL, nyz - c0);
  const int ne = nx * PL;
  for (int e = threadIdx.x; e < ne; e += PNT) {
    const int x = e / PL, l = e - x * PL;
    vt[e] = l < lines ? vloc[(int64_t)x * nyz + c0 + l] * vscale : (R)0;
  }
  const int bb0 = blockIdx.y * BANDS_PER_WG;
  const int bb1 = min(nbands, bb0 + BANDS_PER_WG);
  const int nput = m.nax * PL;
  for (int band = bb0; band < bb1; ++band) {
    for (int e = threadIdx.x; e < ne; e += PNT) b0[e] = cmk<C>(0, 0);
    __syncthreads();
    C* base = P + (int64_t)band * m.nax * nyz + c0;
    for (int e = threadIdx.x; e < nput; e += PNT) {
      const int a = e / PL, l = e - a * PL;
      if (l < lines) b0[m.ax_x[a] * PL + l] = base[(int64_t)a * nyz + l];
    }
    __syncthreads();
    C* res = fft_lds<+1, true, PL, PNT, NC>(b0, b1, twl, plan, PL);
    C* other = res == b0 ? b1 : b0;
    for (int e = threadIdx.x; e < ne; e += PNT) res[e] = cscale(res[e], vt[e]);
    __syncthreads();
    res = fft_lds<-1, true, PL, PNT, NC>(res, other, twl, plan, PL);
    for (int e = threadIdx.x; e < nput; e += PNT) {
      const int a = e / PL, l = e - a * PL;
      if (l < lines) base[(int64_t)a * nyz + l] = res[m.ax_x[a] * PL + l];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- x pass, density
// rho[x*nyz + c] = keep*rho + sum_b occ[b] |IFFT_x(active rows)|^2
template <typename R, int NC>
__global__ __launch_bounds__(PNT) void k_planes_x_density(const cx_t<R>* __restrict__ P, PwMaps m, int64_t nyz,
                                                          int nbands, FftPlan plan, const cx_t<R>* __restrict__ tw,
                                                          const double* __restrict__ occ, double* __restrict__ rho,
                                                          int keep) {
  using C = cx_t<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int nx = NC > 0 ? NC : plan.n;
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + nx;
  C* b1 = b0 + (size_t)PL * nx;
  double* acc = reinterpret_cast<double*>(b1 + (size_t)PL * nx);
  for (int t = threadIdx.x; t < nx; t += PNT) twl[t] = tw[t];
  const int64_t c0 = (int64_t)blockIdx.x * PL;
  const int lines = (int)min((int64_t)PL, nyz - c0);
  const int ne = nx * PL;
  for (int e = threadIdx.x; e < ne; e += PNT) acc[e] = 0.0;
  const int nput = m.nax * PL;
  for (int band = 0; band < nbands; ++band) {
    const double w = occ[band];
    if (w == 0.0) continue;
    __syncthreads();
    for (int e = threadIdx.x; e < ne; e += PNT) b0[e] = cmk<C>(0, 0);
    __syncthreads();
    const C* base = P + (int64_t)band * m.nax * nyz + c0;
    for (int e = threadIdx.x; e < nput; e += PNT) {
      const int a = e / PL, l = e - a * PL;
      if (l < lines) b0[m.ax_x[a] * PL + l] = base[(int64_t)a * nyz + l];
    }
    __syncthreads();
    C* res = fft_lds<+1, true, PL, PNT, NC>(b0, b1, twl, plan, PL);
    for (int e = threadIdx.x; e < ne; e += PNT) {
      const double re = (double)res[e].x, im = (double)res[e].y;
      acc[e] += w * (re * re + im * im);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < ne; e += PNT) {
    const int x = e / PL, l = e - x * PL;
    if (l < lines) {
      double* p = rho + (int64_t)x * nyz + c0 + l;
      *p = (keep ? *p : 0.0) + acc[e];
    }
  }
}

// ---------------------------------------------------------------- launchers
template <typename K> static void allow_lds_pw(K kern, size_t bytes) {
  if (bytes > 65536) DEREK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

static void check_lds_pw(size_t bytes, int n) {
  DEREK_REQUIRE(bytes <= 160 * 1024, "FFT length " + std::to_string(n) + " exceeds the LDS budget");
}

static int sticks_per_block(int nz) { return std::max(1, 2048 / nz); }

#define PW_DISPATCH(n, CALL) \
  switch (n) {               \
    case 32: CALL(32); break; \
    case 72: CALL(72); break; \
    case 108: CALL(108); break; \
    case 144: CALL(144); break; \
    default: CALL(0); break; \
  }

template <typename R>
void sticks_z_inv(const cx_t<R>* psi, int64_t ldp, int nb, const PwMaps& m, const Axis<R>* ax, int nz, cx_t<R>* S,
                  hipStream_t s) {
  const int lpb = sticks_per_block(nz);
  const size_t lds = sizeof(cx_t<R>) * ((size_t)nz + 2 * (size_t)lpb * nz);
  check_lds_pw(lds, nz);
  dim3 grid((unsigned)ceil_div(m.nst, lpb), nb);
#define C_(NC)                                                                                     \
  {                                                                                                \
    auto kern = k_sticks_z_inv<R, NC>;                                                             \
    allow_lds_pw(kern, lds);                                                                       \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, psi, ldp, m, lpb, ax[2].plan, ax[2].tw, S); \
  }
  PW_DISPATCH(nz, C_)
#undef C_
  DEREK_HIP(hipGetLastError());
}

template <typename R>
void planes_y(const cx_t<R>* S_in, cx_t<R>* P, int nb, const PwMaps& m, const Axis<R>* ax, int ny, int nz,
              int sign, cx_t<R>* S_out, hipStream_t s) {
  const int ztiles = (int)ceil_div(nz, PL);
  const size_t lds = sizeof(cx_t<R>) * ((size_t)ny + 2 * (size_t)PL * ny);
  check_lds_pw(lds, ny);
  dim3 grid((unsigned)(m.nax * ztiles), nb);
  if (sign > 0) {
#define C_(NC)                                                                                          \
  {                                                                                                     \
    auto kern = k_planes_y_inv<R, NC>;                                                                  \
    allow_lds_pw(kern, lds);                                                                            \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, S_in, m, nz, ztiles, ax[1].plan, ax[1].tw, P);   \
  }
    PW_DISPATCH(ny, C_)
#undef C_
  } else {
#define C_(NC)                                                                                          \
  {                                                                                                     \
    auto kern = k_planes_y_fwd<R, NC>;                                                                  \
    allow_lds_pw(kern, lds);                                                                            \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, P, m, nz, ztiles, ax[1].plan, ax[1].tw, S_out);  \
  }
    PW_DISPATCH(ny, C_)
#undef C_
  }
  DEREK_HIP(hipGetLastError());
}

template <typename R>
void sticks_z_fwd(const cx_t<R>* S, int nb, const PwMaps& m, const Axis<R>* ax, int nz, const cx_t<R>* psi,
                  int64_t ldp, const double* kin, cx_t<R>* hpsi, int64_t ldh, hipStream_t s) {
  const int lpb = sticks_per_block(nz);
  const size_t lds = sizeof(cx_t<R>) * ((size_t)nz + 2 * (size_t)lpb * nz);
  check_lds_pw(lds, nz);
  dim3 grid((unsigned)ceil_div(m.nst, lpb), nb);
#define C_(NC)                                                                                                    \
  {                                                                                                               \
    auto kern = k_sticks_z_fwd<R, NC>;                                                                            \
    allow_lds_pw(kern, lds);                                                                                      \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, S, m, lpb, ax[2].plan, ax[2].tw, psi, ldp, kin, hpsi, ldh); \
  }
  PW_DISPATCH(nz, C_)
#undef C_
  DEREK_HIP(hipGetLastError());
}

template <typename R>
void planes_x_vmul(cx_t<R>* P, int nb, const PwMaps& m, const Axis<R>* ax, int64_t nyz, const R* vloc, R vscale,
                   hipStream_t s) {
  const int nx = ax[0].n;
  const size_t lds = sizeof(cx_t<R>) * ((size_t)nx + 2 * (size_t)PL * nx) + sizeof(R) * (size_t)PL * nx;
  check_lds_pw(lds, nx);
  dim3 grid((unsigned)ceil_div(nyz, PL), (unsigned)ceil_div(nb, BANDS_PER_WG));
#define C_(NC)                                                                                            \
  {                                                                                                       \
    auto kern = k_planes_x_vmul<R, NC>;                                                                   \
    allow_lds_pw(kern, lds);                                                                              \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, P, m, nyz, nb, ax[0].plan, ax[0].tw, vloc, vscale); \
  }
  PW_DISPATCH(nx, C_)
#undef C_
  DEREK_HIP(hipGetLastError());
}

template <typename R>
void planes_x_density(const cx_t<R>* P, int nb, const PwMaps& m, const Axis<R>* ax, int64_t nyz, const double* occ,
                      double* rho, int keep, hipStream_t s) {
  const int nx = ax[0].n;
  const size_t lds = sizeof(cx_t<R>) * ((size_t)nx + 2 * (size_t)PL * nx) + sizeof(double) * (size_t)PL * nx;
  check_lds_pw(lds, nx);
  dim3 grid((unsigned)ceil_div(nyz, PL));
#define C_(NC)                                                                                               \
  {                                                                                                          \
    auto kern = k_planes_x_density<R, NC>;                                                                   \
    allow_lds_pw(kern, lds);                                                                                 \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, P, m, nyz, nb, ax[0].plan, ax[0].tw, occ, rho, keep); \
  }
  PW_DISPATCH(nx, C_)
#undef C_
  DEREK_HIP(hipGetLastError());
}

#define INST(R)                                                                                                   \
  template void sticks_z_inv<R>(const cx_t<R>*, int64_t, int, const PwMaps&, const Axis<R>*, int, cx_t<R>*,      \
                                hipStream_t);                                                                     \
  template void planes_y<R>(const cx_t<R>*, cx_t<R>*, int, const PwMaps&, const Axis<R>*, int, int, int, cx_t<R>*, \
                            hipStream_t);                                                                         \
  template void sticks_z_fwd<R>(const cx_t<R>*, int, const PwMaps&, const Axis<R>*, int, const cx_t<R>*, int64_t, \
                                const double*, cx_t<R>*, int64_t, hipStream_t);                                   \
  template void planes_x_vmul<R>(cx_t<R>*, int, const PwMaps&, const Axis<R>*, int64_t, const R*, R, hipStream_t); \
  template void planes_x_density<R>(const cx_t<R>*, int, const PwMaps&, const Axis<R>*, int64_t, const double*,   \
                                    double*, int, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace derek
