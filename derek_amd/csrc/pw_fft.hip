// Sphere-pruned FFT pipelines for H|psi> and the density.
//
// The plane waves of one k-point occupy few z columns ("sticks", one per
// distinct (x, y) FFT bin) and few x planes of the grid.  The inverse
// transform runs z on the sticks only (the plane-wave scatter fused into the
// LDS load), y on the active x planes only (sticks scattered into y rows),
// x over every (y, z) line reading/writing only the active x rows; the
// forward transform mirrors it and fuses the gather and the kinetic term
// into the last z pass.  One launch covers a batch of (k-point, band)
// items; each k-point has its own sphere (maps), workgroups past a
// k-point's sticks or planes exit at once.  Buffers, per item:
//   S[item][stick][nz]           stick lines (nst_cap sticks per item)
//   P[item][active x][ny][nz]    active planes (nax_cap planes per item)
#include "launch.hpp"
#include "pw_fft.hpp"

namespace derek {

constexpr int PNT = 256;        // threads per workgroup
constexpr int PL = 16;          // z (or yz) columns per tile

// ---------------------------------------------------------------- z inverse on sticks
// psi[b][perm[p]] -> S[item][s][:], p over the plane waves of the block's sticks.
template <typename R, int NC>
__global__ __launch_bounds__(PNT) void k_sticks_z_inv(PwBatch B, PwBlock<R> psi, int lpb, FftPlan plan,
                                                      const cx_t<R>* __restrict__ tw, cx_t<R>* __restrict__ S) {
  using C = cx_t<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
  const int s0 = blockIdx.x * lpb;
  if (s0 >= m.nst) return;
  const int nz = NC > 0 ? NC : plan.n;
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + nz;
  C* b1 = b0 + (size_t)lpb * nz;
  for (int t = threadIdx.x; t < nz; t += PNT) twl[t] = tw[t];
  const int band = blockIdx.y;
  const int lines = min(lpb, m.nst - s0);
  const int ne = lines * nz;
  for (int e = threadIdx.x; e < ne; e += PNT) b0[e] = cmk<C>(0, 0);
  __syncthreads();
  const int p0 = m.st_pw[s0], p1 = m.st_pw[s0 + lines];
  const C* pb = psi.p + blockIdx.z * psi.ks + (int64_t)band * psi.ld;
  for (int p = p0 + threadIdx.x; p < p1; p += PNT) {
    const int line = m.pw_st[p] - s0;
    b0[line * nz + m.pw_z[p]] = pb[m.perm[p]];
  }
  __syncthreads();
  C* res = fft_lds<+1, false, 1, PNT, NC>(b0, b1, twl, plan, lines);
  const int64_t item = (int64_t)blockIdx.z * B.nbb + band;
  C* dst = S + (item * B.nst_cap + s0) * nz;
  for (int e = threadIdx.x; e < ne; e += PNT) dst[e] = res[e];
}

// ---------------------------------------------------------------- z forward on sticks + gather
// hpsi[b][perm[p]] = S-line FFT at pw_z[p] + kin[perm[p]] * psi[b][perm[p]]
template <typename R, int NC>
__global__ __launch_bounds__(PNT) void k_sticks_z_fwd(PwBatch B, const cx_t<R>* __restrict__ S, int lpb,
                                                      FftPlan plan, const cx_t<R>* __restrict__ tw, PwBlock<R> psi,
                                                      PwBlock<R> hpsi) {
  using C = cx_t<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
  const int s0 = blockIdx.x * lpb;
  if (s0 >= m.nst) return;
  const int nz = NC > 0 ? NC : plan.n;
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + nz;
  C* b1 = b0 + (size_t)lpb * nz;
  for (int t = threadIdx.x; t < nz; t += PNT) twl[t] = tw[t];
  const int band = blockIdx.y;
  const int lines = min(lpb, m.nst - s0);
  const int ne = lines * nz;
  const int64_t item = (int64_t)blockIdx.z * B.nbb + band;
  const C* src = S + (item * B.nst_cap + s0) * nz;
  for (int e = threadIdx.x; e < ne; e += PNT) b0[e] = src[e];
  __syncthreads();
  C* res = fft_lds<-1, false, 1, PNT, NC>(b0, b1, twl, plan, lines);
  const int p0 = m.st_pw[s0], p1 = m.st_pw[s0 + lines];
  const C* pb = psi.p + blockIdx.z * psi.ks + (int64_t)band * psi.ld;
  C* hb = hpsi.p + blockIdx.z * hpsi.ks + (int64_t)band * hpsi.ld;
  for (int p = p0 + threadIdx.x; p < p1; p += PNT) {
    const int line = m.pw_st[p] - s0;
    const int g = m.perm[p];
    const C v = res[line * nz + m.pw_z[p]];
    const C x = pb[g];
    const R t = (R)m.kin[g];
    hb[g] = cmk<C>(v.x + t * x.x, v.y + t * x.y);
  }
}

// ---------------------------------------------------------------- y passes on active planes
// Inverse: sticks of plane a -> full y rows of P[item][a][:, z0:z0+PL].
template <typename R, int NC>
__global__ __launch_bounds__(PNT) void k_planes_y_inv(PwBatch B, const cx_t<R>* __restrict__ S, int nz, int ztiles,
                                                      FftPlan plan, const cx_t<R>* __restrict__ tw,
                                                      cx_t<R>* __restrict__ P) {
  using C = cx_t<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
  const int a = blockIdx.x / ztiles;
  if (a >= m.nax) return;
  const int ny = NC > 0 ? NC : plan.n;
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + ny;
  C* b1 = b0 + (size_t)PL * ny;
  for (int t = threadIdx.x; t < ny; t += PNT) twl[t] = tw[t];
  const int band = blockIdx.y;
  const int z0 = (blockIdx.x - a * ztiles) * PL;
  const int lines = min(PL, nz - z0);
  const int ne = ny * PL;
  for (int e = threadIdx.x; e < ne; e += PNT) b0[e] = cmk<C>(0, 0);
  __syncthreads();
  const int64_t item = (int64_t)blockIdx.z * B.nbb + band;
  const C* Si = S + item * B.nst_cap * nz;
  const int sa = m.pl_st[a], sb = m.pl_st[a + 1];
  const int nput = (sb - sa) * PL;
  for (int e = threadIdx.x; e < nput; e += PNT) {
    const int j = sa + e / PL, l = e % PL;
    if (l < lines) b0[m.st_y[j] * PL + l] = Si[(int64_t)j * nz + z0 + l];
  }
  __syncthreads();
  C* res = fft_lds<+1, true, PL, PNT, NC>(b0, b1, twl, plan, PL);
  C* dst = P + ((item * B.nax_cap + a) * ny) * nz + z0;
  for (int e = threadIdx.x; e < ne; e += PNT) {
    const int y = e / PL, l = e - y * PL;
    if (l < lines) dst[(int64_t)y * nz + l] = res[e];
  }
}

// Forward: full y rows of P -> stick lines of plane a.
template <typename R, int NC>
__global__ __launch_bounds__(PNT) void k_planes_y_fwd(PwBatch B, const cx_t<R>* __restrict__ P, int nz, int ztiles,
                                                      FftPlan plan, const cx_t<R>* __restrict__ tw,
                                                      cx_t<R>* __restrict__ S) {
  using C = cx_t<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
  const int a = blockIdx.x / ztiles;
  if (a >= m.nax) return;
  const int ny = NC > 0 ? NC : plan.n;
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + ny;
  C* b1 = b0 + (size_t)PL * ny;
  for (int t = threadIdx.x; t < ny; t += PNT) twl[t] = tw[t];
  const int band = blockIdx.y;
  const int z0 = (blockIdx.x - a * ztiles) * PL;
  const int lines = min(PL, nz - z0);
  const int ne = ny * PL;
  const int64_t item = (int64_t)blockIdx.z * B.nbb + band;
  const C* src = P + ((item * B.nax_cap + a) * ny) * nz + z0;
  for (int e = threadIdx.x; e < ne; e += PNT) {
    const int y = e / PL, l = e - y * PL;
    b0[e] = l < lines ? src[(int64_t)y * nz + l] : cmk<C>(0, 0);
  }
  __syncthreads();
  C* res = fft_lds<-1, true, PL, PNT, NC>(b0, b1, twl, plan, PL);
  C* Si = S + item * B.nst_cap * nz;
  const int sa = m.pl_st[a], sb = m.pl_st[a + 1];
  const int nput = (sb - sa) * PL;
  for (int e = threadIdx.x; e < nput; e += PNT) {
    const int j = sa + e / PL, l = e % PL;
    if (l < lines) Si[(int64_t)j * nz + z0 + l] = res[m.st_y[j] * PL + l];
  }
}

// ---------------------------------------------------------------- x pass, V multiply
// For PL consecutive (y,z) columns: per band, load the active x rows,
// inverse FFT along x, multiply by V*vscale, forward FFT, store active rows.
// The V tile is loaded once for BPW bands of one k-point.
template <typename R, int NC, int BPW>
__global__ __launch_bounds__(PNT) void k_planes_x_vmul(PwBatch B, cx_t<R>* __restrict__ P, int64_t nyz,
                                                       FftPlan plan, const cx_t<R>* __restrict__ tw,
                                                       const R* __restrict__ vloc, R vscale) {
  using C = cx_t<R>;
  extern __shared__ __align__(16) unsigned char smem[];
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
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
  const int bb0 = blockIdx.y * BPW;
  const int bb1 = min(B.nbb, bb0 + BPW);
  const int nput = m.nax * PL;
  for (int band = bb0; band < bb1; ++band) {
    for (int e = threadIdx.x; e < ne; e += PNT) b0[e] = cmk<C>(0, 0);
    __syncthreads();
    const int64_t item = (int64_t)blockIdx.z * B.nbb + band;
    C* base = P + item * B.nax_cap * nyz + c0;
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
// rho[x*nyz + c] = keep*rho + sum_items occ[kz*occ_ks + b] |IFFT_x(active rows)|^2
template <typename R, int NC>
__global__ __launch_bounds__(PNT) void k_planes_x_density(PwBatch B, const cx_t<R>* __restrict__ P, int64_t nyz,
                                                          FftPlan plan, const cx_t<R>* __restrict__ tw,
                                                          const double* __restrict__ occ, int occ_ks,
                                                          double* __restrict__ rho, int keep) {
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
  for (int kz = 0; kz < B.nkc; ++kz) {
    const PwMaps m = B.mt[B.k0 + kz];
    const int nput = m.nax * PL;
    for (int band = 0; band < B.nbb; ++band) {
      const double w = occ[(int64_t)kz * occ_ks + band];
      if (w == 0.0) continue;
      __syncthreads();
      for (int e = threadIdx.x; e < ne; e += PNT) b0[e] = cmk<C>(0, 0);
      __syncthreads();
      const int64_t item = (int64_t)kz * B.nbb + band;
      const C* base = P + item * B.nax_cap * nyz + c0;
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
  if (bytes > 65536) allow_dynamic_lds(reinterpret_cast<const void*>(kern), bytes);
}

static void check_lds_pw(size_t bytes, int n) {
  DEREK_REQUIRE(bytes <= 160 * 1024, "FFT length " + std::to_string(n) + " exceeds the LDS budget");
}

static int sticks_per_block(int nz) { return std::max(1, 2048 / nz); }

static void check_batch(const PwBatch& B) {
  DEREK_REQUIRE(B.mt && B.nkc >= 1 && B.nbb >= 1 && B.nkc <= 65535 && B.nbb <= 65535, "empty pipeline batch");
}

#define PW_DISPATCH(n, CALL) \
  switch (n) {               \
    case 32: CALL(32); break; \
    case 72: CALL(72); break; \
    case 108: CALL(108); break; \
    case 144: CALL(144); break; \
    default: CALL(0); break; \
  }

template <typename R>
void sticks_z_inv(const PwBatch& B, PwBlock<R> psi, const Axis<R>* ax, int nz, cx_t<R>* S, hipStream_t s) {
  check_batch(B);
  const int lpb = sticks_per_block(nz);
  const size_t lds = sizeof(cx_t<R>) * ((size_t)nz + 2 * (size_t)lpb * nz);
  check_lds_pw(lds, nz);
  dim3 grid((unsigned)ceil_div(B.nst_cap, lpb), B.nbb, B.nkc);
#define C_(NC)                                                                                  \
  {                                                                                             \
    auto kern = k_sticks_z_inv<R, NC>;                                                          \
    allow_lds_pw(kern, lds);                                                                    \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, psi, lpb, ax[2].plan, ax[2].tw, S);    \
  }
  PW_DISPATCH(nz, C_)
#undef C_
  DEREK_HIP(hipGetLastError());
}

template <typename R>
void planes_y(const PwBatch& B, cx_t<R>* S, cx_t<R>* P, const Axis<R>* ax, int ny, int nz, int sign, hipStream_t s) {
  check_batch(B);
  const int ztiles = (int)ceil_div(nz, PL);
  const size_t lds = sizeof(cx_t<R>) * ((size_t)ny + 2 * (size_t)PL * ny);
  check_lds_pw(lds, ny);
  dim3 grid((unsigned)(B.nax_cap * ztiles), B.nbb, B.nkc);
  if (sign > 0) {
#define C_(NC)                                                                                 \
  {                                                                                            \
    auto kern = k_planes_y_inv<R, NC>;                                                         \
    allow_lds_pw(kern, lds);                                                                   \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, S, nz, ztiles, ax[1].plan, ax[1].tw, P); \
  }
    PW_DISPATCH(ny, C_)
#undef C_
  } else {
#define C_(NC)                                                                                 \
  {                                                                                            \
    auto kern = k_planes_y_fwd<R, NC>;                                                         \
    allow_lds_pw(kern, lds);                                                                   \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, P, nz, ztiles, ax[1].plan, ax[1].tw, S); \
  }
    PW_DISPATCH(ny, C_)
#undef C_
  }
  DEREK_HIP(hipGetLastError());
}

template <typename R>
void sticks_z_fwd(const PwBatch& B, const cx_t<R>* S, const Axis<R>* ax, int nz, PwBlock<R> psi, PwBlock<R> hpsi,
                  hipStream_t s) {
  check_batch(B);
  const int lpb = sticks_per_block(nz);
  const size_t lds = sizeof(cx_t<R>) * ((size_t)nz + 2 * (size_t)lpb * nz);
  check_lds_pw(lds, nz);
  dim3 grid((unsigned)ceil_div(B.nst_cap, lpb), B.nbb, B.nkc);
#define C_(NC)                                                                                       \
  {                                                                                                  \
    auto kern = k_sticks_z_fwd<R, NC>;                                                               \
    allow_lds_pw(kern, lds);                                                                         \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, S, lpb, ax[2].plan, ax[2].tw, psi, hpsi);   \
  }
  PW_DISPATCH(nz, C_)
#undef C_
  DEREK_HIP(hipGetLastError());
}

template <typename R, int BPW>
static void planes_x_vmul_bpw(const PwBatch& B, cx_t<R>* P, const Axis<R>* ax, int64_t nyz, const R* vloc,
                              R vscale, hipStream_t s) {
  const int nx = ax[0].n;
  const size_t lds = sizeof(cx_t<R>) * ((size_t)nx + 2 * (size_t)PL * nx) + sizeof(R) * (size_t)PL * nx;
  check_lds_pw(lds, nx);
  dim3 grid((unsigned)ceil_div(nyz, PL), (unsigned)ceil_div(B.nbb, BPW), B.nkc);
#define C_(NC)                                                                                        \
  {                                                                                                   \
    auto kern = k_planes_x_vmul<R, NC, BPW>;                                                          \
    allow_lds_pw(kern, lds);                                                                          \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, P, nyz, ax[0].plan, ax[0].tw, vloc, vscale); \
  }
  PW_DISPATCH(nx, C_)
#undef C_
}

template <typename R>
void planes_x_vmul(const PwBatch& B, cx_t<R>* P, const Axis<R>* ax, int64_t nyz, const R* vloc, R vscale,
                   int bands_per_wg, hipStream_t s) {
  check_batch(B);
  switch (bands_per_wg) {
    case 1: planes_x_vmul_bpw<R, 1>(B, P, ax, nyz, vloc, vscale, s); break;
    case 2: planes_x_vmul_bpw<R, 2>(B, P, ax, nyz, vloc, vscale, s); break;
    case 4: planes_x_vmul_bpw<R, 4>(B, P, ax, nyz, vloc, vscale, s); break;
    case 8: planes_x_vmul_bpw<R, 8>(B, P, ax, nyz, vloc, vscale, s); break;
    default: throw ArgError("bands per x-pass workgroup must be 1, 2, 4 or 8");
  }
  DEREK_HIP(hipGetLastError());
}

template <typename R>
void planes_x_density(const PwBatch& B, const cx_t<R>* P, const Axis<R>* ax, int64_t nyz, const double* occ,
                      int occ_ks, double* rho, int keep, hipStream_t s) {
  check_batch(B);
  const int nx = ax[0].n;
  const size_t lds = sizeof(cx_t<R>) * ((size_t)nx + 2 * (size_t)PL * nx) + sizeof(double) * (size_t)PL * nx;
  check_lds_pw(lds, nx);
  dim3 grid((unsigned)ceil_div(nyz, PL));
#define C_(NC)                                                                                                  \
  {                                                                                                             \
    auto kern = k_planes_x_density<R, NC>;                                                                      \
    allow_lds_pw(kern, lds);                                                                                    \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, P, nyz, ax[0].plan, ax[0].tw, occ, occ_ks, rho, keep); \
  }
  PW_DISPATCH(nx, C_)
#undef C_
  DEREK_HIP(hipGetLastError());
}

#define INST(R)                                                                                                   \
  template void sticks_z_inv<R>(const PwBatch&, PwBlock<R>, const Axis<R>*, int, cx_t<R>*, hipStream_t);         \
  template void planes_y<R>(const PwBatch&, cx_t<R>*, cx_t<R>*, const Axis<R>*, int, int, int, hipStream_t);      \
  template void sticks_z_fwd<R>(const PwBatch&, const cx_t<R>*, const Axis<R>*, int, PwBlock<R>, PwBlock<R>,      \
                                hipStream_t);                                                                     \
  template void planes_x_vmul<R>(const PwBatch&, cx_t<R>*, const Axis<R>*, int64_t, const R*, R, int, hipStream_t); \
  template void planes_x_density<R>(const PwBatch&, const cx_t<R>*, const Axis<R>*, int64_t, const double*, int,  \
                                    double*, int, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace derek
