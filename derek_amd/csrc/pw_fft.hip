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
  const int ls = contig_stride(nz);
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + nz;
  C* b1 = b0 + (size_t)lpb * ls;
  for (int t = threadIdx.x; t < nz; t += PNT) twl[t] = tw[t];
  const int band = blockIdx.y;
  const int lines = min(lpb, m.nst - s0);
  const int ne = lines * nz;
  for (int e = threadIdx.x; e < lines * ls; e += PNT) b0[e] = cmk<C>(0, 0);
  __syncthreads();
  const int p0 = m.st_pw[s0], p1 = m.st_pw[s0 + lines];
  const C* pb = psi.p + blockIdx.z * psi.ks + (int64_t)band * psi.ld;
  for (int p = p0 + threadIdx.x; p < p1; p += PNT) {
    const int line = m.pw_st[p] - s0;
    b0[line * ls + m.pw_z[p]] = pb[m.perm[p]];
  }
  __syncthreads();
  C* res = fft_lds<+1, false, 1, PNT, NC>(b0, b1, twl, plan, lines);
  const int64_t item = (int64_t)blockIdx.z * B.nbb + band;
  C* dst = S + (item * B.nst_cap + s0) * nz;
  for (int e = threadIdx.x; e < ne; e += PNT) {
    const int line = e / nz;
    dst[e] = res[line * ls + (e - line * nz)];
  }
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
  const int ls = contig_stride(nz);
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + nz;
  C* b1 = b0 + (size_t)lpb * ls;
  for (int t = threadIdx.x; t < nz; t += PNT) twl[t] = tw[t];
  const int band = blockIdx.y;
  const int lines = min(lpb, m.nst - s0);
  const int ne = lines * nz;
  const int64_t item = (int64_t)blockIdx.z * B.nbb + band;
  const C* src = S + (item * B.nst_cap + s0) * nz;
  for (int e = threadIdx.x; e < ne; e += PNT) {
    const int line = e / nz;
    b0[line * ls + (e - line * nz)] = src[e];
  }
  __syncthreads();
  C* res = fft_lds<-1, false, 1, PNT, NC>(b0, b1, twl, plan, lines);
  const int p0 = m.st_pw[s0], p1 = m.st_pw[s0 + lines];
  const C* pb = psi.p + blockIdx.z * psi.ks + (int64_t)band * psi.ld;
  C* hb = hpsi.p + blockIdx.z * hpsi.ks + (int64_t)band * hpsi.ld;
  for (int p = p0 + threadIdx.x; p < p1; p += PNT) {
    const int line = m.pw_st[p] - s0;
    const int g = m.perm[p];
    const C v = res[line * ls + m.pw_z[p]];
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

// ---------------------------------------------------------------- two-stage column passes
// For the compile-time lengths N = R1*R2 (two-stage plans) the y and x
// passes keep the transform in registers between the stages and use one LDS
// buffer (16 columns x N) for the single transpose each transform needs:
//   S1: thread (l, j1 < R2): v_r = in[j1 + R2 r] (r < R1), DFT_R1 -> Y[j1 R1 + k]
//   S2: thread (l, j2 < R1): u_r = Y[j2 + R1 r] w^(j2 r), DFT_R2 -> X[j2 + R1 k]
// S1 reads straight from global memory (through the sticks' or active
// planes' index map) and S2 writes straight to it.  The x pass multiplies X
// by V in registers and runs the forward transform by decimation in
// frequency on the same threads:
//   A: T[j2][k2] = w^(-j2 k2) DFT_R2(X[j2 + R1 r]) -> LDS
//   B: thread (l, k2 < R2): F[k2 + R2 k1] = DFT_R1 over j2 of T[j2][k2]
// V and the stage twiddles of each thread are loaded once; the inputs of
// the next band are loaded while the current one is transformed.
// Inputs of the next band are prefetched into registers only for complex64
// (for complex128 the extra registers would halve the waves per SIMD).
template <typename R> constexpr bool kPrefetch = sizeof(R) == 4;

template <int N> struct Split {
  static constexpr FftPlan p = make_plan(N);
  static_assert(p.ns == 2, "two-stage plan expected");
  static constexpr int R1 = p.r[0], R2 = p.r[1];
  static constexpr int NT = ((PL * (R1 > R2 ? R1 : R2) + 63) / 64) * 64;
  static constexpr size_t lds(size_t csize) { return csize * (size_t)N * (PL + 1) + sizeof(int) * N; }
};

template <typename C, int N>
__device__ __forceinline__ void load_twiddles(C* twl, const C* __restrict__ tw) {
  for (int t = threadIdx.x; t < N; t += Split<N>::NT) twl[t] = tw[t];
}

// x pass: inverse x transform of the active rows, * V * vscale, forward
// transform, active rows written back; BPW bands of one k-point per block
// (V is loaded once per block).
// Waves per SIMD the x pass is compiled for (register budget per lane
// 512 / waves): the most that each length fits without spilling.
template <typename R, int N> constexpr int kXWaves = (sizeof(R) == 8 ? 4 : 5) - (N >= 144 ? 1 : 0);

template <typename R, int N, int BPW>
__global__ __launch_bounds__(Split<N>::NT) __attribute__((amdgpu_waves_per_eu(kXWaves<R, N>))) void k_px_vmul2(PwBatch B, cx_t<R>* __restrict__ P, int64_t nyz,
                                                          const cx_t<R>* __restrict__ tw,
                                                          const R* __restrict__ vloc, R vscale) {
  using C = cx_t<R>;
  using SP = Split<N>;
  constexpr int R1 = SP::R1, R2 = SP::R2, NT = SP::NT;
  extern __shared__ __align__(16) unsigned char smem[];
  C* twl = reinterpret_cast<C*>(smem);
  C* buf = twl + N;
  uint32_t* inv = reinterpret_cast<uint32_t*>(buf + N * PL);  // byte offset of row x in a band, or kOob
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
  const int tid = threadIdx.x, l = tid % PL, jj = tid / PL;
  const int64_t c0 = (int64_t)blockIdx.x * PL;
  const bool lok = c0 + l < nyz;
  load_twiddles<C, N>(twl, tw);
  for (int x = tid; x < N; x += NT) inv[x] = kOob;
  __syncthreads();
  for (int a = tid; a < m.nax; a += NT) inv[m.ax_x[a]] = (uint32_t)a * (uint32_t)nyz * (uint32_t)sizeof(C);
  __syncthreads();
  // V at x = j2 + R1 r: held in registers across the band loop, or, where
  // that would spill (complex128 at N = 144), re-read per band (an L2 hit)
  constexpr bool kVregs = !(sizeof(R) == 8 && N >= 144);
  R vr[kVregs ? R2 : 1];
  if (kVregs && jj < R1) {
#pragma unroll
    for (int r = 0; r < R2; ++r) vr[r] = lok ? vloc[(int64_t)(jj + R1 * r) * nyz + c0 + l] * vscale : (R)0;
  }
  const int bb0 = blockIdx.y * BPW;
  const int bb1 = min(B.nbb, bb0 + BPW);
  // S1 input rows are looked up in LDS per band (no index registers held
  // across the band loop)
  // lane offset within a row; kOob makes every access of a lane past the
  // last column (and of the S1 rows j >= R2) fall outside the buffer
  const uint32_t lo = lok ? (uint32_t)l * (uint32_t)sizeof(C) : kOob;
  const uint32_t band_bytes = (uint32_t)((B.nax_cap * nyz - c0) * (int64_t)sizeof(C));
  auto band_rsrc = [&](int band) {
    return buf_rsrc(P + ((int64_t)blockIdx.z * B.nbb + band) * B.nax_cap * nyz + c0, band_bytes);
  };
  auto load_s1 = [&](int band, C* v) {
    const auto rs = band_rsrc(band);
    int j = jj;
    asm volatile("" : "+v"(j));
    const int jr = j < R2 ? j : 0;
    const uint32_t lj = j < R2 ? lo : kOob;
#pragma unroll
    // the planes stream through this pass (read once, written once): measured
    // 5 % faster at 108^3 with the non-temporal hint, the y passes around it too
    for (int r = 0; r < R1; ++r) v[r] = buf_ld<C, kNt>(rs, inv[jr + R2 * r] + lj);
  };
  C pre[kPrefetch<R> ? R1 : 1];
  if (kPrefetch<R> && bb0 < bb1) load_s1(bb0, pre);
  for (int band = bb0; band < bb1; ++band) {
    C v[R1];
    if constexpr (kPrefetch<R>) {
#pragma unroll
      for (int r = 0; r < R1; ++r) v[r] = pre[r];
      if (band + 1 < bb1) load_s1(band + 1, pre);
    } else {
      load_s1(band, v);
    }
    if (jj < R2) {
      dftR<R1, +1>(v, twl, N);
#pragma unroll
      for (int k = 0; k < R1; ++k) buf[(jj * R1 + k) * PL + l] = v[k];
    }
    __syncthreads();
    C u[R2];
    // an opaque copy of jj: the twiddle, V and row-index loads below are
    // redone per band instead of being hoisted into registers held across
    // the band loop
    int jx = jj;
    asm volatile("" : "+v"(jx));
    if (jx < R1) {
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const C y = buf[(jx + R1 * r) * PL + l];
        u[r] = r ? cmul(y, cconj(twl[(jx * r) % N])) : y;
      }
      dftR<R2, +1>(u, twl, N);
#pragma unroll
      for (int r = 0; r < R2; ++r) u[r] = cscale(u[r], kVregs ? vr[kVregs ? r : 0] : (lok ? vloc[(int64_t)(jx + R1 * r) * nyz + c0 + l] * vscale : (R)0));
      dftR<R2, -1>(u, twl, N);
#pragma unroll
      for (int k2 = 1; k2 < R2; ++k2) u[k2] = cmul(u[k2], twl[(jx * k2) % N]);
    }
    __syncthreads();
    if (jj < R1) {
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) buf[(jj * R2 + k2) * PL + l] = u[k2];
    }
    __syncthreads();
    if (jj < R2) {
      C t[R1];
#pragma unroll
      for (int j2 = 0; j2 < R1; ++j2) t[j2] = buf[(j2 * R2 + jj) * PL + l];
      dftR<R1, -1>(t, twl, N);
      const auto rs = band_rsrc(band);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) buf_st<C, kNt>(rs, inv[jx + R2 * k1] + lo, t[k1]);
    }
    __syncthreads();
  }
}

// x pass for the density: rho[x*nyz + c] = keep*rho + sum_items occ |IFFT_x|^2,
// accumulated in registers (thread (l, j2) owns x = j2 + R1 r).
template <typename R, int N>
__global__ __launch_bounds__(Split<N>::NT) void k_px_density2(PwBatch B, const cx_t<R>* __restrict__ P, int64_t nyz,
                                                             const cx_t<R>* __restrict__ tw,
                                                             const double* __restrict__ occ, int occ_ks,
                                                             double* __restrict__ rho, int keep) {
  using C = cx_t<R>;
  using SP = Split<N>;
  constexpr int R1 = SP::R1, R2 = SP::R2, NT = SP::NT;
  extern __shared__ __align__(16) unsigned char smem[];
  C* twl = reinterpret_cast<C*>(smem);
  C* buf = twl + N;
  int* inv = reinterpret_cast<int*>(buf + N * PL);
  const int tid = threadIdx.x, l = tid % PL, jj = tid / PL;
  const int64_t c0 = (int64_t)blockIdx.x * PL;
  const bool lok = c0 + l < nyz;
  load_twiddles<C, N>(twl, tw);
  double acc[R2];
#pragma unroll
  for (int r = 0; r < R2; ++r) acc[r] = 0.0;
  for (int kz = 0; kz < B.nkc; ++kz) {
    const PwMaps m = B.mt[B.k0 + kz];
    __syncthreads();
    for (int x = tid; x < N; x += NT) inv[x] = (int)kOob;
    __syncthreads();
    for (int a = tid; a < m.nax; a += NT) inv[m.ax_x[a]] = a * (int)nyz * (int)sizeof(C);
    __syncthreads();
    // byte offsets of this thread's rows in a band (kOob: inactive row or lane)
    uint32_t arow[R1];
#pragma unroll
    for (int r = 0; r < R1; ++r) arow[r] = (jj < R2 && lok) ? (uint32_t)inv[jj + R2 * r] + (uint32_t)l * sizeof(C) : kOob;
    const uint32_t band_bytes = (uint32_t)((B.nax_cap * nyz - c0) * (int64_t)sizeof(C));
    auto load_s1 = [&](int band, C* v) {
      const auto rs = buf_rsrc(P + ((int64_t)kz * B.nbb + band) * B.nax_cap * nyz + c0, band_bytes);
#pragma unroll
      for (int r = 0; r < R1; ++r) v[r] = buf_ld<C>(rs, arow[r]);
    };
    C pre[kPrefetch<R> ? R1 : 1];
    if (kPrefetch<R>) load_s1(0, pre);
    for (int band = 0; band < B.nbb; ++band) {
      const double w = occ[(int64_t)kz * occ_ks + band];
      C v[R1];
      if constexpr (kPrefetch<R>) {
#pragma unroll
        for (int r = 0; r < R1; ++r) v[r] = pre[r];
        if (band + 1 < B.nbb) load_s1(band + 1, pre);
      } else if (w != 0.0) {
        load_s1(band, v);
      }
      if (w == 0.0) continue;  // uniform over the block
      if (jj < R2) {
        dftR<R1, +1>(v, twl, N);
#pragma unroll
        for (int k = 0; k < R1; ++k) buf[(jj * R1 + k) * PL + l] = v[k];
      }
      __syncthreads();
      if (jj < R1) {
        C u[R2];
#pragma unroll
        for (int r = 0; r < R2; ++r) {
          const C y = buf[(jj + R1 * r) * PL + l];
          u[r] = r ? cmul(y, cconj(twl[(jj * r) % N])) : y;
        }
        dftR<R2, +1>(u, twl, N);
#pragma unroll
        for (int r = 0; r < R2; ++r) {
          const double re = (double)u[r].x, im = (double)u[r].y;
          acc[r] += w * (re * re + im * im);
        }
      }
      __syncthreads();
    }
  }
  if (jj < R1 && lok) {
#pragma unroll
    for (int r = 0; r < R2; ++r) {
      double* p = rho + (int64_t)(jj + R1 * r) * nyz + c0 + l;
      *p = (keep ? *p : 0.0) + acc[r];
    }
  }
}

// y pass, inverse: sticks of plane a (zero rows elsewhere) -> full y rows of P.
template <typename R, int N, int BPW>
__global__ __launch_bounds__(Split<N>::NT) void k_py_inv2(PwBatch B, const cx_t<R>* __restrict__ S, int nz,
                                                         int ztiles, const cx_t<R>* __restrict__ tw,
                                                         cx_t<R>* __restrict__ P) {
  using C = cx_t<R>;
  using SP = Split<N>;
  constexpr int R1 = SP::R1, R2 = SP::R2, NT = SP::NT;
  extern __shared__ __align__(16) unsigned char smem[];
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
  const int a = blockIdx.x / ztiles;
  if (a >= m.nax) return;
  C* twl = reinterpret_cast<C*>(smem);
  C* buf = twl + N;
  int* inv = reinterpret_cast<int*>(buf + N * PL);
  const int tid = threadIdx.x, l = tid % PL, jj = tid / PL;
  const int z0 = (blockIdx.x - a * ztiles) * PL;
  const bool lok = z0 + l < nz;
  load_twiddles<C, N>(twl, tw);
  for (int y = tid; y < N; y += NT) inv[y] = (int)kOob;
  __syncthreads();
  const int sa = m.pl_st[a], sb = m.pl_st[a + 1];
  for (int j = sa + tid; j < sb; j += NT) inv[m.st_y[j]] = j * nz * (int)sizeof(C);
  __syncthreads();
  // byte offsets of this thread's stick lines in a band (kOob: no stick)
  const uint32_t lo = lok ? (uint32_t)l * (uint32_t)sizeof(C) : kOob;
  uint32_t srow[R1];
#pragma unroll
  for (int r = 0; r < R1; ++r) srow[r] = jj < R2 ? (uint32_t)inv[jj + R2 * r] + lo : kOob;
  const int bb0 = blockIdx.y * BPW;
  const int bb1 = min(B.nbb, bb0 + BPW);
  const uint32_t st_bytes = (uint32_t)(((int64_t)B.nst_cap * nz - z0) * (int64_t)sizeof(C));
  auto load_s1 = [&](int band, C* v) {
    const auto rs = buf_rsrc(S + ((int64_t)blockIdx.z * B.nbb + band) * B.nst_cap * nz + z0, st_bytes);
#pragma unroll
    for (int r = 0; r < R1; ++r) v[r] = buf_ld<C>(rs, srow[r]);
  };
  C pre[kPrefetch<R> ? R1 : 1];
  if (kPrefetch<R> && bb0 < bb1) load_s1(bb0, pre);
  for (int band = bb0; band < bb1; ++band) {
    C v[R1];
    if constexpr (kPrefetch<R>) {
#pragma unroll
      for (int r = 0; r < R1; ++r) v[r] = pre[r];
      if (band + 1 < bb1) load_s1(band + 1, pre);
    } else {
      load_s1(band, v);
    }
    if (jj < R2) {
      dftR<R1, +1>(v, twl, N);
#pragma unroll
      for (int k = 0; k < R1; ++k) buf[(jj * R1 + k) * PL + l] = v[k];
    }
    __syncthreads();
    if (jj < R1) {
      C u[R2];
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const C y = buf[(jj + R1 * r) * PL + l];
        u[r] = r ? cmul(y, cconj(twl[(jj * r) % N])) : y;
      }
      dftR<R2, +1>(u, twl, N);
      const auto rd = buf_rsrc(P + (((int64_t)blockIdx.z * B.nbb + band) * B.nax_cap + a) * N * nz + z0,
                               (uint32_t)(((int64_t)N * nz - z0) * (int64_t)sizeof(C)));
#pragma unroll
      for (int k = 0; k < R2; ++k) buf_st<C>(rd, (uint32_t)((jj + R1 * k) * nz) * (uint32_t)sizeof(C) + lo, u[k]);
    }
    __syncthreads();
  }
}

// y pass, forward: full y rows of P -> stick lines of plane a.
template <typename R, int N, int BPW>
__global__ __launch_bounds__(Split<N>::NT) void k_py_fwd2(PwBatch B, const cx_t<R>* __restrict__ P, int nz,
                                                         int ztiles, const cx_t<R>* __restrict__ tw,
                                                         cx_t<R>* __restrict__ S) {
  using C = cx_t<R>;
  using SP = Split<N>;
  constexpr int R1 = SP::R1, R2 = SP::R2, NT = SP::NT;
  extern __shared__ __align__(16) unsigned char smem[];
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
  const int a = blockIdx.x / ztiles;
  if (a >= m.nax) return;
  C* twl = reinterpret_cast<C*>(smem);
  C* buf = twl + N;
  int* inv = reinterpret_cast<int*>(buf + N * PL);
  const int tid = threadIdx.x, l = tid % PL, jj = tid / PL;
  const int z0 = (blockIdx.x - a * ztiles) * PL;
  const bool lok = z0 + l < nz;
  load_twiddles<C, N>(twl, tw);
  for (int y = tid; y < N; y += NT) inv[y] = (int)kOob;
  __syncthreads();
  const int sa = m.pl_st[a], sb = m.pl_st[a + 1];
  for (int j = sa + tid; j < sb; j += NT) inv[m.st_y[j]] = j * nz * (int)sizeof(C);
  __syncthreads();
  // byte offsets of this thread's output stick lines (kOob: no stick)
  const uint32_t lo = lok ? (uint32_t)l * (uint32_t)sizeof(C) : kOob;
  uint32_t srow[R2];
  if (jj < R1) {
#pragma unroll
    for (int r = 0; r < R2; ++r) srow[r] = (uint32_t)inv[jj + R1 * r] + lo;
  }
  const int bb0 = blockIdx.y * BPW;
  const int bb1 = min(B.nbb, bb0 + BPW);
  const uint32_t pl_bytes = (uint32_t)(((int64_t)N * nz - z0) * (int64_t)sizeof(C));
  const uint32_t ls1 = jj < R2 ? lo : kOob;
  auto load_s1 = [&](int band, C* v) {
    const auto rs = buf_rsrc(P + (((int64_t)blockIdx.z * B.nbb + band) * B.nax_cap + a) * N * nz + z0, pl_bytes);
#pragma unroll
    for (int r = 0; r < R1; ++r) v[r] = buf_ld<C>(rs, (uint32_t)((jj + R2 * r) * nz) * (uint32_t)sizeof(C) + ls1);
  };
  C pre[kPrefetch<R> ? R1 : 1];
  if (kPrefetch<R> && bb0 < bb1) load_s1(bb0, pre);
  for (int band = bb0; band < bb1; ++band) {
    C v[R1];
    if constexpr (kPrefetch<R>) {
#pragma unroll
      for (int r = 0; r < R1; ++r) v[r] = pre[r];
      if (band + 1 < bb1) load_s1(band + 1, pre);
    } else {
      load_s1(band, v);
    }
    if (jj < R2) {
      dftR<R1, -1>(v, twl, N);
#pragma unroll
      for (int k = 0; k < R1; ++k) buf[(jj * R1 + k) * PL + l] = v[k];
    }
    __syncthreads();
    if (jj < R1) {
      C u[R2];
#pragma unroll
      for (int r = 0; r < R2; ++r) {
        const C y = buf[(jj + R1 * r) * PL + l];
        u[r] = r ? cmul(y, twl[(jj * r) % N]) : y;
      }
      dftR<R2, -1>(u, twl, N);
      const auto rd = buf_rsrc(S + ((int64_t)blockIdx.z * B.nbb + band) * B.nst_cap * nz + z0,
                               (uint32_t)(((int64_t)B.nst_cap * nz - z0) * (int64_t)sizeof(C)));
#pragma unroll
      for (int k = 0; k < R2; ++k) buf_st<C>(rd, srow[k], u[k]);
    }
    __syncthreads();
  }
}

// z passes on sticks for the two-stage lengths: SL sticks per workgroup,
// TPL = max(R1, R2) threads per stick; the sticks are padded lines of one
// LDS buffer (the plane-wave scatter / gather needs them there anyway), the
// two DFT stages run in registers with one transpose through that buffer.
// SL sticks per workgroup, 16 (FP64) or 32 (FP32): with the stick capacity
// a multiple of 16, every workgroup's segment of a z-major row is two whole
// 128-byte lines, so no line is written by two workgroups.  (32 FP64 sticks
// need 57 KB of LDS at N = 108 and cut the z passes to two workgroups per
// CU, which was slower.)
template <int N, int SL_> struct StickSplit {
  static constexpr int R1 = Split<N>::R1, R2 = Split<N>::R2;
  static constexpr int TPL = R1 > R2 ? R1 : R2;
  static constexpr int SL = SL_;       // sticks per workgroup
  static constexpr int NT = SL * TPL;
  // intermediate pitch between the stages: sub-transform j's outputs start
  // at slot j (R1 + 1), so the lanes' stage-1 stores stride by an odd count
  // (a stride of R1 = 12 complex128 is 48 dwords: lanes j and j + 4 would
  // share banks)
  static constexpr int IP = R1 + 1;
  // padded line pitch (odd), room for the padded intermediate
  static constexpr int LS = ((N + 1 > R2 * IP ? N + 1 : R2 * IP)) | 1;
  static constexpr size_t lds(size_t csize) { return csize * ((size_t)N + (size_t)SL * LS); }
};

// sticks per workgroup of a batch: its choice, or 16 (FP64) / 32 (FP32)
template <typename R> int stick_block(const PwBatch& B) { return B.zsl ? B.zsl : (sizeof(R) == 8 ? 16 : 32); }

// S1 + S2 of one transform on a padded LDS line, in place; thread (line, j)
// with j < TPL, `act` false for lines past the block's last stick
template <int SGN, typename C, int N>
__device__ __forceinline__ void stick_two_stage(C* line, const C* twl, int j, bool act) {
  using SS = StickSplit<N, 1>;
  constexpr int R1 = SS::R1, R2 = SS::R2, IP = SS::IP;
  C v[R1];
  if (act && j < R2) {
#pragma unroll
    for (int r = 0; r < R1; ++r) v[r] = line[j + R2 * r];
    dftR<R1, SGN>(v, twl, N);
  }
  __syncthreads();
  if (act && j < R2) {
#pragma unroll
    for (int k = 0; k < R1; ++k) line[j * IP + k] = v[k];
  }
  __syncthreads();
  C u[R2];
  if (act && j < R1) {
#pragma unroll
    for (int r = 0; r < R2; ++r) {
      const C y = line[r * IP + j];
      u[r] = r ? cmul(y, SGN > 0 ? cconj(twl[(j * r) % N]) : twl[(j * r) % N]) : y;
    }
    dftR<R2, SGN>(u, twl, N);
  }
  __syncthreads();
  if (act && j < R1) {
#pragma unroll
    for (int k = 0; k < R2; ++k) line[j + R1 * k] = u[k];
  }
  __syncthreads();
}

// (major, minor) task of a workgroup of a (gridDim.x majors x gridDim.y
// minors) grid, for workgroups that share data through their major index
// (the complex128 fused pass: the V rows of a z plane).  The dispatcher deals
// workgroups to the 8 XCDs in turn (linear id mod 8); the tasks are numbered
// major-first and each XCD takes a contiguous range of them, so the
// workgroups of one major index run on one XCD at about the same time and
// read the shared data through its L2 once.  (The z passes keep the stick
// block fastest: grouping the bands of a stick block also removed their
// re-read maps, but their stick-row writes, one item stride apart, then
// landed together and the passes ran 5-10 % slower.)
__device__ __forceinline__ void xcd_task(int& major, int& minor) {
  const int T = (int)(gridDim.x * gridDim.y);
  const int L = (int)(blockIdx.x + gridDim.x * blockIdx.y);
  const int x = L % 8, s = L / 8;
  const int t = x * (T / 8) + min(x, T % 8) + s;
  major = t / (int)gridDim.y;
  minor = t % (int)gridDim.y;
}

// Streaming (non-temporal) 16- or 8-byte access of one complex value, for the
// stick rows that the inverse z pass writes and the forward z pass reads:
// each is touched once.  (Measured at 108^3: the inverse z pass 87 -> 63 us
// per 80-band launch; the same hint on the y/x pass's stick rows slowed the
// complex64 forward z pass, whose input then no longer came from the cache.)
template <typename C> __device__ __forceinline__ void nt_store(C* p, C v) {
  using V = decltype(v.x) __attribute__((ext_vector_type(2)));
  V t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<V*>(p));
}
template <typename C> __device__ __forceinline__ C nt_load(const C* p) {
  using V = decltype(p->x) __attribute__((ext_vector_type(2)));
  const V t = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
  C r;
  r.x = t.x;
  r.y = t.y;
  return r;
}

// psi[b][perm[p]] -> S[item][s][:] for the block's sticks
template <typename R, int N, bool ZM, int SL_>
__global__ __launch_bounds__((StickSplit<N, SL_>::NT)) void k_sticks_z_inv2(PwBatch B, PwBlock<R> psi,
                                                                     const cx_t<R>* __restrict__ tw,
                                                                     cx_t<R>* __restrict__ S) {
  using C = cx_t<R>;
  using SS = StickSplit<N, SL_>;
  constexpr int NT = SS::NT, SL = SS::SL, LS = SS::LS, TPL = SS::TPL;
  extern __shared__ __align__(16) unsigned char smem[];
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
  const int s0 = blockIdx.x * SL;
  if (s0 >= m.nst) return;
  C* twl = reinterpret_cast<C*>(smem);
  C* buf = twl + N;
  const int lines = min(SL, m.nst - s0);
  const int band = blockIdx.y;
  for (int t = threadIdx.x; t < N; t += NT) twl[t] = tw[t];
  for (int e = threadIdx.x; e < lines * LS; e += NT) buf[e] = cmk<C>(0, 0);
  __syncthreads();
  const int p0 = m.st_pw[s0], p1 = m.st_pw[s0 + lines];
  const C* pb = psi.p + blockIdx.z * psi.ks + (int64_t)band * psi.ld;
  for (int p = p0 + threadIdx.x; p < p1; p += NT) buf[(m.pw_st[p] - s0) * LS + m.pw_z[p]] = pb[m.perm[p]];
  __syncthreads();
  const int line = threadIdx.x / TPL, j = threadIdx.x % TPL;
  stick_two_stage<+1, C, N>(buf + line * LS, twl, j, line < lines);
  if constexpr (ZM) {
    // z-major sticks: row z of the item holds every stick at that z
    C* dst = S + ((int64_t)blockIdx.z * B.nbb + band) * B.nst_cap * N + s0;
    for (int e = threadIdx.x; e < lines * N; e += NT) {
      const int z = e / lines, q = e - z * lines;
      nt_store(dst + (int64_t)z * B.nst_cap + q, buf[q * LS + z]);
    }
  } else {
    C* dst = S + (((int64_t)blockIdx.z * B.nbb + band) * B.nst_cap + s0) * N;
    for (int e = threadIdx.x; e < lines * N; e += NT) {
      const int q = e / N;
      nt_store(dst + e, buf[q * LS + (e - q * N)]);
    }
  }
}

// hpsi[b][perm[p]] = (z FFT of the stick)[pw_z[p]] + kin[perm[p]] psi[b][perm[p]]
template <typename R, int N, bool ZM, int SL_>
__global__ __launch_bounds__((StickSplit<N, SL_>::NT)) void k_sticks_z_fwd2(PwBatch B, const cx_t<R>* __restrict__ S,
                                                                     const cx_t<R>* __restrict__ tw, PwBlock<R> psi,
                                                                     PwBlock<R> hpsi) {
  using C = cx_t<R>;
  using SS = StickSplit<N, SL_>;
  constexpr int NT = SS::NT, SL = SS::SL, LS = SS::LS, TPL = SS::TPL;
  extern __shared__ __align__(16) unsigned char smem[];
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
  const int s0 = blockIdx.x * SL;
  if (s0 >= m.nst) return;
  C* twl = reinterpret_cast<C*>(smem);
  C* buf = twl + N;
  const int lines = min(SL, m.nst - s0);
  const int band = blockIdx.y;
  for (int t = threadIdx.x; t < N; t += NT) twl[t] = tw[t];
  if constexpr (ZM) {
    const C* src = S + ((int64_t)blockIdx.z * B.nbb + band) * B.nst_cap * N + s0;
    for (int e = threadIdx.x; e < lines * N; e += NT) {
      const int z = e / lines, q = e - z * lines;
      buf[q * LS + z] = nt_load(src + (int64_t)z * B.nst_cap + q);
    }
  } else {
    const C* src = S + (((int64_t)blockIdx.z * B.nbb + band) * B.nst_cap + s0) * N;
    for (int e = threadIdx.x; e < lines * N; e += NT) {
      const int q = e / N;
      buf[q * LS + (e - q * N)] = nt_load(src + e);
    }
  }
  __syncthreads();
  const int line = threadIdx.x / TPL, j = threadIdx.x % TPL;
  stick_two_stage<-1, C, N>(buf + line * LS, twl, j, line < lines);
  const int p0 = m.st_pw[s0], p1 = m.st_pw[s0 + lines];
  const C* pb = psi.p + blockIdx.z * psi.ks + (int64_t)band * psi.ld;
  C* hb = hpsi.p + blockIdx.z * hpsi.ks + (int64_t)band * hpsi.ld;
  for (int p = p0 + threadIdx.x; p < p1; p += NT) {
    const int g = m.perm[p];
    const C v = buf[(m.pw_st[p] - s0) * LS + m.pw_z[p]];
    const C x = pb[g];
    const R t = (R)m.kin[g];
    hb[g] = cmk<C>(v.x + t * x.x, v.y + t * x.y);
  }
}

// ---------------------------------------------------------------- launchers
template <typename K> static void allow_lds_pw(K kern, size_t bytes) {
  if (bytes > 65536) allow_dynamic_lds(reinterpret_cast<const void*>(kern), bytes);
}

static void check_lds_pw(size_t bytes, int n) {
  DEREK_REQUIRE(bytes <= 160 * 1024, "FFT length " + std::to_string(n) + " exceeds the LDS budget");
}

static int sticks_per_block(int nz) { return std::max(1, 1024 / nz); }

static void check_batch(const PwBatch& B) {
  DEREK_REQUIRE(B.mt && B.nkc >= 1 && B.nbb >= 1 && B.nkc <= 65535 && B.nbb <= 65535, "empty pipeline batch");
}

// the column passes address one item's sticks / planes by 32-bit byte offsets
template <typename R> static void check_offsets(const PwBatch& B, int64_t nyz, int nz) {
  const int64_t cs = (int64_t)sizeof(cx_t<R>);
  DEREK_REQUIRE((int64_t)B.nst_cap * nz * cs < (int64_t)kOob && (int64_t)B.nax_cap * nyz * cs < (int64_t)kOob,
                "one item's sticks or planes exceed 2 GiB");
}

template <typename R, int N, int SL>
static void sticks_z_inv2_sl(const PwBatch& B, PwBlock<R> psi, const Axis<R>* ax, cx_t<R>* S, bool zmajor,
                             hipStream_t s) {
  using SS = StickSplit<N, SL>;
  const size_t lds = SS::lds(sizeof(cx_t<R>));
  check_lds_pw(lds, N);
  auto kern = zmajor ? k_sticks_z_inv2<R, N, true, SL> : k_sticks_z_inv2<R, N, false, SL>;
  allow_lds_pw(kern, lds);
  dim3 grid((unsigned)ceil_div(B.nst_cap, SS::SL), B.nbb, B.nkc);
  hipLaunchKernelGGL(kern, grid, dim3(SS::NT), lds, s, B, psi, ax[2].tw, S);
}

template <typename R, int N, int SL>
static void sticks_z_fwd2_sl(const PwBatch& B, const cx_t<R>* S, const Axis<R>* ax, PwBlock<R> psi, PwBlock<R> hpsi,
                             bool zmajor, hipStream_t s) {
  using SS = StickSplit<N, SL>;
  const size_t lds = SS::lds(sizeof(cx_t<R>));
  check_lds_pw(lds, N);
  auto kern = zmajor ? k_sticks_z_fwd2<R, N, true, SL> : k_sticks_z_fwd2<R, N, false, SL>;
  allow_lds_pw(kern, lds);
  dim3 grid((unsigned)ceil_div(B.nst_cap, SS::SL), B.nbb, B.nkc);
  hipLaunchKernelGGL(kern, grid, dim3(SS::NT), lds, s, B, S, ax[2].tw, psi, hpsi);
}

template <typename R, int N>
static void sticks_z_inv2(const PwBatch& B, PwBlock<R> psi, const Axis<R>* ax, cx_t<R>* S, bool zmajor,
                          hipStream_t s) {
  if (stick_block<R>(B) == 32) sticks_z_inv2_sl<R, N, 32>(B, psi, ax, S, zmajor, s);
  else sticks_z_inv2_sl<R, N, 16>(B, psi, ax, S, zmajor, s);
}

template <typename R, int N>
static void sticks_z_fwd2(const PwBatch& B, const cx_t<R>* S, const Axis<R>* ax, PwBlock<R> psi, PwBlock<R> hpsi,
                          bool zmajor, hipStream_t s) {
  if (stick_block<R>(B) == 32) sticks_z_fwd2_sl<R, N, 32>(B, S, ax, psi, hpsi, zmajor, s);
  else sticks_z_fwd2_sl<R, N, 16>(B, S, ax, psi, hpsi, zmajor, s);
}

template <typename R>
void sticks_z_inv(const PwBatch& B, PwBlock<R> psi, const Axis<R>* ax, int nz, cx_t<R>* S, bool zmajor,
                  hipStream_t s) {
  check_batch(B);
  switch (nz) {
    case 32: sticks_z_inv2<R, 32>(B, psi, ax, S, zmajor, s); DEREK_HIP(hipGetLastError()); return;
    case 72: sticks_z_inv2<R, 72>(B, psi, ax, S, zmajor, s); DEREK_HIP(hipGetLastError()); return;
    case 108: sticks_z_inv2<R, 108>(B, psi, ax, S, zmajor, s); DEREK_HIP(hipGetLastError()); return;
    case 144: sticks_z_inv2<R, 144>(B, psi, ax, S, zmajor, s); DEREK_HIP(hipGetLastError()); return;
    default: break;
  }
  DEREK_REQUIRE(!zmajor, "z-major sticks need a z length with a compiled kernel");
  const int lpb = sticks_per_block(nz);
  const size_t lds = sizeof(cx_t<R>) * ((size_t)nz + 2 * (size_t)lpb * contig_stride(nz));
  check_lds_pw(lds, nz);
  dim3 grid((unsigned)ceil_div(B.nst_cap, lpb), B.nbb, B.nkc);
#define C_(NC)                                                                                  \
  {                                                                                             \
    auto kern = k_sticks_z_inv<R, NC>;                                                          \
    allow_lds_pw(kern, lds);                                                                    \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, psi, lpb, ax[2].plan, ax[2].tw, S);    \
  }
  C_(0)
#undef C_
  DEREK_HIP(hipGetLastError());
}

template <typename R, int N, int BPW>
static void planes_y2(const PwBatch& B, cx_t<R>* S, cx_t<R>* P, const Axis<R>* ax, int nz, int sign, hipStream_t s) {
  const int ztiles = (int)ceil_div(nz, PL);
  const size_t lds = Split<N>::lds(sizeof(cx_t<R>));
  dim3 grid((unsigned)(B.nax_cap * ztiles), (unsigned)ceil_div(B.nbb, BPW), B.nkc);
  if (sign > 0) {
    auto kern = k_py_inv2<R, N, BPW>;
    allow_lds_pw(kern, lds);
    hipLaunchKernelGGL(kern, grid, dim3(Split<N>::NT), lds, s, B, S, nz, ztiles, ax[1].tw, P);
  } else {
    auto kern = k_py_fwd2<R, N, BPW>;
    allow_lds_pw(kern, lds);
    hipLaunchKernelGGL(kern, grid, dim3(Split<N>::NT), lds, s, B, P, nz, ztiles, ax[1].tw, S);
  }
}

template <typename R, int N>
static void planes_y2_bpw(const PwBatch& B, cx_t<R>* S, cx_t<R>* P, const Axis<R>* ax, int nz, int sign, int bpw,
                          hipStream_t s) {
  switch (bpw) {
    case 1: planes_y2<R, N, 1>(B, S, P, ax, nz, sign, s); break;
    case 2: planes_y2<R, N, 2>(B, S, P, ax, nz, sign, s); break;
    case 4: planes_y2<R, N, 4>(B, S, P, ax, nz, sign, s); break;
    case 8: planes_y2<R, N, 8>(B, S, P, ax, nz, sign, s); break;
    default: throw ArgError("bands per workgroup must be 1, 2, 4 or 8");
  }
}

template <typename R>
void planes_y(const PwBatch& B, cx_t<R>* S, cx_t<R>* P, const Axis<R>* ax, int ny, int nz, int sign, int bpw,
              hipStream_t s) {
  check_batch(B);
  check_offsets<R>(B, (int64_t)ny * nz, nz);
  switch (ny) {
    case 32: planes_y2_bpw<R, 32>(B, S, P, ax, nz, sign, bpw, s); break;
    case 72: planes_y2_bpw<R, 72>(B, S, P, ax, nz, sign, bpw, s); break;
    case 108: planes_y2_bpw<R, 108>(B, S, P, ax, nz, sign, bpw, s); break;
    case 144: planes_y2_bpw<R, 144>(B, S, P, ax, nz, sign, bpw, s); break;
    default: {
      const int ztiles = (int)ceil_div(nz, PL);
      const size_t lds = sizeof(cx_t<R>) * ((size_t)ny + 2 * (size_t)PL * ny);
      check_lds_pw(lds, ny);
      dim3 grid((unsigned)(B.nax_cap * ztiles), B.nbb, B.nkc);
      if (sign > 0) {
        auto kern = k_planes_y_inv<R, 0>;
        allow_lds_pw(kern, lds);
        hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, S, nz, ztiles, ax[1].plan, ax[1].tw, P);
      } else {
        auto kern = k_planes_y_fwd<R, 0>;
        allow_lds_pw(kern, lds);
        hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, P, nz, ztiles, ax[1].plan, ax[1].tw, S);
      }
    }
  }
  DEREK_HIP(hipGetLastError());
}

template <typename R>
void sticks_z_fwd(const PwBatch& B, const cx_t<R>* S, const Axis<R>* ax, int nz, PwBlock<R> psi, PwBlock<R> hpsi,
                  bool zmajor, hipStream_t s) {
  check_batch(B);
  switch (nz) {
    case 32: sticks_z_fwd2<R, 32>(B, S, ax, psi, hpsi, zmajor, s); DEREK_HIP(hipGetLastError()); return;
    case 72: sticks_z_fwd2<R, 72>(B, S, ax, psi, hpsi, zmajor, s); DEREK_HIP(hipGetLastError()); return;
    case 108: sticks_z_fwd2<R, 108>(B, S, ax, psi, hpsi, zmajor, s); DEREK_HIP(hipGetLastError()); return;
    case 144: sticks_z_fwd2<R, 144>(B, S, ax, psi, hpsi, zmajor, s); DEREK_HIP(hipGetLastError()); return;
    default: break;
  }
  DEREK_REQUIRE(!zmajor, "z-major sticks need a z length with a compiled kernel");
  const int lpb = sticks_per_block(nz);
  const size_t lds = sizeof(cx_t<R>) * ((size_t)nz + 2 * (size_t)lpb * contig_stride(nz));
  check_lds_pw(lds, nz);
  dim3 grid((unsigned)ceil_div(B.nst_cap, lpb), B.nbb, B.nkc);
#define C_(NC)                                                                                       \
  {                                                                                                  \
    auto kern = k_sticks_z_fwd<R, NC>;                                                               \
    allow_lds_pw(kern, lds);                                                                         \
    hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, S, lpb, ax[2].plan, ax[2].tw, psi, hpsi);   \
  }
  C_(0)
#undef C_
  DEREK_HIP(hipGetLastError());
}

template <typename R, int BPW>
static void planes_x_vmul_bpw(const PwBatch& B, cx_t<R>* P, const Axis<R>* ax, int64_t nyz, const R* vloc,
                              R vscale, hipStream_t s) {
  const int nx = ax[0].n;
  dim3 grid((unsigned)ceil_div(nyz, PL), (unsigned)ceil_div(B.nbb, BPW), B.nkc);
#define C2_(NC)                                                                                           \
  {                                                                                                       \
    auto kern = k_px_vmul2<R, NC, BPW>;                                                                   \
    const size_t lds = Split<NC>::lds(sizeof(cx_t<R>));                                                   \
    allow_lds_pw(kern, lds);                                                                              \
    hipLaunchKernelGGL(kern, grid, dim3(Split<NC>::NT), lds, s, B, P, nyz, ax[0].tw, vloc, vscale);       \
  }
  switch (nx) {
    case 32: C2_(32); break;
    case 72: C2_(72); break;
    case 108: C2_(108); break;
    case 144: C2_(144); break;
    default: {
      const size_t lds = sizeof(cx_t<R>) * ((size_t)nx + 2 * (size_t)PL * nx) + sizeof(R) * (size_t)PL * nx;
      check_lds_pw(lds, nx);
      auto kern = k_planes_x_vmul<R, 0, BPW>;
      allow_lds_pw(kern, lds);
      hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, P, nyz, ax[0].plan, ax[0].tw, vloc, vscale);
    }
  }
#undef C2_
}

template <typename R>
void planes_x_vmul(const PwBatch& B, cx_t<R>* P, const Axis<R>* ax, int64_t nyz, const R* vloc, R vscale,
                   int bands_per_wg, hipStream_t s) {
  check_batch(B);
  check_offsets<R>(B, nyz, 1);
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
  check_offsets<R>(B, nyz, 1);
  const int nx = ax[0].n;
  dim3 grid((unsigned)ceil_div(nyz, PL));
#define C2_(NC)                                                                                                 \
  {                                                                                                             \
    auto kern = k_px_density2<R, NC>;                                                                           \
    const size_t lds = Split<NC>::lds(sizeof(cx_t<R>));                                                         \
    allow_lds_pw(kern, lds);                                                                                    \
    hipLaunchKernelGGL(kern, grid, dim3(Split<NC>::NT), lds, s, B, P, nyz, ax[0].tw, occ, occ_ks, rho, keep);   \
  }
  switch (nx) {
    case 32: C2_(32); break;
    case 72: C2_(72); break;
    case 108: C2_(108); break;
    case 144: C2_(144); break;
    default: {
      const size_t lds = sizeof(cx_t<R>) * ((size_t)nx + 2 * (size_t)PL * nx) + sizeof(double) * (size_t)PL * nx;
      check_lds_pw(lds, nx);
      auto kern = k_planes_x_density<R, 0>;
      allow_lds_pw(kern, lds);
      hipLaunchKernelGGL(kern, grid, dim3(PNT), lds, s, B, P, nyz, ax[0].plan, ax[0].tw, occ, occ_ks, rho, keep);
    }
  }
#undef C2_
  DEREK_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- fused y / x pass per z plane
// One workgroup per (z plane, group of bands of one k-point) for nx = ny = N.
// The sticks are z-major (row z of an item holds every stick at that z) and
// are transformed in place:
//   A  y inverse of each active x plane into A[a][.] (LDS);
//   B  x lines in rounds of LR: inverse transform of the active rows,
//      * V(x, y, z) / (nx ny nz), forward transform, active rows back into A;
//   C  y forward of each active plane, stick rows written back.
// Each length-N transform is two register stages, N = R1 R2: R2 threads
// (digit j2) take R1-point DFTs over the inputs j2 + R2 r, R1 threads (digit
// k1) finish with R2-point DFTs; the result with index k1 + R1 k2 sits in
// slot k1 R2P + k2 of its line (R2P = R2 rounded up to odd, so that lanes
// striding by R2P hit different LDS banks).  Every stage writes the slots it
// read, so one barrier per stage and no second buffer.  The workgroup size is
// a multiple of R1 and R2: each thread keeps its digits, its twiddles and its
// x rows in registers for the whole launch.  Compared with the separate y and
// x passes the active planes stay on chip: HBM sees the stick rows once each
// way and V once per band group.
// VM: parameters of k_slab_vmul (false: k_slab_density, whose FP64
// accumulators need more registers per thread)
template <typename R, int N, bool VM = false> struct Slab {
  static constexpr int R1 = Split<N>::R1, R2 = Split<N>::R2;
  static constexpr int R2P = R2 % 2 ? R2 : R2 + 1;
  static constexpr int L = R1 * R2P;              // slots of a line
  // line pitches of the plane area A and the x-line area X.  A 16-byte LDS
  // access of 16 lanes is conflict-free when their slot indices differ mod
  // 16; the lanes of one access span several planes or x lines, so the
  // pitches decide the conflicts.  For complex128 at 108 the pitches 113 / 108
  // came from a model of every access of the kernel (phase by phase, 16
  // lanes per pass over 64 four-byte banks): 25 % fewer LDS passes than
  // 109 / 109 and 22 % fewer than 108 / 109, within the LDS budget of 49
  // active planes.  (Measured: an unpadded A is 3-4 % faster for complex64
  // at 108^3.)
  static constexpr bool kPitch108 = sizeof(R) == 8 && N == 108;
  static constexpr int LPA = kPitch108 ? 113 : (sizeof(R) == 8 ? (L % 2 ? L : L + 1) : L);
  static constexpr int LPX = kPitch108 ? 108 : (L % 2 ? L : L + 1);
  static_assert(LPA >= L && LPX >= L, "line pitches below the slots of a line");
  // 360 threads at N = 72 (not 288): the R1- and R2-thread stages then take
  // 40 and 45 planes per pass, so the 33 active planes of a 16 bohr cell at
  // Ecut 20 Ha need one pass per y stage instead of two
  // complex64 H|psi> at 144 runs 840 threads (70 planes per pass: one pass
  // per y stage for the 69 active planes of a 30 bohr cell at Ecut 25 Ha,
  // and 14 waves on the CU that its LDS working set allows)
  // FP64 H|psi> at 72 runs 720 threads with all 72 x lines in one round
  // (7 barrier phases per band instead of 13; one 136 KB workgroup per CU)
  static constexpr bool kWide72 = VM && N == 72 && sizeof(R) == 8;
  static constexpr int NT =
      N == 32 ? 256 : (N == 72 ? (kWide72 ? 720 : 360) : (VM && N == 144 && sizeof(R) == 4 ? 840 : 504));
  // x lines per round: the complex64 108 case is sized for two workgroups per
  // CU; complex64 at 72 takes the 72 lines in two rounds
  static constexpr int LR =
      N == 32 ? 32 : (N == 72 ? (kWide72 ? 72 : (sizeof(R) == 4 ? 36 : 24)) : (N == 108 && sizeof(R) == 4 ? 27 : 36));
  static constexpr int NR = (N + LR - 1) / LR;
  static constexpr int LJ = NT / R2, LK = NT / R1;  // lines per pass of the R2- / R1-thread stages
  static_assert(NT % R1 == 0 && NT % R2 == 0, "workgroup size must keep each thread's digits");
  static_assert(R1 % 7 && R2 % 7 && R1 % 11 && R2 % 11 && R1 % 13 && R2 % 13,
                "the fused kernels use register DFTs with compile-time roots only");
  static_assert(LR <= LJ && LR <= LK, "one pass per stage of a round");
  static constexpr size_t lds(int nax_cap) {
    return sizeof(cx_t<R>) * ((size_t)nax_cap * LPA + (size_t)LR * LPX) + sizeof(int) * N +
           sizeof(int16_t) * (((size_t)nax_cap * N + 1) / 2 * 2);
  }
};

// Per-thread state of the fused kernels: LDS areas, the thread's digits
// (j2 in the R2-thread stages, k1 in the R1-thread stages, and its line in
// each), its twiddles exp(-2 pi i k1 r / N) and the LDS rows of the active
// planes at x = j2 + R2 r (-1 if none).
template <typename R, int N, bool VM = false> struct SlabThread {
  using C = cx_t<R>;
  using SL = Slab<R, N, VM>;
  static constexpr int R1 = SL::R1, R2 = SL::R2, R2P = SL::R2P, NT = SL::NT, LR = SL::LR;
  static constexpr int LPA = SL::LPA, LPX = SL::LPX;
  static constexpr int LJ = SL::LJ, LK = SL::LK;
  C* A;           // [nax_cap][LPA]
  C* X;           // [LR][LPX]
  int16_t* sty;   // stick of (plane, y bin), -1 if none
  int nax, j2, lj, k1, lk;
  C tk[R2];
  int xr[R1];

  // LDS tables of the k-point's maps (two barriers)
  __device__ __forceinline__ void init(const PwBatch& B, const PwMaps& m, const C* __restrict__ tw,
                                       unsigned char* smem) {
    const int tid = threadIdx.x;
    A = reinterpret_cast<C*>(smem);
    X = A + (size_t)B.nax_cap * LPA;
    int* xa = reinterpret_cast<int*>(X + LR * LPX);  // active plane of x bin, -1 if none
    sty = reinterpret_cast<int16_t*>(xa + N);
    nax = m.nax;
    j2 = tid % R2, lj = tid / R2, k1 = tid % R1, lk = tid / R1;
    for (int x = tid; x < N; x += NT) xa[x] = -1;
    for (int i = tid; i < nax * N; i += NT) sty[i] = (int16_t)m.pl_y[i];
#pragma unroll
    for (int r = 1; r < R2; ++r) tk[r] = tw[(k1 * r) % N];
    __syncthreads();
    for (int a = tid; a < nax; a += NT) xa[m.ax_x[a]] = a;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R1; ++r) {
      const int a = xa[j2 + R2 * r];
      xr[r] = a >= 0 ? a * LPA : -1;
    }
  }
  __device__ static int pos(int y) { return (y % R1) * R2P + y / R1; }  // slot of index y
  __device__ static uint32_t stick_off(int st) { return st >= 0 ? (uint32_t)st * (uint32_t)sizeof(C) : kOob; }
  // the sticks y = j2 + R2 r of plane a from a stick row
  template <typename Rs> __device__ __forceinline__ void load_a1(const Rs& rs, int a, C* v) const {
    const int16_t* ty = sty + a * N + j2;
#pragma unroll
    for (int r = 0; r < R1; ++r) v[r] = buf_ld<C>(rs, stick_off(ty[R2 * r]));
  }
  // y inverse of every active plane from a stick row into A (two barriers);
  // with PRE, `pre` holds the first item's inputs
  template <bool PRE, typename Rs> __device__ __forceinline__ void y_inverse(const Rs& rs, const C (&pre)[R1]) {
    // A1: y = j2 + R2 r from the sticks, output k in slot k R2P + j2
    for (int a = lj; a < nax; a += LJ) {
      C v[R1];
      if (PRE && a == lj) {
#pragma unroll
        for (int r = 0; r < R1; ++r) v[r] = pre[r];
      } else {
        load_a1(rs, a, v);
      }
      dftR<R1, +1>(v, (const C*)nullptr, N);
      C* row = A + a * LPA + j2;
#pragma unroll
      for (int k = 0; k < R1; ++k) row[k * R2P] = v[k];
    }
    __syncthreads();
    // A2: y = k1 + R1 k2 into slot k1 R2P + k2
    for (int a = lk; a < nax; a += LK) {
      C* row = A + a * LPA + k1 * R2P;
      C u[R2];
#pragma unroll
      for (int r = 0; r < R2; ++r) u[r] = r ? cmul(row[r], cconj(tk[r])) : row[r];
      dftR<R2, +1>(u, (const C*)nullptr, N);
#pragma unroll
      for (int k = 0; k < R2; ++k) row[k] = u[k];
    }
    __syncthreads();
  }
  // B1: x line y0 + lj (lj < ln), active rows x = j2 + R2 r, first stage of
  // the inverse into X (barrier after)
  __device__ __forceinline__ void x_inverse1(int y0, int ln) {
    if (lj < ln) {
      const int py = pos(y0 + lj);
      C v[R1];
#pragma unroll
      for (int r = 0; r < R1; ++r) v[r] = xr[r] >= 0 ? A[xr[r] + py] : cmk<C>(0, 0);
      dftR<R1, +1>(v, (const C*)nullptr, N);
      C* row = X + lj * LPX + j2;
#pragma unroll
      for (int k = 0; k < R1; ++k) row[k * R2P] = v[k];
    }
    __syncthreads();
  }
  // second stage of the inverse of line lk: x = k1 + R1 k2 in u
  __device__ __forceinline__ void x_inverse2(C* u) const {
    const C* row = X + lk * LPX + k1 * R2P;
#pragma unroll
    for (int r = 0; r < R2; ++r) u[r] = r ? cmul(row[r], cconj(tk[r])) : row[r];
    dftR<R2, +1>(u, (const C*)nullptr, N);
  }
};

template <typename R, int N>
__global__ __launch_bounds__((Slab<R, N, true>::NT)) void k_slab_vmul(PwBatch B, cx_t<R>* __restrict__ S, int nz, int bpw,
                                                              const cx_t<R>* __restrict__ tw,
                                                              const R* __restrict__ vt) {
  using C = cx_t<R>;
  using ST = SlabThread<R, N, true>;
  constexpr int R1 = ST::R1, R2 = ST::R2, LR = ST::LR, NR = Slab<R, N, true>::NR, LJ = ST::LJ, LK = ST::LK;
  constexpr int LPA = ST::LPA, LPX = ST::LPX;
  extern __shared__ __align__(16) unsigned char smem[];
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
  // complex128 runs one workgroup per CU and waits on its V rows: planes
  // grouped per XCD (measured 4-5 % faster at 108^3; complex64, two
  // workgroups per CU, keeps the plane-fastest order, which measured 1-2 %
  // faster for it)
  int z = blockIdx.x, grp = blockIdx.y;
  if constexpr (sizeof(R) == 8) xcd_task(z, grp);
  ST t;
  t.init(B, m, tw, smem);
  const int nax = t.nax, j2 = t.j2, lj = t.lj, k1 = t.k1, lk = t.lk;
  C* A = t.A;
  C* X = t.X;
  const R* vz = vt + (int64_t)z * N * N;  // V at (x, y, z) in vz[y*N + x]
  const int bb0 = grp * bpw, bb1 = min(B.nbb, bb0 + bpw);
  const uint32_t row_bytes = (uint32_t)B.nst_cap * (uint32_t)sizeof(C);
  auto row_rsrc = [&](int band) {
    return buf_rsrc(S + (((int64_t)blockIdx.z * B.nbb + band) * nz + z) * B.nst_cap, row_bytes);
  };
  // the first A1 item's sticks of the next band are loaded during phase C
  // (complex128 at one workgroup per CU, where the registers are free)
  constexpr bool kPre = sizeof(R) == 8 && N >= 100;
  C pre[R1];
  if (kPre && lj < nax) t.load_a1(row_rsrc(bb0), lj, pre);
  for (int band = bb0; band < bb1; ++band) {
    const auto rs = row_rsrc(band);
    t.template y_inverse<kPre>(rs, pre);
#pragma unroll 1
    for (int q = 0; q < NR; ++q) {
      const int y0 = q * LR, ln = min(LR, N - y0);
      const bool b2 = lk < ln;
      // V of this thread's B2 line, loaded before B1 and used after its barrier
      R vq[R2];
      {
        const R* vy = vz + (y0 + (b2 ? lk : 0)) * N + k1;
#pragma unroll
        for (int k = 0; k < R2; ++k) vq[k] = vy[R1 * k];
      }
      t.x_inverse1(y0, ln);
      // B2: finish the inverse (x = k1 + R1 k2), * V, start the forward transform
      if (b2) {
        C u[R2];
        t.x_inverse2(u);
#pragma unroll
        for (int k = 0; k < R2; ++k) u[k] = cscale(u[k], vq[k]);
        dftR<R2, -1>(u, (const C*)nullptr, N);
        C* row = X + lk * LPX + k1 * ST::R2P;
#pragma unroll
        for (int k = 0; k < R2; ++k) row[k] = k ? cmul(u[k], t.tk[k]) : u[k];
      }
      __syncthreads();
      // B3: finish the forward transform, x = j2 + R2 m1, active rows back into A
      if (lj < ln) {
        const int py = ST::pos(y0 + lj);
        const C* row = X + lj * LPX + j2;
        C v[R1];
#pragma unroll
        for (int j = 0; j < R1; ++j) v[j] = row[j * ST::R2P];
        dftR<R1, -1>(v, (const C*)nullptr, N);
#pragma unroll
        for (int k = 0; k < R1; ++k)
          if (t.xr[k] >= 0) A[t.xr[k] + py] = v[k];
      }
      __syncthreads();
    }
    if (kPre && band + 1 < bb1 && lj < nax) t.load_a1(row_rsrc(band + 1), lj, pre);
    // C1: y = j2 + R2 r, output k1 back into the slot of y = j2 + R2 k1
    for (int a = lj; a < nax; a += LJ) {
      C* row = A + a * LPA;
      C v[R1];
#pragma unroll
      for (int r = 0; r < R1; ++r) v[r] = row[ST::pos(j2 + R2 * r)];
      dftR<R1, -1>(v, (const C*)nullptr, N);
#pragma unroll
      for (int k = 0; k < R1; ++k) row[ST::pos(j2 + R2 * k)] = v[k];
    }
    __syncthreads();
    // C2: y = k1 + R1 k2 to the sticks
    for (int a = lk; a < nax; a += LK) {
      const C* row = A + a * LPA;
      const int16_t* ty = t.sty + a * N + k1;
      C u[R2];
#pragma unroll
      for (int j = 0; j < R2; ++j) {
        const C w = row[ST::pos(j + R2 * k1)];
        u[j] = j ? cmul(w, t.tk[j]) : w;
      }
      dftR<R2, -1>(u, (const C*)nullptr, N);
#pragma unroll
      for (int k = 0; k < R2; ++k) buf_st<C>(rs, ST::stick_off(ty[R1 * k]), u[k]);
    }
    __syncthreads();
  }
}

// Density on z-major sticks: per z plane and group of bands of one k-point,
// the y and x inverse transforms as in k_slab_vmul, then
// part[p][z][y][x] = sum over the group's bands of occ |psi(r)|^2 in FP64
// (p = k-point slot * groups + group); k_slab_rho_sum adds the partials.
template <typename R, int N>
__global__ __launch_bounds__((Slab<R, N>::NT)) void k_slab_density(PwBatch B, const cx_t<R>* __restrict__ S, int nz,
                                                                 int bpg, const cx_t<R>* __restrict__ tw,
                                                                 const double* __restrict__ occ, int occ_ks,
                                                                 double* __restrict__ part) {
  using C = cx_t<R>;
  using ST = SlabThread<R, N>;
  constexpr int R1 = ST::R1, R2 = ST::R2, LR = ST::LR, NR = Slab<R, N>::NR;
  extern __shared__ __align__(16) unsigned char smem[];
  const PwMaps m = B.mt[B.k0 + blockIdx.z];
  const int z = blockIdx.x;
  ST t;
  t.init(B, m, tw, smem);
  double acc[NR][R2];
#pragma unroll
  for (int q = 0; q < NR; ++q)
#pragma unroll
    for (int k = 0; k < R2; ++k) acc[q][k] = 0.0;
  const int bb0 = blockIdx.y * bpg, bb1 = min(B.nbb, bb0 + bpg);
  const uint32_t row_bytes = (uint32_t)B.nst_cap * (uint32_t)sizeof(C);
  for (int band = bb0; band < bb1; ++band) {
    const double w = occ[(int64_t)blockIdx.z * occ_ks + band];
    if (w == 0.0) continue;  // uniform over the workgroup
    const auto rs = buf_rsrc(S + (((int64_t)blockIdx.z * B.nbb + band) * nz + z) * B.nst_cap, row_bytes);
    const C none[R1] = {};
    t.template y_inverse<false>(rs, none);
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const int y0 = q * LR, ln = min(LR, N - y0);
      t.x_inverse1(y0, ln);
      if (t.lk < ln) {
        C u[R2];
        t.x_inverse2(u);
#pragma unroll
        for (int k = 0; k < R2; ++k) {
          const double re = (double)u[k].x, im = (double)u[k].y;
          acc[q][k] += w * (re * re + im * im);
        }
      }
      __syncthreads();
    }
  }
  double* pz = part + ((int64_t)(blockIdx.z * gridDim.y + blockIdx.y) * nz + z) * N * N;
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const int y = q * LR + t.lk;
    if (t.lk < min(LR, N - q * LR)) {
#pragma unroll
      for (int k = 0; k < R2; ++k) pz[(int64_t)y * N + t.k1 + R1 * k] = acc[q][k];
    }
  }
}

// rho[(x*ny + y)*nz + z] = keep*rho + sum_p part[p][z][y][x] (p in order),
// 32 x 32 (x, z) tiles of one y through LDS
__global__ __launch_bounds__(256) void k_slab_rho_sum(const double* __restrict__ part, int np, int nx, int ny, int nz,
                                                      double* __restrict__ rho, int keep) {
  __shared__ double tile[32][33];
  const int x0 = blockIdx.x * 32, z0 = blockIdx.y * 32, y = blockIdx.z;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  const int64_t ng = (int64_t)nx * ny * nz;
  for (int i = ty; i < 32; i += 8) {
    const int z = z0 + i, x = x0 + tx;
    double s = 0.0;
    if (z < nz && x < nx) {
      const double* p = part + ((int64_t)z * ny + y) * nx + x;
      // in order, with the loads of eight partials in flight
#pragma unroll 8
      for (int j = 0; j < np; ++j) s += p[j * ng];
    }
    tile[i][tx] = s;
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int x = x0 + i, z = z0 + tx;
    if (x < nx && z < nz) {
      double* r = rho + ((int64_t)x * ny + y) * nz + z;
      *r = (keep ? *r : 0.0) + tile[tx][i];
    }
  }
}

// both fused kernels (H|psi> and density) fit the LDS for nax_cap planes
template <typename R, int N> static bool slab_fits(int nax_cap) {
  return Slab<R, N>::lds(nax_cap) <= 160 * 1024 && Slab<R, N, true>::lds(nax_cap) <= 160 * 1024;
}

template <typename R> bool slab_supported(int nx, int ny, int nz, int nax_cap, int nst_cap) {
  if (nx != ny || nax_cap < 1 || nax_cap > nx || nst_cap > 32767) return false;
  if (nz != 32 && nz != 72 && nz != 108 && nz != 144) return false;  // z-major stick kernels
  switch (nx) {
    case 32: return slab_fits<R, 32>(nax_cap);
    case 72: return slab_fits<R, 72>(nax_cap);
    case 108: return slab_fits<R, 108>(nax_cap);
    case 144: return slab_fits<R, 144>(nax_cap);
    default: return false;
  }
}

template <typename R, int N>
static void slab_vmul_n(const PwBatch& B, cx_t<R>* S, const Axis<R>* ax, int nz, const R* vt, int bpw, hipStream_t s) {
  using SL = Slab<R, N, true>;
  const size_t lds = SL::lds(B.nax_cap);
  check_lds_pw(lds, N);
  auto kern = k_slab_vmul<R, N>;
  allow_lds_pw(kern, lds);
  dim3 grid((unsigned)nz, (unsigned)ceil_div(B.nbb, bpw), B.nkc);
  hipLaunchKernelGGL(kern, grid, dim3(SL::NT), lds, s, B, S, nz, bpw, ax[0].tw, vt);
}

template <typename R>
void slab_vmul(const PwBatch& B, cx_t<R>* S, const Axis<R>* ax, int nz, const R* vt, int bpw, hipStream_t s) {
  check_batch(B);
  DEREK_REQUIRE(slab_supported<R>(ax[0].n, ax[1].n, nz, B.nax_cap, B.nst_cap),
                "grid not supported by the fused y/x pass");
  DEREK_REQUIRE(bpw >= 1, "bands per workgroup must be positive");
  DEREK_REQUIRE((int64_t)B.nst_cap * (int64_t)sizeof(cx_t<R>) < (int64_t)kOob, "stick row exceeds 2 GiB");
  switch (ax[0].n) {
    case 32: slab_vmul_n<R, 32>(B, S, ax, nz, vt, bpw, s); break;
    case 72: slab_vmul_n<R, 72>(B, S, ax, nz, vt, bpw, s); break;
    case 108: slab_vmul_n<R, 108>(B, S, ax, nz, vt, bpw, s); break;
    case 144: slab_vmul_n<R, 144>(B, S, ax, nz, vt, bpw, s); break;
    default: throw ArgError("grid not supported by the fused y/x pass");
  }
  DEREK_HIP(hipGetLastError());
}

template <typename R, int N> static int slab_density_groups_n(const PwBatch& B, int nz) {
  static const int cus = [] {
    int dev = 0;
    DEREK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    DEREK_HIP(hipGetDeviceProperties(&p, dev));
    return std::max(1, p.multiProcessorCount);
  }();
  const size_t lds = Slab<R, N>::lds(B.nax_cap);
  allow_lds_pw(k_slab_density<R, N>, lds);
  int per = 0;
  DEREK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k_slab_density<R, N>),
                                                         Slab<R, N>::NT, lds));
  // about four rounds of resident workgroups
  const int64_t want = 4LL * std::max(1, per) * cus;
  const int64_t g = (want + (int64_t)nz * B.nkc - 1) / ((int64_t)nz * B.nkc);
  const int bpg = (int)ceil_div(B.nbb, std::max<int64_t>(1, std::min<int64_t>(g, B.nbb)));
  return (int)ceil_div(B.nbb, bpg);
}

template <typename R> int slab_density_groups(const PwBatch& B, const Axis<R>* ax, int nz) {
  check_batch(B);
  DEREK_REQUIRE(slab_supported<R>(ax[0].n, ax[1].n, nz, B.nax_cap, B.nst_cap),
                "grid not supported by the fused y/x pass");
  switch (ax[0].n) {
    case 32: return slab_density_groups_n<R, 32>(B, nz);
    case 72: return slab_density_groups_n<R, 72>(B, nz);
    case 108: return slab_density_groups_n<R, 108>(B, nz);
    case 144:
      // complex128 planes of 144^3 do not fit the LDS (slab_supported is false)
      if constexpr (sizeof(R) == 4) return slab_density_groups_n<R, 144>(B, nz);
      throw ArgError("grid not supported by the fused y/x pass");
    default: throw ArgError("grid not supported by the fused y/x pass");
  }
}

template <typename R, int N>
static void slab_density_n(const PwBatch& B, const cx_t<R>* S, const Axis<R>* ax, int nz, const double* occ, int occ_ks,
                           double* part, int groups, hipStream_t s) {
  using SL = Slab<R, N>;
  const size_t lds = SL::lds(B.nax_cap);
  check_lds_pw(lds, N);
  auto kern = k_slab_density<R, N>;
  allow_lds_pw(kern, lds);
  dim3 grid((unsigned)nz, (unsigned)groups, B.nkc);
  hipLaunchKernelGGL(kern, grid, dim3(SL::NT), lds, s, B, S, nz, (int)ceil_div(B.nbb, groups), ax[0].tw, occ, occ_ks,
                     part);
}

template <typename R>
void slab_density(const PwBatch& B, const cx_t<R>* S, const Axis<R>* ax, int nz, const double* occ, int occ_ks,
                  double* part, int groups, double* rho, int keep, hipStream_t s) {
  check_batch(B);
  DEREK_REQUIRE(slab_supported<R>(ax[0].n, ax[1].n, nz, B.nax_cap, B.nst_cap),
                "grid not supported by the fused y/x pass");
  DEREK_REQUIRE(groups >= 1 && groups <= B.nbb && ceil_div(B.nbb, ceil_div(B.nbb, groups)) == groups,
                "band groups must come from slab_density_groups");
  switch (ax[0].n) {
    case 32: slab_density_n<R, 32>(B, S, ax, nz, occ, occ_ks, part, groups, s); break;
    case 72: slab_density_n<R, 72>(B, S, ax, nz, occ, occ_ks, part, groups, s); break;
    case 108: slab_density_n<R, 108>(B, S, ax, nz, occ, occ_ks, part, groups, s); break;
    case 144:
      if constexpr (sizeof(R) == 4) {
        slab_density_n<R, 144>(B, S, ax, nz, occ, occ_ks, part, groups, s);
        break;
      }
      throw ArgError("grid not supported by the fused y/x pass");
    default: throw ArgError("grid not supported by the fused y/x pass");
  }
  const int nx = ax[0].n, ny = ax[1].n;
  dim3 grid((unsigned)ceil_div(nx, 32), (unsigned)ceil_div(nz, 32), (unsigned)ny);
  hipLaunchKernelGGL(k_slab_rho_sum, grid, dim3(256), 0, s, part, B.nkc * groups, nx, ny, nz, rho, keep);
  DEREK_HIP(hipGetLastError());
}

#define INST(R)                                                                                                   \
  template void sticks_z_inv<R>(const PwBatch&, PwBlock<R>, const Axis<R>*, int, cx_t<R>*, bool, hipStream_t);   \
  template void planes_y<R>(const PwBatch&, cx_t<R>*, cx_t<R>*, const Axis<R>*, int, int, int, int, hipStream_t); \
  template void sticks_z_fwd<R>(const PwBatch&, const cx_t<R>*, const Axis<R>*, int, PwBlock<R>, PwBlock<R>,      \
                                bool, hipStream_t);                                                               \
  template bool slab_supported<R>(int, int, int, int, int);                                                       \
  template void slab_vmul<R>(const PwBatch&, cx_t<R>*, const Axis<R>*, int, const R*, int, hipStream_t);          \
  template int slab_density_groups<R>(const PwBatch&, const Axis<R>*, int);                                       \
  template void slab_density<R>(const PwBatch&, const cx_t<R>*, const Axis<R>*, int, const double*, int, double*,  \
                                int, double*, int, hipStream_t);                                                  \
  template void planes_x_vmul<R>(const PwBatch&, cx_t<R>*, const Axis<R>*, int64_t, const R*, R, int, hipStream_t); \
  template void planes_x_density<R>(const PwBatch&, const cx_t<R>*, const Axis<R>*, int64_t, const double*, int,  \
                                    double*, int, hipStream_t);
INST(double)
INST(float)
#undef INST

}  // namespace derek
