// Tall-skinny complex GEMMs on the CDNA4 matrix cores, batched over k-points.
//
// Overlap  C = X^H Y  (X: npw x m, Y: npw x n, npw >> m, n): split-K over
// the plane-wave index; each workgroup owns a 64x64 tile of C and a chunk of
// rows of one k-point (blockIdx.z), stages 16-row slabs of X and Y through
// LDS and accumulates three real products per complex product (Gauss's
// method: 25% fewer matrix-core operations than the four-product form) with
// 16x16x4 MFMAs (f64, or f32 for complex64 data); the next slab is loaded
// into registers while the current one is multiplied.  A second kernel sums
// the chunk partials in a fixed order in FP64 (deterministic).
// Rotation Z = X U: each workgroup owns 64 rows x 64 columns of Z; U slabs
// (FP64, rounded to R) are staged through LDS, X is read straight into the
// MFMA A operand; three real products per complex product as well.
//
// 16x16x4 MFMA operand layout (lane l of a 64-wide wave):
//   A (16x4): A[l%16][l/16]      B (4x16): B[l/16][l%16]
//   D (16x16), f64: register r holds D[l/16 + 4*r][l%16]
//   D (16x16), f32: register r holds D[4*(l/16) + r][l%16]
#include "launch.hpp"

namespace derek {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename R> struct Mfma;
template <> struct Mfma<double> {
  using acc = d4;
  __device__ static __forceinline__ d4 op(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
  using acc = f4;
  __device__ static __forceinline__ f4 op(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int OV_T = 64;     // C tile edge
constexpr int OV_KS = 16;    // rows per LDS slab
constexpr int OV_PAD = 66;   // LDS row pitch

struct OvGeom {
  int m, n, tm, tn, ntiles, nchunk;
  int64_t npw, kch;
  bool herm;
};

// Row chunks per tile: at least two full rounds of workgroups over the
// chip's resident slots (`slots` = resident workgroups per CU x CUs), then,
// among the next few counts, the one that leaves the smallest idle tail in
// the last round; every chunk keeps at least 256 rows.
static OvGeom ov_geom(int m, int n, int64_t npw, bool herm, int batch, int slots) {
  OvGeom g{};
  g.m = m;
  g.n = n;
  g.npw = npw;
  g.herm = herm;
  g.tm = (int)ceil_div(m, OV_T);
  g.tn = (int)ceil_div(n, OV_T);
  g.ntiles = herm ? g.tm * (g.tm + 1) / 2 : g.tm * g.tn;
  const int64_t units = (int64_t)g.ntiles * std::max(1, batch);
  const int64_t cmax = std::max<int64_t>(1, npw / 256);
  const int64_t c0 = std::min(cmax, std::max<int64_t>(1, ceil_div(2 * (int64_t)slots, units)));
  int64_t best = c0;
  double best_eff = -1.0;
  for (int64_t c = c0; c <= std::min(cmax, c0 + 16); ++c) {
    const int64_t w = units * c;
    const double eff = (double)w / (double)(ceil_div(w, (int64_t)slots) * slots);
    if (eff > best_eff + 1e-9) {
      best_eff = eff;
      best = c;
    }
  }
  int64_t kch = ceil_div(std::max<int64_t>(npw, 1), best);
  kch = std::max<int64_t>(OV_KS, ceil_div(kch, OV_KS) * OV_KS);
  g.kch = kch;
  g.nchunk = (int)std::max<int64_t>(1, ceil_div(npw, kch));
  return g;
}

// at least two waves per SIMD (register budget 256 per lane incl. accumulators)
template <typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_overlap(const cx_t<R>* __restrict__ X, int64_t ldx, int64_t sx,
                                                 const cx_t<R>* __restrict__ Y, int64_t ldy, int64_t sy,
                                                 const int* __restrict__ npwk, OvGeom g,
                                                 double2* __restrict__ part) {
  using M = Mfma<R>;
  using A4 = typename M::acc;
  using C = cx_t<R>;
  // slab operands for the three real products of the complex product
  // conj(x) y (Gauss):  P1 = xr yr,  P2 = xi yi,  P3 = (xr - xi)(yr + yi);
  // Re = P1 + P2,  Im = P3 - P1 + P2
  __shared__ R xr[OV_KS][OV_PAD], xi[OV_KS][OV_PAD], xs[OV_KS][OV_PAD];
  __shared__ R yr[OV_KS][OV_PAD], yi[OV_KS][OV_PAD], ys[OV_KS][OV_PAD];
  const int kz = blockIdx.z;
  X += kz * sx;
  Y += kz * sy;
  const int64_t npw = npwk ? (int64_t)npwk[kz] : g.npw;
  // tile coordinates
  int ti, tj;
  int t = blockIdx.x;
  if (g.herm) {
    ti = 0;
    while (t >= g.tm - ti) { t -= g.tm - ti; ++ti; }
    tj = ti + t;
  } else {
    ti = t % g.tm;
    tj = t / g.tm;
  }
  const int i0 = ti * OV_T, j0 = tj * OV_T;
  const int64_t r0 = (int64_t)blockIdx.y * g.kch;
  const int64_t r1 = std::min<int64_t>(npw, r0 + g.kch);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = wave & 1, wj = wave >> 1;
  A4 p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      p1[a][b] = A4{0, 0, 0, 0};
      p2[a][b] = A4{0, 0, 0, 0};
      p3[a][b] = A4{0, 0, 0, 0};
    }
  const int lrow = threadIdx.x & 15, lcol = threadIdx.x >> 4;  // 16 rows x 16 column groups
  // the next slab is loaded into registers while the current one is multiplied
  C px[4], py[4];
  auto gload = [&](int64_t k0) {
    const int64_t gr = k0 + lrow;
    const bool rok = gr < r1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = lcol + 16 * q;
      px[q] = (rok && i0 + c < g.m) ? X[(int64_t)(i0 + c) * ldx + gr] : cmk<C>(0, 0);
      py[q] = (rok && j0 + c < g.n) ? Y[(int64_t)(j0 + c) * ldy + gr] : cmk<C>(0, 0);
    }
  };
  if (r0 < r1) gload(r0);
  for (int64_t k0 = r0; k0 < r1; k0 += OV_KS) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = lcol + 16 * q;
      xr[lrow][c] = px[q].x;
      xi[lrow][c] = px[q].y;
      xs[lrow][c] = px[q].x - px[q].y;
      yr[lrow][c] = py[q].x;
      yi[lrow][c] = py[q].y;
      ys[lrow][c] = py[q].x + py[q].y;
    }
    __syncthreads();
    if (k0 + OV_KS < r1) gload(k0 + OV_KS);
#pragma unroll
    for (int kk = 0; kk < OV_KS / 4; ++kk) {
      const int k = kk * 4 + (lane >> 4);
      R ar[2], ai[2], as[2], br[2], bi[2], bs[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int ii = wi * 32 + a * 16 + (lane & 15);
        ar[a] = xr[k][ii];
        ai[a] = xi[k][ii];
        as[a] = xs[k][ii];
        const int jj = wj * 32 + a * 16 + (lane & 15);
        br[a] = yr[k][jj];
        bi[a] = yi[k][jj];
        bs[a] = ys[k][jj];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          p1[a][b] = M::op(ar[a], br[b], p1[a][b]);
          p2[a][b] = M::op(ai[a], bi[b], p2[a][b]);
          p3[a][b] = M::op(as[a], bs[b], p3[a][b]);
        }
    }
    __syncthreads();
  }
  double2* P = part + ((int64_t)kz * g.nchunk + blockIdx.y) * g.m * g.n;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wi * 32 + a * 16 + M::row(lane, r);
        const int j = j0 + wj * 32 + b * 16 + (lane & 15);
        const double q1 = (double)p1[a][b][r], q2 = (double)p2[a][b][r], q3 = (double)p3[a][b][r];
        if (i < g.m && j < g.n) P[(int64_t)j * g.m + i] = make_double2(q1 + q2, q3 - q1 + q2);
      }
}

// resident overlap workgroups on the current device (CUs x per-CU occupancy)
template <typename R> static int ov_slots() {
  static const int slots = [] {
    int dev = 0;
    DEREK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    DEREK_HIP(hipGetDeviceProperties(&p, dev));
    int per = 0;
    DEREK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(k_overlap<R>), 256, 0));
    return std::max(1, per) * std::max(1, p.multiProcessorCount);
  }();
  return slots;
}

int64_t overlap_work_elems(int m, int n, int64_t npw, int batch) {
  int64_t chunks = 1;
  for (int slots : {ov_slots<double>(), ov_slots<float>()})
    for (bool h : {false, m == n}) chunks = std::max<int64_t>(chunks, ov_geom(m, n, npw, h, batch, slots).nchunk);
  return (int64_t)batch * chunks * m * n;
}

// C[kz] = sum over chunks, columns spread over the workgroups of blockIdx.x
__global__ void k_overlap_reduce(const double2* __restrict__ part, OvGeom g, double2* __restrict__ C, int ldc,
                                 int64_t sc) {
  const int kz = blockIdx.y;
  const int64_t mn = (int64_t)g.m * g.n;
  const double2* pk = part + (int64_t)kz * g.nchunk * mn;
  double2* Ck = C + kz * sc;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mn; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.m), j = (int)(e / g.m);
    // Hermitian: the strict lower triangle is the conjugate of the upper one
    // (exactly Hermitian output, real diagonal)
    const bool mirror = g.herm && i > j;
    const int64_t src = mirror ? (int64_t)i * g.m + j : (int64_t)j * g.m + i;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < g.nchunk; ++c) {
      const double2 v = pk[(int64_t)c * mn + src];
      sr += v.x;
      si += v.y;
    }
    if (g.herm && i == j) si = 0.0;
    Ck[(int64_t)j * ldc + i] = mirror ? make_double2(sr, -si) : make_double2(sr, si);
  }
}

template <typename R>
void overlap(const cx_t<R>* X, int64_t ldx, int64_t sx, const cx_t<R>* Y, int64_t ldy, int64_t sy, int m, int n,
             int64_t npw, const int* npwk, double2* C, int ldc, int64_t sc, int batch, double2* work,
             int64_t work_cap, bool hermitian, hipStream_t s) {
  DEREK_REQUIRE(m > 0 && n > 0 && batch > 0, "overlap: empty block");
  DEREK_REQUIRE(!hermitian || m == n, "overlap: hermitian needs a square result");
  OvGeom g = ov_geom(m, n, npw, hermitian, batch, ov_slots<R>());
  DEREK_REQUIRE((int64_t)batch * g.nchunk * m * n <= work_cap, "overlap: partial-sum buffer too small");
  hipLaunchKernelGGL(k_overlap<R>, dim3(g.ntiles, g.nchunk, batch), dim3(256), 0, s, X, ldx, sx, Y, ldy, sy, npwk, g,
                     work);
  DEREK_HIP(hipGetLastError());
  const int64_t mn = (int64_t)m * n;
  hipLaunchKernelGGL(k_overlap_reduce, dim3((unsigned)std::min<int64_t>(512, ceil_div(mn, 256)), batch), dim3(256), 0,
                     s, work, g, C, ldc, sc);
  DEREK_HIP(hipGetLastError());
}

constexpr int RT_ROWS = 64;
constexpr int RT_COLS = 64;
constexpr int RT_KS = 16;

template <typename R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_rotate(const cx_t<R>* __restrict__ X, int64_t ldx, int64_t sx,
                                                const double2* __restrict__ U, int ldu, int64_t su, int m, int n,
                                                int64_t npw, const int* __restrict__ npwk, cx_t<R>* __restrict__ Z,
                                                int64_t ldz, int64_t sz) {
  using M = Mfma<R>;
  using A4 = typename M::acc;
  using C = cx_t<R>;
  // three real products of x u (Gauss): P1 = xr ur, P2 = xi ui,
  // P3 = (xr + xi)(ur + ui);  Re = P1 - P2,  Im = P3 - P1 - P2
  __shared__ R ur[RT_KS][RT_COLS + 2], ui[RT_KS][RT_COLS + 2], us[RT_KS][RT_COLS + 2];
  const int kz = blockIdx.z;
  const int64_t nrow = npwk ? (int64_t)npwk[kz] : npw;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t g0 = (int64_t)blockIdx.x * RT_ROWS + wave * 16;
  if ((int64_t)blockIdx.x * RT_ROWS >= nrow) return;
  X += kz * sx;
  U += kz * su;
  Z += kz * sz;
  const int j0 = blockIdx.y * RT_COLS;
  const int64_t grow = g0 + (lane & 15);
  const bool rok = grow < nrow;
  A4 p1[4], p2[4], p3[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    p1[b] = A4{0, 0, 0, 0};
    p2[b] = A4{0, 0, 0, 0};
    p3[b] = A4{0, 0, 0, 0};
  }
  // next slab's U (staging thread -> (k = t%16, column group t/16)) and X
  // (A operand: row grow, k = 4 kk + lane/16) held in registers
  const int sk = threadIdx.x & 15, cg = threadIdx.x >> 4;
  double2 pu[4];
  C pxv[RT_KS / 4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = cg + 16 * q;
      pu[q] = (k0 + sk < m && j0 + c < n) ? U[(int64_t)(j0 + c) * ldu + k0 + sk] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int kk = 0; kk < RT_KS / 4; ++kk) {
      const int k = k0 + kk * 4 + (lane >> 4);
      pxv[kk] = (rok && k < m) ? X[(int64_t)k * ldx + grow] : cmk<C>(0, 0);
    }
  };
  gload(0);
  for (int k0 = 0; k0 < m; k0 += RT_KS) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = cg + 16 * q;
      const R vr = (R)pu[q].x, vi = (R)pu[q].y;
      ur[sk][c] = vr;
      ui[sk][c] = vi;
      us[sk][c] = vr + vi;
    }
    C xv[RT_KS / 4];
#pragma unroll
    for (int kk = 0; kk < RT_KS / 4; ++kk) xv[kk] = pxv[kk];
    __syncthreads();
    if (k0 + RT_KS < m) gload(k0 + RT_KS);
#pragma unroll
    for (int kk = 0; kk < RT_KS / 4; ++kk) {
      const int k = kk * 4 + (lane >> 4);
      const R xs = xv[kk].x + xv[kk].y;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int c = b * 16 + (lane & 15);
        p1[b] = M::op(xv[kk].x, ur[k][c], p1[b]);
        p2[b] = M::op(xv[kk].y, ui[k][c], p2[b]);
        p3[b] = M::op(xs, us[k][c], p3[b]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int j = j0 + b * 16 + (lane & 15);
    if (j >= n) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t gg = (int64_t)blockIdx.x * RT_ROWS + wave * 16 + M::row(lane, r);
      if (gg < nrow) Z[(int64_t)j * ldz + gg] = cmk<C>(p1[b][r] - p2[b][r], p3[b][r] - p1[b][r] - p2[b][r]);
    }
  }
}

template <typename R>
void rotate(const cx_t<R>* X, int64_t ldx, int64_t sx, const double2* U, int ldu, int64_t su, int m, int n,
            int64_t npw, const int* npwk, cx_t<R>* Z, int64_t ldz, int64_t sz, int batch, hipStream_t s) {
  DEREK_REQUIRE(m > 0 && n > 0 && batch > 0, "rotate: empty block");
  dim3 grid((unsigned)ceil_div(npw, RT_ROWS), (unsigned)ceil_div(n, RT_COLS), batch);
  hipLaunchKernelGGL(k_rotate<R>, grid, dim3(256), 0, s, X, ldx, sx, U, ldu, su, m, n, npw, npwk, Z, ldz, sz);
  DEREK_HIP(hipGetLastError());
}

template void overlap<double>(const double2*, int64_t, int64_t, const double2*, int64_t, int64_t, int, int, int64_t,
                              const int*, double2*, int, int64_t, int, double2*, int64_t, bool, hipStream_t);
template void overlap<float>(const float2*, int64_t, int64_t, const float2*, int64_t, int64_t, int, int, int64_t,
                             const int*, double2*, int, int64_t, int, double2*, int64_t, bool, hipStream_t);
template void rotate<double>(const double2*, int64_t, int64_t, const double2*, int, int64_t, int, int, int64_t,
                             const int*, double2*, int64_t, int64_t, int, hipStream_t);
template void rotate<float>(const float2*, int64_t, int64_t, const double2*, int, int64_t, int, int, int64_t,
                            const int*, float2*, int64_t, int64_t, int, hipStream_t);

}  // namespace derek
