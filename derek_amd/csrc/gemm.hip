// Tall-skinny complex GEMMs on the CDNA4 matrix cores.
//
// Overlap  C = X^H Y  (X: npw x m, Y: npw x n, npw >> m, n): split-K over
// the plane-wave index; each workgroup owns a 64x64 tile of C and a chunk of
// rows, stages 16-row slabs of X and Y through LDS and accumulates the four
// real products of the complex product with v_mfma_f64_16x16x4f64; a second
// kernel sums the chunk partials in a fixed order (deterministic).
// Rotation Z = X U: each workgroup owns 64 rows x 64 columns of Z; U slabs
// are staged through LDS, X is read straight into the MFMA A operand.
//
// MFMA f64 16x16x4 operand layout (lane l of a 64-wide wave):
//   A (16x4): A[l%16][l/16]      B (4x16): B[l/16][l%16]
//   D (16x16): register r holds D[l/16 + 4*r][l%16]
#include "launch.hpp"

namespace derek {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma64(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int OV_T = 64;     // C tile edge
constexpr int OV_KS = 16;    // rows per LDS slab
constexpr int OV_PAD = 66;   // LDS row pitch (doubles)

struct OvGeom {
  int m, n, tm, tn, ntiles, nchunk;
  int64_t npw, kch;
  bool herm;
};

static OvGeom ov_geom(int m, int n, int64_t npw, bool herm) {
  OvGeom g{};
  g.m = m;
  g.n = n;
  g.npw = npw;
  g.herm = herm;
  g.tm = (int)ceil_div(m, OV_T);
  g.tn = (int)ceil_div(n, OV_T);
  g.ntiles = herm ? g.tm * (g.tm + 1) / 2 : g.tm * g.tn;
  int64_t target = std::max<int64_t>(1, 2048 / g.ntiles);
  int64_t kch = ceil_div(std::max<int64_t>(npw, 1), target);
  kch = std::max<int64_t>(256, ceil_div(kch, OV_KS) * OV_KS);
  g.kch = kch;
  g.nchunk = (int)std::max<int64_t>(1, ceil_div(npw, kch));
  return g;
}

int64_t overlap_work_elems(int m, int n, int64_t npw) {
  OvGeom g = ov_geom(m, n, npw, false);
  OvGeom h = ov_geom(m, n, npw, m == n);
  return std::max<int64_t>((int64_t)g.nchunk * m * n, (int64_t)h.nchunk * m * n);
}

__global__ __launch_bounds__(256) void k_overlap_f64(const double2* __restrict__ X, int64_t ldx,
                                                     const double2* __restrict__ Y, int64_t ldy, OvGeom g,
                                                     double2* __restrict__ part) {
  __shared__ double xr[OV_KS][OV_PAD], xi[OV_KS][OV_PAD], yr[OV_KS][OV_PAD], yi[OV_KS][OV_PAD];
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
  const int64_t r1 = std::min<int64_t>(g.npw, r0 + g.kch);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = wave & 1, wj = wave >> 1;
  d4 cr[2][2], ci[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      cr[a][b] = d4{0, 0, 0, 0};
      ci[a][b] = d4{0, 0, 0, 0};
    }
  const int lrow = threadIdx.x & 15, lcol = threadIdx.x >> 4;  // 16 rows x 16 column groups
  for (int64_t k0 = r0; k0 < r1; k0 += OV_KS) {
    const int64_t gr = k0 + lrow;
    const bool rok = gr < r1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = lcol + 16 * q;
      double2 vx = make_double2(0.0, 0.0), vy = make_double2(0.0, 0.0);
      if (rok && i0 + c < g.m) vx = X[(int64_t)(i0 + c) * ldx + gr];
      if (rok && j0 + c < g.n) vy = Y[(int64_t)(j0 + c) * ldy + gr];
      xr[lrow][c] = vx.x;
      xi[lrow][c] = vx.y;
      yr[lrow][c] = vy.x;
      yi[lrow][c] = vy.y;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < OV_KS / 4; ++kk) {
      const int k = kk * 4 + (lane >> 4);
      double ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int ii = wi * 32 + a * 16 + (lane & 15);
        ar[a] = xr[k][ii];
        ai[a] = xi[k][ii];
        const int jj = wj * 32 + a * 16 + (lane & 15);
        br[a] = yr[k][jj];
        bi[a] = yi[k][jj];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          // conj(x) * y = (xr yr + xi yi) + i (xr yi - xi yr)
          cr[a][b] = mfma64(ar[a], br[b], cr[a][b]);
          cr[a][b] = mfma64(ai[a], bi[b], cr[a][b]);
          ci[a][b] = mfma64(ar[a], bi[b], ci[a][b]);
          ci[a][b] = mfma64(-ai[a], br[b], ci[a][b]);
        }
    }
    __syncthreads();
  }
  double2* P = part + (int64_t)blockIdx.y * g.m * g.n;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wi * 32 + a * 16 + (lane >> 4) + 4 * r;
        const int j = j0 + wj * 32 + b * 16 + (lane & 15);
        if (i < g.m && j < g.n) P[(int64_t)j * g.m + i] = make_double2(cr[a][b][r], ci[a][b][r]);
      }
}

__global__ void k_overlap_reduce(const double2* __restrict__ part, OvGeom g, double2* __restrict__ C, int ldc) {
  const int64_t mn = (int64_t)g.m * g.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < mn; e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % g.m), j = (int)(e / g.m);
    // Hermitian: the strict lower triangle is the conjugate of the upper one
    // (exactly Hermitian output, real diagonal)
    const bool mirror = g.herm && i > j;
    const int64_t src = mirror ? (int64_t)i * g.m + j : (int64_t)j * g.m + i;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < g.nchunk; ++c) {
      const double2 v = part[(int64_t)c * mn + src];
      sr += v.x;
      si += v.y;
    }
    if (g.herm && i == j) si = 0.0;
    C[(int64_t)j * ldc + i] = mirror ? make_double2(sr, -si) : make_double2(sr, si);
  }
}

void overlap_f64(const double2* X, int64_t ldx, const double2* Y, int64_t ldy, int m, int n, int64_t npw,
                 double2* C, int ldc, double2* work, bool hermitian, hipStream_t s) {
  DEREK_REQUIRE(m > 0 && n > 0, "overlap: empty block");
  DEREK_REQUIRE(!hermitian || m == n, "overlap: hermitian needs a square result");
  OvGeom g = ov_geom(m, n, npw, hermitian);
  hipLaunchKernelGGL(k_overlap_f64, dim3(g.ntiles, g.nchunk), dim3(256), 0, s, X, ldx, Y, ldy, g, work);
  DEREK_HIP(hipGetLastError());
  const int64_t mn = (int64_t)m * n;
  hipLaunchKernelGGL(k_overlap_reduce, dim3((unsigned)std::min<int64_t>(4096, ceil_div(mn, 256))), dim3(256), 0, s,
                     work, g, C, ldc);
  DEREK_HIP(hipGetLastError());
}

constexpr int RT_ROWS = 64;
constexpr int RT_COLS = 64;
constexpr int RT_KS = 16;

__global__ __launch_bounds__(256) void k_rotate_f64(const double2* __restrict__ X, int64_t ldx,
                                                    const double2* __restrict__ U, int ldu, int m, int n,
                                                    int64_t npw, double2* __restrict__ Z, int64_t ldz) {
  __shared__ double ur[RT_KS][RT_COLS + 2], ui[RT_KS][RT_COLS + 2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t g0 = (int64_t)blockIdx.x * RT_ROWS + wave * 16;
  const int j0 = blockIdx.y * RT_COLS;
  const int64_t grow = g0 + (lane & 15);
  const bool rok = grow < npw;
  d4 cr[4], ci[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    cr[b] = d4{0, 0, 0, 0};
    ci[b] = d4{0, 0, 0, 0};
  }
  for (int k0 = 0; k0 < m; k0 += RT_KS) {
    // stage U[k0:k0+16, j0:j0+64]: thread -> (k = t%16, column group t/16)
    {
      const int kk = threadIdx.x & 15, cg = threadIdx.x >> 4;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int c = cg + 16 * q;
        double2 v = make_double2(0.0, 0.0);
        if (k0 + kk < m && j0 + c < n) v = U[(int64_t)(j0 + c) * ldu + k0 + kk];
        ur[kk][c] = v.x;
        ui[kk][c] = v.y;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < RT_KS / 4; ++kk) {
      const int k = kk * 4 + (lane >> 4);
      double2 xv = make_double2(0.0, 0.0);
      if (rok && k0 + k < m) xv = X[(int64_t)(k0 + k) * ldx + grow];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const double vr = ur[k][b * 16 + (lane & 15)];
        const double vi = ui[k][b * 16 + (lane & 15)];
        // x * u = (xr ur - xi ui) + i (xr ui + xi ur)
        cr[b] = mfma64(xv.x, vr, cr[b]);
        cr[b] = mfma64(-xv.y, vi, cr[b]);
        ci[b] = mfma64(xv.x, vi, ci[b]);
        ci[b] = mfma64(xv.y, vr, ci[b]);
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
      const int64_t gg = g0 + (lane >> 4) + 4 * r;
      if (gg < npw) Z[(int64_t)j * ldz + gg] = make_double2(cr[b][r], ci[b][r]);
    }
  }
}

void rotate_f64(const double2* X, int64_t ldx, const double2* U, int ldu, int m, int n, int64_t npw, double2* Z,
                int64_t ldz, hipStream_t s) {
  DEREK_REQUIRE(m > 0 && n > 0, "rotate: empty block");
  dim3 grid((unsigned)ceil_div(npw, RT_ROWS), (unsigned)ceil_div(n, RT_COLS));
  hipLaunchKernelGGL(k_rotate_f64, grid, dim3(256), 0, s, X, ldx, U, ldu, m, n, npw, Z, ldz);
  DEREK_HIP(hipGetLastError());
}

}  // namespace derek
