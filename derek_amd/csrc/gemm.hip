// Tall-skinny complex GEMMs on the CDNA4 matrix cores, batched over k-points.
//
// Overlap  C = X^H Y  (X: npw x m, Y: npw x n, npw >> m, n): split-K over
// the plane-wave index; each workgroup owns a 64x64 tile of C and a chunk of
// rows of one k-point, stages 16-row slabs of X and Y through
// LDS and accumulates three real products per complex product (Gauss's
// method: 25% fewer matrix-core operations than the four-product form) with
// 16x16x4 MFMAs (f64, or f32 for complex64 data); the next slab is loaded
// into registers while the current one is multiplied.  A second kernel sums
// the chunk partials in a fixed order in FP64 (deterministic).
// Rotation Z = X U: each workgroup owns 64 rows x 64 columns of Z; U slabs
// (FP64, rounded to R) are staged through LDS, X is read straight into the
// MFMA A operand; three real products per complex product as well.
// Work items are dealt so that the workgroups that read the same rows (the
// tiles of one row chunk, the column blocks of one row block) run on one
// XCD, which then fetches those rows from HBM once into its L2.
//
// 16x16x4 MFMA operand layout (lane l of a 64-wide wave):
//   A (16x4): A[l%16][l/16]      B (4x16): B[l/16][l%16]
//   D (16x16), f64: register r holds D[l/16 + 4*r][l%16]
//   D (16x16), f32: register r holds D[4*(l/16) + r][l%16]
#include "launch.hpp"
#include "ov_geom.hpp"

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

constexpr int OV_PAD = 66;   // LDS row pitch

constexpr int kXcds = 8;  // MI355X: 8 XCDs, each with its own L2

// Work item (0 .. total-1) of the workgroup with linear id L such that every
// XCD runs one contiguous range of items: the dispatcher deals workgroups to
// the XCDs round-robin, so id L runs on XCD L % 8 as its (L / 8)-th.
// Consecutive items share operand rows, which then stay in that XCD's L2.
__device__ __forceinline__ int64_t xcd_item(int64_t L, int64_t total) {
  const int64_t q = total / kXcds, r = total % kXcds;
  const int64_t x = L % kXcds, slot = L / kXcds;
  return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
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
  // work item: tile fastest, then row chunk, then k-point; the tiles of one
  // chunk share an XCD and read the same rows through its L2
  const int64_t w = xcd_item(blockIdx.x, gridDim.x);
  const int chunk = (int)((w / g.ntiles) % g.nchunk);
  const int kz = (int)(w / ((int64_t)g.ntiles * g.nchunk));
  X += kz * sx;
  Y += kz * sy;
  const int64_t npw = npwk ? (int64_t)npwk[kz] : g.npw;
  // tile coordinates
  int ti, tj;
  int t = (int)(w % g.ntiles);
  if (g.herm) {
    ti = 0;
    while (t >= g.tm - ti) { t -= g.tm - ti; ++ti; }
    tj = ti + t;
  } else {
    ti = t % g.tm;
    tj = t / g.tm;
  }
  const int i0 = ti * OV_T, j0 = tj * OV_T;
  const int64_t r0 = (int64_t)chunk * g.kch;
  const int64_t r1 = std::min<int64_t>(npw, r0 + g.kch);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = wave & 1, wj = wave >> 1;
  // 16 x 16 blocks wholly outside C (their operands are zero padding and
  // their results are never stored) skip the matrix cores; wave-uniform
  bool live[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) live[a][b] = i0 + wi * 32 + a * 16 < g.m && j0 + wj * 32 + b * 16 < g.n;
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
        for (int b = 0; b < 2; ++b)
          if (live[a][b]) {
            p1[a][b] = M::op(ar[a], br[b], p1[a][b]);
            p2[a][b] = M::op(ai[a], bi[b], p2[a][b]);
            p3[a][b] = M::op(as[a], bs[b], p3[a][b]);
          }
    }
    __syncthreads();
  }
  double2* P = part + ((int64_t)kz * g.nchunk + chunk) * g.m * g.n;
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

// ---- complex64 data on the BF16 matrix cores --------------------------
// Each FP32 operand value v is split into NT bf16 terms t_0 + t_1 + ...,
// t_j = bf16(v - t_0 - ... - t_{j-1}) rounded to nearest even (each
// subtraction is exact in FP32).  A real product keeps the term products
// t_a s_b with a + b < NT: NT = 1 is plain BF16 (8 significant bits),
// NT = 2 three products (about 16 bits), NT = 3 six products (about FP32's
// 24).  Accumulation is FP32 on the matrix cores, chunk sums FP64 as above.
// The 16x16x32 BF16 MFMA takes 8 consecutive k per lane:
//   A (16x32): A[l%16][8*(l/16) + e]   B (32x16): B[8*(l/16) + e][l%16]
//   D as the f32 16x16x4 form (register r holds D[4*(l/16) + r][l%16]).
// Operands are split once, when their slab is staged into LDS ([column][k
// slot], slots contiguous, so a lane's 8 values are one 16-byte read).
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int BF_P = 6;    // LDS pitch of a column in bf8 units (96 bytes; measured 3 % faster than 80 or 64)

__device__ __forceinline__ f4 mfma_bf16(bf8 a, bf8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// out[t] = term t of the 8 values v (v is consumed)
template <int NT> __device__ __forceinline__ void bf_split(float (&v)[8], bf8 (&out)[NT]) {
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const __bf16 h = (__bf16)v[e];
      out[t][e] = h;
      if (t + 1 < NT) v[e] -= (float)h;
    }
}

// acc += sum over term pairs (a, b), a + b < NT, of A_a B_b
template <int NT> __device__ __forceinline__ f4 bf_products(const bf8 (&A)[NT], const bf8 (&B)[NT], f4 acc) {
#pragma unroll
  for (int s = NT - 1; s >= 0; --s)  // smallest terms first
#pragma unroll
    for (int a = 0; a <= s; ++a) acc = mfma_bf16(A[a], B[s - a], acc);
  return acc;
}

// Stage two complex values of one column as the three Gauss planes (re, im,
// re + sgn*im), each split into NT terms, into the adjacent k slots 2*slot
// and 2*slot + 1 of LDS column `col` (one 4-byte store per plane and term).
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
template <int NT, int TT>
__device__ __forceinline__ void bf_stage2(float2 v0, float2 v1, float sgn, bf8 (*L)[NT][TT][BF_P], int col, int slot) {
  float a[3][2] = {{v0.x, v1.x}, {v0.y, v1.y}, {v0.x + sgn * v0.y, v1.x + sgn * v1.y}};
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      bf2 h;
      h[0] = (__bf16)a[p][0];
      h[1] = (__bf16)a[p][1];
      reinterpret_cast<bf2*>(&L[p][t][col][0])[slot] = h;
      if (t + 1 < NT) {
        a[p][0] -= (float)h[0];
        a[p][1] -= (float)h[1];
      }
    }
}

// TT = 64: 4 wavefronts, each a 32 x 32 block of the tile; TT = 128: 8
// wavefronts, each 64 rows x 32 columns.  The larger tile reads every operand
// row half as often per matrix-core operation (fewer tiles share a row chunk)
// and splits and stages half as many values per product.
template <int TT> constexpr int kBfWaves = TT == 64 ? 4 : 8;
template <int NT, int TT>
__global__ __launch_bounds__(64 * kBfWaves<TT>) __attribute__((amdgpu_waves_per_eu(2))) void k_overlap_bf16(
    const float2* __restrict__ X, int64_t ldx, int64_t sx, const float2* __restrict__ Y, int64_t ldy, int64_t sy,
    const int* __restrict__ npwk, OvGeom g, double2* __restrict__ part) {
  static_assert(TT == 64 || TT == 128, "tile edge 64 or 128");
  constexpr int AB = TT / 32;  // 16-row blocks per wavefront
  // planes: conj(x) y = P1 + P2 + i (P3 - P1 + P2), P1 = xr yr, P2 = xi yi,
  // P3 = (xr - xi)(yr + yi)
  __shared__ bf8 lx[3][NT][TT][BF_P], ly[3][NT][TT][BF_P];
  const int64_t w = xcd_item(blockIdx.x, gridDim.x);
  const int chunk = (int)((w / g.ntiles) % g.nchunk);
  const int kz = (int)(w / ((int64_t)g.ntiles * g.nchunk));
  X += kz * sx;
  Y += kz * sy;
  const int64_t npw = npwk ? (int64_t)npwk[kz] : g.npw;
  int ti, tj;
  int t = (int)(w % g.ntiles);
  if (g.herm) {
    ti = 0;
    while (t >= g.tm - ti) { t -= g.tm - ti; ++ti; }
    tj = ti + t;
  } else {
    ti = t % g.tm;
    tj = t / g.tm;
  }
  const int i0 = ti * TT, j0 = tj * TT;
  const int64_t r0 = (int64_t)chunk * g.kch;
  const int64_t r1 = std::min<int64_t>(npw, r0 + g.kch);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = wave & 1, wj = wave >> 1;
  const int ri = wi * (TT / 2), cj = wj * 32;  // the wavefront's rows and columns in the tile
  // 16 x 16 blocks wholly outside C (their operands are zero padding and
  // their results are never stored) skip the matrix cores; wave-uniform
  bool live[AB][2];
#pragma unroll
  for (int a = 0; a < AB; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) live[a][b] = i0 + ri + a * 16 < g.m && j0 + cj + b * 16 < g.n;
  f4 p1[AB][2], p2[AB][2], p3[AB][2];
#pragma unroll
  for (int a = 0; a < AB; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      p1[a][b] = f4{0, 0, 0, 0};
      p2[a][b] = f4{0, 0, 0, 0};
      p3[a][b] = f4{0, 0, 0, 0};
    }
  // staging: wave w, pass q: tile column 16 w + 4 q + lane/16, slab rows
  // sr = lane%16 and sr + 16, so the lanes of a column read 128 contiguous
  // bytes.  The two rows go to the adjacent k slots 2 sr and 2 sr + 1; both
  // operands use the same slot order, so the sum over k is unchanged.
  const int sr = lane & 15;
  float2 px[4][2], py[4][2];
  auto gload = [&](int64_t k0) {
    const int64_t ga = k0 + sr, gb = k0 + sr + 16;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 16 * wave + 4 * q + (lane >> 4);
      const bool xok = i0 + c < g.m, yok = j0 + c < g.n;
      const float2* xc = X + (int64_t)(xok ? i0 + c : 0) * ldx;
      const float2* yc = Y + (int64_t)(yok ? j0 + c : 0) * ldy;
      px[q][0] = (xok && ga < r1) ? xc[ga] : make_float2(0.f, 0.f);
      px[q][1] = (xok && gb < r1) ? xc[gb] : make_float2(0.f, 0.f);
      py[q][0] = (yok && ga < r1) ? yc[ga] : make_float2(0.f, 0.f);
      py[q][1] = (yok && gb < r1) ? yc[gb] : make_float2(0.f, 0.f);
    }
  };
  if (r0 < r1) gload(r0);
  const int kg = lane >> 4;
  for (int64_t k0 = r0; k0 < r1; k0 += BF_KS) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 16 * wave + 4 * q + (lane >> 4);
      bf_stage2<NT, TT>(px[q][0], px[q][1], -1.f, lx, c, sr);
      bf_stage2<NT, TT>(py[q][0], py[q][1], 1.f, ly, c, sr);
    }
    __syncthreads();
    if (k0 + BF_KS < r1) gload(k0 + BF_KS);
    bf8 B[2][3][NT];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int jj = cj + b * 16 + (lane & 15);
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < NT; ++q) B[b][p][q] = ly[p][q][jj][kg];
    }
#pragma unroll
    for (int a = 0; a < AB; ++a) {
      const int ii = ri + a * 16 + (lane & 15);
      bf8 A[3][NT];
#pragma unroll
      for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < NT; ++q) A[p][q] = lx[p][q][ii][kg];
#pragma unroll
      for (int b = 0; b < 2; ++b)
        if (live[a][b]) {
          p1[a][b] = bf_products<NT>(A[0], B[b][0], p1[a][b]);
          p2[a][b] = bf_products<NT>(A[1], B[b][1], p2[a][b]);
          p3[a][b] = bf_products<NT>(A[2], B[b][2], p3[a][b]);
        }
    }
    __syncthreads();
  }
  double2* P = part + ((int64_t)kz * g.nchunk + chunk) * g.m * g.n;
#pragma unroll
  for (int a = 0; a < AB; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + ri + a * 16 + Mfma<float>::row(lane, r);
        const int j = j0 + cj + b * 16 + (lane & 15);
        const double q1 = (double)p1[a][b][r], q2 = (double)p2[a][b][r], q3 = (double)p3[a][b][r];
        if (i < g.m && j < g.n) P[(int64_t)j * g.m + i] = make_double2(q1 + q2, q3 - q1 + q2);
      }
}

// resident 256-thread workgroups of `kern` on the current device (CUs x
// per-CU occupancy)
static int resident_slots(const void* kern, int threads) {
  int dev = 0;
  DEREK_HIP(hipGetDevice(&dev));
  hipDeviceProp_t p;
  DEREK_HIP(hipGetDeviceProperties(&p, dev));
  int per = 0;
  DEREK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, 0));
  return std::max(1, per) * std::max(1, p.multiProcessorCount);
}
template <typename R> static int ov_slots() {
  static const int slots = resident_slots(reinterpret_cast<const void*>(k_overlap<R>), 256);
  return slots;
}
template <int NT, int TT> static int ov_slots_bf16() {
  static const int slots = resident_slots(reinterpret_cast<const void*>(k_overlap_bf16<NT, TT>), 64 * kBfWaves<TT>);
  return slots;
}

// resident slots of every overlap kernel (FP64, FP32, BF16 with 1..3 terms
// on 64- and 128-wide tiles)
constexpr int kOvKernels = 8;
static void overlap_slots(int out[kOvKernels]) {
  out[0] = ov_slots<double>();
  out[1] = ov_slots<float>();
  out[2] = ov_slots_bf16<1, 64>();
  out[3] = ov_slots_bf16<2, 64>();
  out[4] = ov_slots_bf16<3, 64>();
  out[5] = ov_slots_bf16<1, 128>();
  out[6] = ov_slots_bf16<2, 128>();
  out[7] = out[6];
}

int64_t overlap_work_elems(int m, int n, int64_t npw, int batch) {
  int slots[kOvKernels];
  overlap_slots(slots);
  return ov_work_elems(m, n, npw, batch, slots, kOvKernels);
}

int64_t engine_overlap_work_elems(int nb, int nk, int kc, int64_t npw) {
  int slots[kOvKernels];
  overlap_slots(slots);
  return ov_engine_work_elems(nb, nk, kc, npw, slots, kOvKernels);
}

// C[kz] = sum over chunks.  A workgroup takes RD_E consecutive elements;
// its RD_G thread groups each sum every RD_G-th chunk (in increasing
// order), and the group sums are added in group order through LDS, so the
// result does not depend on the launch.
constexpr int RD_E = 32, RD_G = 8;
__global__ __launch_bounds__(RD_E * RD_G) void k_overlap_reduce(const double2* __restrict__ part, OvGeom g,
                                                                double2* __restrict__ C, int ldc, int64_t sc) {
  __shared__ double2 red[RD_G][RD_E];
  const int kz = blockIdx.y;
  const int64_t mn = (int64_t)g.m * g.n;
  const double2* pk = part + (int64_t)kz * g.nchunk * mn;
  double2* Ck = C + kz * sc;
  const int el = threadIdx.x % RD_E, grp = threadIdx.x / RD_E;
  for (int64_t e0 = (int64_t)blockIdx.x * RD_E; e0 < mn; e0 += (int64_t)gridDim.x * RD_E) {
    const int64_t e = e0 + el;
    const int i = e < mn ? (int)(e % g.m) : 0, j = e < mn ? (int)(e / g.m) : 0;
    // Hermitian: only the upper triangle's partials are summed (read in
    // order, so the loads stay coalesced); each sum also gives its mirror
    // element, the conjugate (exactly Hermitian output, real diagonal)
    const bool ok = e < mn && !(g.herm && i > j);
    double sr = 0.0, si = 0.0;
    if (ok)
      for (int c = grp; c < g.nchunk; c += RD_G) {
        const double2 v = pk[(int64_t)c * mn + e];
        sr += v.x;
        si += v.y;
      }
    red[grp][el] = make_double2(sr, si);
    __syncthreads();
    if (grp == 0 && ok) {
      for (int q = 1; q < RD_G; ++q) {
        sr += red[q][el].x;
        si += red[q][el].y;
      }
      if (g.herm && i == j) si = 0.0;
      Ck[(int64_t)j * ldc + i] = make_double2(sr, si);
      if (g.herm && i < j) Ck[(int64_t)i * ldc + j] = make_double2(sr, -si);
    }
    __syncthreads();
  }
}

template <typename R>
void overlap(const cx_t<R>* X, int64_t ldx, int64_t sx, const cx_t<R>* Y, int64_t ldy, int64_t sy, int m, int n,
             int64_t npw, const int* npwk, double2* C, int ldc, int64_t sc, int batch, double2* work,
             int64_t work_cap, bool hermitian, hipStream_t s, int bf16_terms, int64_t max_chunks, int rounds,
             int tile) {
  DEREK_REQUIRE(m > 0 && n > 0 && batch > 0, "overlap: empty block");
  DEREK_REQUIRE(!hermitian || m == n, "overlap: hermitian needs a square result");
  DEREK_REQUIRE(bf16_terms == 0 || (std::is_same<R, float>::value && bf16_terms >= 1 && bf16_terms <= 3),
                "overlap: bf16 terms must be 0, or 1..3 for complex64 data");
  DEREK_REQUIRE(tile == 64 || tile == 128, "overlap: tile must be 64 or 128");
  // the 128-wide tile exists for the BF16 kernels with 1 or 2 terms (three
  // terms of a 128-wide tile need 180 KB of LDS)
  const int tt = bf16_terms && bf16_terms < 3 ? tile : 64;
  const int slots = bf16_terms == 1   ? (tt == 128 ? ov_slots_bf16<1, 128>() : ov_slots_bf16<1, 64>())
                    : bf16_terms == 2 ? (tt == 128 ? ov_slots_bf16<2, 128>() : ov_slots_bf16<2, 64>())
                    : bf16_terms == 3 ? ov_slots_bf16<3, 64>()
                                      : ov_slots<R>();
  DEREK_REQUIRE(rounds == 1 || rounds == 2 || rounds == 4, "overlap: rounds must be 1, 2 or 4");
  OvGeom g = ov_geom(m, n, npw, hermitian, batch, slots, bf16_terms ? BF_KS : OV_KS, rounds, tt);
  ov_limit_chunks(g, max_chunks, bf16_terms ? BF_KS : OV_KS);
  DEREK_REQUIRE((int64_t)batch * g.nchunk * m * n <= work_cap, "overlap: partial-sum buffer too small");
  const int64_t items = (int64_t)g.ntiles * g.nchunk * batch;
  DEREK_REQUIRE(items < (int64_t)1 << 31, "overlap: too many workgroups");
  const dim3 grid((unsigned)items);
  if constexpr (std::is_same<R, float>::value) {
    const dim3 b64(64 * kBfWaves<64>), b128(64 * kBfWaves<128>);
    if (bf16_terms == 1 && tt == 64)
      hipLaunchKernelGGL((k_overlap_bf16<1, 64>), grid, b64, 0, s, X, ldx, sx, Y, ldy, sy, npwk, g, work);
    else if (bf16_terms == 2 && tt == 64)
      hipLaunchKernelGGL((k_overlap_bf16<2, 64>), grid, b64, 0, s, X, ldx, sx, Y, ldy, sy, npwk, g, work);
    else if (bf16_terms == 3 && tt == 64)
      hipLaunchKernelGGL((k_overlap_bf16<3, 64>), grid, b64, 0, s, X, ldx, sx, Y, ldy, sy, npwk, g, work);
    else if (bf16_terms == 1)
      hipLaunchKernelGGL((k_overlap_bf16<1, 128>), grid, b128, 0, s, X, ldx, sx, Y, ldy, sy, npwk, g, work);
    else if (bf16_terms == 2)
      hipLaunchKernelGGL((k_overlap_bf16<2, 128>), grid, b128, 0, s, X, ldx, sx, Y, ldy, sy, npwk, g, work);
  }
  if (bf16_terms == 0)
    hipLaunchKernelGGL(k_overlap<R>, grid, dim3(256), 0, s, X, ldx, sx, Y, ldy, sy, npwk, g, work);
  DEREK_HIP(hipGetLastError());
  const int64_t mn = (int64_t)m * n;
  hipLaunchKernelGGL(k_overlap_reduce, dim3((unsigned)std::min<int64_t>(4096, ceil_div(mn, RD_E)), batch),
                     dim3(RD_E * RD_G), 0, s, work, g, C, ldc, sc);
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
  // work item: column block fastest, then row block, then k-point; the
  // column blocks of one row block share an XCD and read its X rows through
  // that XCD's L2
  const int ncb = (int)ceil_div(n, RT_COLS);
  const int64_t nrb = ceil_div(npw, RT_ROWS);
  const int64_t w = xcd_item(blockIdx.x, gridDim.x);
  const int64_t rb = (w / ncb) % nrb;
  const int kz = (int)(w / (ncb * nrb));
  const int64_t nrow = npwk ? (int64_t)npwk[kz] : npw;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t g0 = rb * RT_ROWS + wave * 16;
  if (rb * RT_ROWS >= nrow) return;
  X += kz * sx;
  U += kz * su;
  Z += kz * sz;
  const int j0 = (int)(w % ncb) * RT_COLS;
  // 16-column blocks past n, and a wavefront whose 16 rows are all past the
  // k-point's plane waves, skip the matrix cores (wave-uniform)
  bool live[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) live[b] = j0 + b * 16 < n && rb * RT_ROWS + wave * 16 < nrow;
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
    R xs[RT_KS / 4];
#pragma unroll
    for (int kk = 0; kk < RT_KS / 4; ++kk) xs[kk] = xv[kk].x + xv[kk].y;
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (live[b]) {
        const int c = b * 16 + (lane & 15);
#pragma unroll
        for (int kk = 0; kk < RT_KS / 4; ++kk) {
          const int k = kk * 4 + (lane >> 4);
          p1[b] = M::op(xv[kk].x, ur[k][c], p1[b]);
          p2[b] = M::op(xv[kk].y, ui[k][c], p2[b]);
          p3[b] = M::op(xs[kk], us[k][c], p3[b]);
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
      const int64_t gg = rb * RT_ROWS + wave * 16 + M::row(lane, r);
      if (gg < nrow) Z[(int64_t)j * ldz + gg] = cmk<C>(p1[b][r] - p2[b][r], p3[b][r] - p1[b][r] - p2[b][r]);
    }
  }
}

// Z = X U with complex64 X on the BF16 matrix cores (terms as in
// k_overlap_bf16): U's slab (rounded to FP32, split) is staged in LDS, X's
// A operand (8 consecutive k of the lane's row) is split in registers.
template <int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_rotate_bf16(
    const float2* __restrict__ X, int64_t ldx, int64_t sx, const double2* __restrict__ U, int ldu, int64_t su, int m,
    int n, int64_t npw, const int* __restrict__ npwk, float2* __restrict__ Z, int64_t ldz, int64_t sz) {
  // x u = P1 - P2 + i (P3 - P1 - P2), P1 = xr ur, P2 = xi ui, P3 = (xr + xi)(ur + ui)
  __shared__ bf8 lu[3][NT][64][BF_P];
  const int ncb = (int)ceil_div(n, RT_COLS);
  const int64_t nrb = ceil_div(npw, RT_ROWS);
  const int64_t w = xcd_item(blockIdx.x, gridDim.x);
  const int64_t rb = (w / ncb) % nrb;
  const int kz = (int)(w / (ncb * nrb));
  const int64_t nrow = npwk ? (int64_t)npwk[kz] : npw;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (rb * RT_ROWS >= nrow) return;
  X += kz * sx;
  U += kz * su;
  Z += kz * sz;
  const int j0 = (int)(w % ncb) * RT_COLS;
  // 16-column blocks past n, and a wavefront whose 16 rows are all past the
  // k-point's plane waves, skip the matrix cores (wave-uniform)
  bool live[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) live[b] = j0 + b * 16 < n && rb * RT_ROWS + wave * 16 < nrow;
  const int64_t grow = rb * RT_ROWS + wave * 16 + (lane & 15);
  const bool rok = grow < nrow;
  const int kg = lane >> 4;
  f4 p1[4], p2[4], p3[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    p1[b] = f4{0, 0, 0, 0};
    p2[b] = f4{0, 0, 0, 0};
    p3[b] = f4{0, 0, 0, 0};
  }
  // U staging: wave w, pass q: slab column 16 w + 4 q + lane/16, rows
  // sr = lane%16 and sr + 16 (consecutive lanes, consecutive rows) into the
  // k slots 2 sr, 2 sr + 1.  The A operand's element e (slot 8 kg + e) is
  // therefore X's column k0 + 4 kg + e/2 + 16 (e%2).
  const int sr = lane & 15;
  double2 pu[4][2];
  float2 pxv[8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 16 * wave + 4 * q + (lane >> 4);
      const bool uok = j0 + c < n;
      const double2* uc = U + (int64_t)(uok ? j0 + c : 0) * ldu;
      pu[q][0] = (uok && k0 + sr < m) ? uc[k0 + sr] : make_double2(0.0, 0.0);
      pu[q][1] = (uok && k0 + sr + 16 < m) ? uc[k0 + sr + 16] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + 4 * kg + (e >> 1) + 16 * (e & 1);
      pxv[e] = (rok && k < m) ? X[(int64_t)k * ldx + grow] : make_float2(0.f, 0.f);
    }
  };
  gload(0);
  for (int k0 = 0; k0 < m; k0 += BF_KS) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = 16 * wave + 4 * q + (lane >> 4);
      bf_stage2<NT, 64>(make_float2((float)pu[q][0].x, (float)pu[q][0].y),
                    make_float2((float)pu[q][1].x, (float)pu[q][1].y), 1.f, lu, c, sr);
    }
    bf8 A[3][NT];
    {
      float r[8], i[8], q[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        r[e] = pxv[e].x;
        i[e] = pxv[e].y;
        q[e] = pxv[e].x + pxv[e].y;
      }
      bf_split<NT>(r, A[0]);
      bf_split<NT>(i, A[1]);
      bf_split<NT>(q, A[2]);
    }
    __syncthreads();
    if (k0 + BF_KS < m) gload(k0 + BF_KS);
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (live[b]) {
        const int c = b * 16 + (lane & 15);
        bf8 B[3][NT];
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
          for (int q = 0; q < NT; ++q) B[p][q] = lu[p][q][c][kg];
        p1[b] = bf_products<NT>(A[0], B[0], p1[b]);
        p2[b] = bf_products<NT>(A[1], B[1], p2[b]);
        p3[b] = bf_products<NT>(A[2], B[2], p3[b]);
      }
    __syncthreads();
  }
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int j = j0 + b * 16 + (lane & 15);
    if (j >= n) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t gg = rb * RT_ROWS + wave * 16 + Mfma<float>::row(lane, r);
      if (gg < nrow) Z[(int64_t)j * ldz + gg] = make_float2(p1[b][r] - p2[b][r], p3[b][r] - p1[b][r] - p2[b][r]);
    }
  }
}

template <typename R>
void rotate(const cx_t<R>* X, int64_t ldx, int64_t sx, const double2* U, int ldu, int64_t su, int m, int n,
            int64_t npw, const int* npwk, cx_t<R>* Z, int64_t ldz, int64_t sz, int batch, hipStream_t s,
            int bf16_terms) {
  DEREK_REQUIRE(m > 0 && n > 0 && batch > 0, "rotate: empty block");
  DEREK_REQUIRE(bf16_terms == 0 || (std::is_same<R, float>::value && bf16_terms >= 1 && bf16_terms <= 3),
                "rotate: bf16 terms must be 0, or 1..3 for complex64 data");
  const int64_t items = ceil_div(npw, RT_ROWS) * ceil_div(n, RT_COLS) * batch;
  DEREK_REQUIRE(items < (int64_t)1 << 31, "rotate: too many workgroups");
  const dim3 grid((unsigned)items);
  if constexpr (std::is_same<R, float>::value) {
    if (bf16_terms == 1)
      hipLaunchKernelGGL(k_rotate_bf16<1>, grid, dim3(256), 0, s, X, ldx, sx, U, ldu, su, m, n, npw, npwk, Z, ldz, sz);
    else if (bf16_terms == 2)
      hipLaunchKernelGGL(k_rotate_bf16<2>, grid, dim3(256), 0, s, X, ldx, sx, U, ldu, su, m, n, npw, npwk, Z, ldz, sz);
    else if (bf16_terms == 3)
      hipLaunchKernelGGL(k_rotate_bf16<3>, grid, dim3(256), 0, s, X, ldx, sx, U, ldu, su, m, n, npw, npwk, Z, ldz, sz);
  }
  if (bf16_terms == 0)
    hipLaunchKernelGGL(k_rotate<R>, grid, dim3(256), 0, s, X, ldx, sx, U, ldu, su, m, n, npw, npwk, Z, ldz, sz);
  DEREK_HIP(hipGetLastError());
}

template void overlap<double>(const double2*, int64_t, int64_t, const double2*, int64_t, int64_t, int, int, int64_t,
                              const int*, double2*, int, int64_t, int, double2*, int64_t, bool, hipStream_t, int,
                              int64_t, int, int);
template void overlap<float>(const float2*, int64_t, int64_t, const float2*, int64_t, int64_t, int, int, int64_t,
                             const int*, double2*, int, int64_t, int, double2*, int64_t, bool, hipStream_t, int,
                             int64_t, int, int);
template void rotate<double>(const double2*, int64_t, int64_t, const double2*, int, int64_t, int, int, int64_t,
                             const int*, double2*, int64_t, int64_t, int, hipStream_t, int);
template void rotate<float>(const float2*, int64_t, int64_t, const double2*, int, int64_t, int, int, int64_t,
                            const int*, float2*, int64_t, int64_t, int, hipStream_t, int);

}  // namespace derek
