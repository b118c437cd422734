// Full-grid FFT passes, LDA XC, Hartree/Kerker and pointwise kernels for gfx950.
#include <mutex>
#include <unordered_map>

#include "launch.hpp"

namespace derek {

constexpr int NT = 256;          // threads per workgroup for the FFT passes
constexpr int TILE_LINES = 16;   // z columns per strided-axis tile

// ------------------------------------------------------------------ FFT kernels
// Contiguous lines (z axis).  Each workgroup owns `lpb` lines.
template <typename Rl, int S, int NC>
__global__ __launch_bounds__(NT) void k_fft_contig(cx_t<Rl>* __restrict__ data, int64_t nlines, int lpb,
                                                   FftPlan plan, const cx_t<Rl>* __restrict__ tw) {
  using C = cx_t<Rl>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = NC > 0 ? NC : plan.n;
  const int ls = contig_stride(N);
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + N;
  C* b1 = b0 + (size_t)lpb * ls;
  for (int t = threadIdx.x; t < N; t += NT) twl[t] = tw[t];
  const int64_t line0 = (int64_t)blockIdx.x * lpb;
  const int64_t left = nlines - line0;
  const int lines = left < lpb ? (int)left : lpb;
  const C* src = data + line0 * N;
  const int ne = lines * N;
  for (int e = threadIdx.x; e < ne; e += NT) {
    const int line = e / N;
    b0[line * ls + (e - line * N)] = src[e];
  }
  __syncthreads();
  C* res = fft_lds<S, false, 1, NT, NC>(b0, b1, twl, plan, lines);
  C* dst = data + line0 * N;
  for (int e = threadIdx.x; e < ne; e += NT) {
    const int line = e / N;
    dst[e] = res[line * ls + (e - line * N)];
  }
}

// Strided axis, tiles of TILE_LINES neighbouring inner indices.
template <typename Rl, int S, int NC>
__global__ __launch_bounds__(NT) void k_fft_tile(cx_t<Rl>* __restrict__ data, int64_t stride, int64_t tiles,
                                                 FftPlan plan, const cx_t<Rl>* __restrict__ tw) {
  using C = cx_t<Rl>;
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = NC > 0 ? NC : plan.n;
  C* twl = reinterpret_cast<C*>(smem);
  C* b0 = twl + N;
  C* b1 = b0 + (size_t)TILE_LINES * N;
  for (int t = threadIdx.x; t < N; t += NT) twl[t] = tw[t];
  const int64_t o = blockIdx.x / tiles;
  const int64_t i0 = (blockIdx.x - o * tiles) * TILE_LINES;
  const int64_t rem = stride - i0;
  const int lines = rem < TILE_LINES ? (int)rem : TILE_LINES;
  C* base = data + o * (int64_t)N * stride + i0;
  const int ne = N * TILE_LINES;
  for (int e = threadIdx.x; e < ne; e += NT) {
    const int a = e / TILE_LINES, l = e - a * TILE_LINES;
    b0[e] = l < lines ? base[(int64_t)a * stride + l] : cmk<C>(0, 0);
  }
  __syncthreads();
  C* res = fft_lds<S, true, TILE_LINES, NT, NC>(b0, b1, twl, plan, TILE_LINES);
  for (int e = threadIdx.x; e < ne; e += NT) {
    const int a = e / TILE_LINES, l = e - a * TILE_LINES;
    if (l < lines) base[(int64_t)a * stride + l] = res[e];
  }
}

void allow_dynamic_lds(const void* kern, size_t bytes) {
  static std::mutex mu;
  static std::unordered_map<const void*, size_t> done;
  std::lock_guard<std::mutex> lock(mu);
  size_t& have = done[kern];
  if (bytes <= have) return;
  DEREK_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  have = bytes;
}

template <typename K> static void allow_lds(K kern, size_t bytes) {
  // gfx950 has 160 KB of LDS per CU; kernels above 64 KB must opt in (once
  // per kernel, before any graph capture replays the launch)
  if (bytes > 65536) allow_dynamic_lds(reinterpret_cast<const void*>(kern), bytes);
}

static void check_lds(size_t bytes, int n) {
  DEREK_REQUIRE(bytes <= 160 * 1024, "FFT length " + std::to_string(n) + " exceeds the LDS budget");
}

template <typename Rl, int S, int NC>
static void launch_contig(cx_t<Rl>* data, int64_t nlines, const Axis<Rl>& ax, hipStream_t s) {
  const int N = ax.n;
  int lpb = 1024 / N;
  if (lpb < 1) lpb = 1;
  const size_t lds = sizeof(cx_t<Rl>) * ((size_t)N + 2 * (size_t)lpb * contig_stride(N));
  check_lds(lds, N);
  auto kern = k_fft_contig<Rl, S, NC>;
  allow_lds(kern, lds);
  const int64_t blocks = ceil_div(nlines, lpb);
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(NT), lds, s, data, nlines, lpb, ax.plan, ax.tw);
  DEREK_HIP(hipGetLastError());
}

template <typename Rl, int S, int NC>
static void launch_tile(cx_t<Rl>* data, int64_t nouter, int64_t stride, const Axis<Rl>& ax, hipStream_t s) {
  const int N = ax.n;
  const size_t lds = sizeof(cx_t<Rl>) * ((size_t)N + 2 * (size_t)TILE_LINES * N);
  check_lds(lds, N);
  auto kern = k_fft_tile<Rl, S, NC>;
  allow_lds(kern, lds);
  const int64_t tiles = ceil_div(stride, TILE_LINES);
  hipLaunchKernelGGL(kern, dim3((unsigned)(nouter * tiles)), dim3(NT), lds, s, data, stride, tiles, ax.plan, ax.tw);
  DEREK_HIP(hipGetLastError());
}

// Lengths with a compile-time plan (the benchmark grids); others run the
// runtime plan.
#define DEREK_FFT_DISPATCH(n, CALL) \
  switch (n) {                      \
    case 32: CALL(32); break;       \
    case 72: CALL(72); break;       \
    case 108: CALL(108); break;     \
    case 144: CALL(144); break;     \
    default: CALL(0); break;        \
  }

template <typename R> void fft_z(cx_t<R>* data, int64_t nlines, const Axis<R>& ax, int sign, hipStream_t s) {
  if (sign < 0) {
#define C_(NC) launch_contig<R, -1, NC>(data, nlines, ax, s)
    DEREK_FFT_DISPATCH(ax.n, C_)
#undef C_
  } else {
#define C_(NC) launch_contig<R, +1, NC>(data, nlines, ax, s)
    DEREK_FFT_DISPATCH(ax.n, C_)
#undef C_
  }
}

template <typename R>
void fft_strided(cx_t<R>* data, int64_t nouter, int64_t stride, const Axis<R>& ax, int sign, hipStream_t s) {
  if (sign < 0) {
#define C_(NC) launch_tile<R, -1, NC>(data, nouter, stride, ax, s)
    DEREK_FFT_DISPATCH(ax.n, C_)
#undef C_
  } else {
#define C_(NC) launch_tile<R, +1, NC>(data, nouter, stride, ax, s)
    DEREK_FFT_DISPATCH(ax.n, C_)
#undef C_
  }
}

template void fft_z<double>(double2*, int64_t, const Axis<double>&, int, hipStream_t);
template void fft_z<float>(float2*, int64_t, const Axis<float>&, int, hipStream_t);
template void fft_strided<double>(double2*, int64_t, int64_t, const Axis<double>&, int, hipStream_t);
template void fft_strided<float>(float2*, int64_t, int64_t, const Axis<float>&, int, hipStream_t);
// ------------------------------------------------------------------ pointwise
static unsigned grid1d(int64_t n, int per = 256) {
  int64_t b = ceil_div(n, per);
  if (b > (int64_t)1 << 30) b = (int64_t)1 << 30;
  return (unsigned)(b < 1 ? 1 : b);
}

template <typename C> __global__ void k_zero(C* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = cmk<C>(0, 0);
}
template <typename R> void zero(cx_t<R>* p, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero<cx_t<R>>, dim3(grid1d(n)), dim3(256), 0, s, p, n);
  DEREK_HIP(hipGetLastError());
}
template void zero<double>(double2*, int64_t, hipStream_t);
template void zero<float>(float2*, int64_t, hipStream_t);

// ------------------------------------------------------------------ LDA XC
// Slater exchange + Perdew-Zunger 1981 correlation, spin-unpolarised, with
// libxc's screening (skip n < 1e-15; exchange zero when n/2 <= 1e-15).
__device__ __forceinline__ void lda_point(double n, double& eps, double& v) {
  eps = 0.0;
  v = 0.0;
  const double thr = 1e-15;
  if (!(n >= thr)) return;
  if (n / 2.0 > thr) {
    const double cx = -0.75 * 0.98474502184269654115;  // (3/pi)^(1/3)
    const double n13 = cbrt(n);
    eps = cx * n13;
    v = (4.0 / 3.0) * cx * n13;
  }
  const double rs = cbrt(3.0 / (4.0 * 3.141592653589793238462643383279502884197 * n));
  double ec, vc;
  if (rs >= 1.0) {
    const double g = -0.1423, b1 = 1.0529, b2 = 0.3334;
    const double sr = sqrt(rs);
    const double den = 1.0 + b1 * sr + b2 * rs;
    ec = g / den;
    vc = ec * (1.0 + (7.0 / 6.0) * b1 * sr + (4.0 / 3.0) * b2 * rs) / den;
  } else {
    const double a = 0.0311, b = -0.048, c = 0.0020, d = -0.0116;
    const double lr = log(rs);
    ec = a * lr + b + c * rs * lr + d * rs;
    vc = a * lr + (b - a / 3.0) + (2.0 / 3.0) * c * rs * lr + ((2.0 * d - c) / 3.0) * rs;
  }
  eps += ec;
  v += vc;
}

constexpr int RED_NT = 256;
constexpr int RED_BLOCKS = 1024;

__device__ __forceinline__ double block_sum(double x, double* sh) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = x;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
  }
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(RED_NT) void k_lda(const double* __restrict__ rho, double* __restrict__ vxc,
                                                double* __restrict__ partial, int64_t n) {
  __shared__ double sh[RED_NT / 64];
  double e = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)RED_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_NT) {
    double eps, v;
    const double r = rho[i];
    lda_point(r, eps, v);
    vxc[i] = v;
    e += r * eps;
  }
  const double t = block_sum(e, sh);
  if (threadIdx.x == 0 && partial) partial[blockIdx.x] = t;
}

int lda_partials(int64_t n) { return (int)std::min<int64_t>(RED_BLOCKS, ceil_div(n, RED_NT)); }

void lda_xc(const double* rho, double* vxc, double* partial, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_lda, dim3(lda_partials(n)), dim3(RED_NT), 0, s, rho, vxc, partial, n);
  DEREK_HIP(hipGetLastError());
}

__global__ void k_lda_eps(const double* __restrict__ rho, double* __restrict__ eps, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double e, v;
    lda_point(rho[i], e, v);
    eps[i] = e;
  }
}
void lda_eps(const double* rho, double* eps, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_lda_eps, dim3(grid1d(n)), dim3(256), 0, s, rho, eps, n);
  DEREK_HIP(hipGetLastError());
}

__global__ void k_real_to_cx(const double* __restrict__ in, double2* __restrict__ out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = make_double2(in[i], 0.0);
}
void real_to_cx(const double* in, double2* out, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_real_to_cx, dim3(grid1d(n)), dim3(256), 0, s, in, out, n);
  DEREK_HIP(hipGetLastError());
}

__global__ void k_cx_real(const double2* __restrict__ in, double* __restrict__ out, double scale, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = in[i].x * scale;
}
void cx_real(const double2* in, double* out, double scale, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_cx_real, dim3(grid1d(n)), dim3(256), 0, s, in, out, scale, n);
  DEREK_HIP(hipGetLastError());
}

__global__ void k_gspace_scale(double2* __restrict__ g, const double* __restrict__ gsq, double q0sq, double scale,
                               int mode, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double q = gsq[i];
    double f;
    if (mode == 0) f = q > 0.0 ? 4.0 * 3.141592653589793238462643383279502884197 / q * scale : 0.0;
    else f = q > 0.0 ? q / (q + q0sq) * scale : 0.0;
    g[i] = make_double2(g[i].x * f, g[i].y * f);
  }
}
void gspace_scale(double2* g, const double* gsq, double q0sq, double scale, int mode, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_gspace_scale, dim3(grid1d(n)), dim3(256), 0, s, g, gsq, q0sq, scale, mode, n);
  DEREK_HIP(hipGetLastError());
}

__global__ void k_add3(const double* a, const double* b, const double* c, double* out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double v = a ? a[i] : 0.0;
    if (b) v += b[i];
    if (c) v += c[i];
    out[i] = v;
  }
}
void add3(const double* a, const double* b, const double* c, double* out, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_add3, dim3(grid1d(n)), dim3(256), 0, s, a, b, c, out, n);
  DEREK_HIP(hipGetLastError());
}

__global__ __launch_bounds__(RED_NT) void k_dot(const double* __restrict__ a, const double* __restrict__ b,
                                                double* __restrict__ partial, int64_t n) {
  __shared__ double sh[RED_NT / 64];
  double e = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)RED_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_NT)
    e += a[i] * b[i];
  const double t = block_sum(e, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}
int dot_nblocks(int64_t n) { return (int)std::min<int64_t>(RED_BLOCKS, ceil_div(n, RED_NT)); }
void dot_partials(const double* a, const double* b, double* partial, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_dot, dim3(dot_nblocks(n)), dim3(RED_NT), 0, s, a, b, partial, n);
  DEREK_HIP(hipGetLastError());
}

__global__ void k_axpy(double* y, const double* x, double a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += a * x[i];
}
void axpy(double* y, const double* x, double a, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(grid1d(n)), dim3(256), 0, s, y, x, a, n);
  DEREK_HIP(hipGetLastError());
}
// z = a x + b y over n complex values (y unread when b = 0): the
// wavefunction +, - and scalar * operators and copies in one streaming pass
// (3 x 16 B per value, 2 when y is unread).
__global__ void k_wave_axpby(double2* __restrict__ z, double2 a, const double2* __restrict__ x, double2 b,
                             const double2* __restrict__ y, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 u = x[i];
    double2 r = make_double2(a.x * u.x - a.y * u.y, a.x * u.y + a.y * u.x);
    if (y) {
      const double2 v = y[i];
      r.x += b.x * v.x - b.y * v.y;
      r.y += b.x * v.y + b.y * v.x;
    }
    z[i] = r;
  }
}
void wave_axpby(double2* z, double2 a, const double2* x, double2 b, const double2* y, int64_t n, hipStream_t s) {
  const bool use_y = y && (b.x != 0.0 || b.y != 0.0);
  hipLaunchKernelGGL(k_wave_axpby, dim3(grid1d(n)), dim3(256), 0, s, z, a, x, b, use_y ? y : nullptr, n);
  DEREK_HIP(hipGetLastError());
}

__global__ void k_sub(double* out, const double* a, const double* b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = a[i] - b[i];
}
void sub(double* out, const double* a, const double* b, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_sub, dim3(grid1d(n)), dim3(256), 0, s, out, a, b, n);
  DEREK_HIP(hipGetLastError());
}

template <typename D, typename S> __global__ void k_convert(D* dst, const S* src, int64_t n) {
  using RD = decltype(D::x);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = cmk<D>((RD)src[i].x, (RD)src[i].y);
}
template <typename Rd, typename Rs> void convert(cx_t<Rd>* dst, const cx_t<Rs>* src, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL((k_convert<cx_t<Rd>, cx_t<Rs>>), dim3(grid1d(n)), dim3(256), 0, s, dst, src, n);
  DEREK_HIP(hipGetLastError());
}
template void convert<float, double>(float2*, const double2*, int64_t, hipStream_t);
template void convert<double, float>(double2*, const float2*, int64_t, hipStream_t);

__global__ void k_convert_real_f(float* dst, const double* src, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}
// t[(z*ny + y)*nx + x] = scale * v[(x*ny + y)*nz + z] (and its FP32 copy when
// t32 is given): 32 x 32 (x, z) tiles of one y through LDS.
__global__ __launch_bounds__(256) void k_transpose_xz(const double* __restrict__ v, double* __restrict__ t64,
                                                      float* __restrict__ t32, int nx, int ny, int nz, double scale) {
  __shared__ double tile[32][33];
  const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32, y = blockIdx.z;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  for (int i = ty; i < 32; i += 8) {
    const int x = x0 + i, z = z0 + tx;
    tile[i][tx] = (x < nx && z < nz) ? v[((int64_t)x * ny + y) * nz + z] * scale : 0.0;
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int z = z0 + i, x = x0 + tx;
    if (z < nz && x < nx) {
      const int64_t o = ((int64_t)z * ny + y) * nx + x;
      const double w = tile[tx][i];
      t64[o] = w;
      if (t32) t32[o] = (float)w;
    }
  }
}

void transpose_xz(const double* v, double* t64, float* t32, int nx, int ny, int nz, double scale, hipStream_t s) {
  dim3 grid((unsigned)ceil_div(nz, 32), (unsigned)ceil_div(nx, 32), (unsigned)ny);
  hipLaunchKernelGGL(k_transpose_xz, grid, dim3(256), 0, s, v, t64, t32, nx, ny, nz, scale);
}

void convert_real_f(float* dst, const double* src, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_convert_real_f, dim3(grid1d(n)), dim3(256), 0, s, dst, src, n);
  DEREK_HIP(hipGetLastError());
}

}  // namespace derek
