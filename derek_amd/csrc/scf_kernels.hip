// Device-side pieces of the batched SCF step: per-band statistics and the
// TPA correction for a batch of k-points, initial wavefunctions, Fermi-Dirac
// occupations, scalar reductions into the step report and the Pulay mixer.
// Nothing here needs the host between launches, so a whole SCF step can be
// enqueued (or captured in a graph) and read back once.
#include "scf_kernels.hpp"

namespace derek {

constexpr int NT = 256;

static unsigned blocks(int64_t n, int per = NT) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, ceil_div(n, per)));
}

__device__ __forceinline__ double wsum(double x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  return x;  // lane 0
}

// Sum over a workgroup of NT threads, result valid in thread 0.
__device__ __forceinline__ double bsum(double x, double* sh) {
  x = wsum(x);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = x;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += sh[w];
  return r;
}

// ------------------------------------------------------------------ band statistics
// partial[((kz*nb + b)*nchunk + c)*2 + {0,1}] = chunk sums of kin|x|^2 and
// |hx - eps x|^2 for band b of k-point k0+kz.
template <typename C>
__global__ __launch_bounds__(NT) void k_band_stats(BandBlock<C> X, BandBlock<C> HX, const PwMaps* __restrict__ mt,
                                                   int k0, const double* __restrict__ eps, int nb, int nchunk,
                                                   double* __restrict__ partial) {
  __shared__ double sh[NT / 64];
  const int kz = blockIdx.z, b = blockIdx.y, c = blockIdx.x;
  const PwMaps m = mt[k0 + kz];
  const C* x = X.p + kz * X.ks + (int64_t)b * X.ld;
  const C* hx = HX.p + kz * HX.ks + (int64_t)b * HX.ld;
  const double e = eps[(int64_t)(k0 + kz) * nb + b];
  double sk = 0.0, sr = 0.0;
  for (int64_t g = c * (int64_t)NT + threadIdx.x; g < m.npw; g += (int64_t)nchunk * NT) {
    const C p = x[g], h = hx[g];
    const double px = p.x, py = p.y;
    sk += m.kin[g] * (px * px + py * py);
    const double rx = (double)h.x - e * px, ry = (double)h.y - e * py;
    sr += rx * rx + ry * ry;
  }
  const double tk = bsum(sk, sh);
  const double tr = bsum(sr, sh);
  if (threadIdx.x == 0) {
    double* p = partial + (((int64_t)kz * nb + b) * nchunk + c) * 2;
    p[0] = tk;
    p[1] = tr;
  }
}

__global__ void k_band_stats_reduce(const double* __restrict__ partial, int nitems, int nchunk, int k0, int nb,
                                    double* __restrict__ ekin, double* __restrict__ rn) {
  for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < nitems; it += gridDim.x * blockDim.x) {
    double sk = 0.0, sr = 0.0;
    for (int c = 0; c < nchunk; ++c) {
      sk += partial[((int64_t)it * nchunk + c) * 2 + 0];
      sr += partial[((int64_t)it * nchunk + c) * 2 + 1];
    }
    ekin[(int64_t)k0 * nb + it] = sk;
    rn[(int64_t)k0 * nb + it] = sqrt(sr);
  }
}

int band_stats_chunks(int64_t npw) { return (int)std::max<int64_t>(1, std::min<int64_t>(64, ceil_div(npw, 4096))); }

template <typename R>
void band_stats(BandBlock<cx_t<R>> X, BandBlock<cx_t<R>> HX, const PwMaps* mt, int k0, int nkc, int64_t max_npw,
                const double* eps, int nb, double* partial, double* ekin, double* rn, hipStream_t s) {
  const int nchunk = band_stats_chunks(max_npw);
  hipLaunchKernelGGL(k_band_stats<cx_t<R>>, dim3(nchunk, nb, nkc), dim3(NT), 0, s, X, HX, mt, k0, eps, nb, nchunk,
                     partial);
  DEREK_HIP(hipGetLastError());
  const int nitems = nkc * nb;
  hipLaunchKernelGGL(k_band_stats_reduce, dim3(blocks(nitems)), dim3(NT), 0, s, partial, nitems, nchunk, k0, nb, ekin,
                     rn);
  DEREK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ TPA correction
// W[b] = K(kin/<T>_b) (HX[b] - eps_b X[b]) where rn_b > drop_tol, else 0.
template <typename C>
__global__ void k_tpa_batch(BandBlock<C> X, BandBlock<C> HX, BandBlock<C> W, const PwMaps* __restrict__ mt, int k0,
                            const double* __restrict__ eps, const double* __restrict__ ekin,
                            const double* __restrict__ rn, int nb, double drop_tol) {
  using R = decltype(C::x);
  const int kz = blockIdx.z, b = blockIdx.y;
  const PwMaps m = mt[k0 + kz];
  const int64_t ib = (int64_t)(k0 + kz) * nb + b;
  const bool act = rn[ib] > drop_tol;
  const double e = eps[ib], ik = 1.0 / fmax(ekin[ib], 0.1);  // <T>_n floored at 0.1 Ha
  const C* x = X.p + kz * X.ks + (int64_t)b * X.ld;
  const C* hx = HX.p + kz * HX.ks + (int64_t)b * HX.ld;
  C* w = W.p + kz * W.ks + (int64_t)b * W.ld;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < m.npw; g += (int64_t)gridDim.x * blockDim.x) {
    if (!act) {
      w[g] = cmk<C>(0, 0);
      continue;
    }
    const C p = x[g], h = hx[g];
    const double t = m.kin[g] * ik;
    const double pp = 27.0 + t * (18.0 + t * (12.0 + 8.0 * t));
    const double k = pp / (pp + 16.0 * t * t * t * t);
    w[g] = cmk<C>((R)(k * ((double)h.x - e * (double)p.x)), (R)(k * ((double)h.y - e * (double)p.y)));
  }
}

template <typename R>
void tpa_batch(BandBlock<cx_t<R>> X, BandBlock<cx_t<R>> HX, BandBlock<cx_t<R>> W, const PwMaps* mt, int k0, int nkc,
               int64_t max_npw, const double* eps, const double* ekin, const double* rn, int nb, double drop_tol,
               hipStream_t s) {
  hipLaunchKernelGGL(k_tpa_batch<cx_t<R>>, dim3(blocks(max_npw, 4 * NT), nb, nkc), dim3(NT), 0, s, X, HX, W, mt, k0,
                     eps, ekin, rn, nb, drop_tol);
  DEREK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ initial wavefunctions
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// psi(g, b, k) = (n1 + i n2) / (1 + kin_g), (n1, n2) Box-Muller normals from
// the SplitMix64 stream (1<<48 | k<<24 | b) of the seed.
template <typename C>
__global__ void k_init_psi_batch(BandBlock<C> X, const PwMaps* __restrict__ mt, int k0, uint64_t seed) {
  using R = decltype(C::x);
  const int kz = blockIdx.z, b = blockIdx.y, k = k0 + kz;
  const PwMaps m = mt[k];
  const uint64_t G = 0x9E3779B97F4A7C15ull;
  const uint64_t stream = (1ull << 48) | ((uint64_t)k << 24) | (uint64_t)b;
  const uint64_t key = mix64((seed * G) ^ mix64(stream + G));
  C* x = X.p + kz * X.ks + (int64_t)b * X.ld;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < m.npw; g += (int64_t)gridDim.x * blockDim.x) {
    const uint64_t z1 = mix64(key + (uint64_t)(2 * g + 1) * G);
    const uint64_t z2 = mix64(key + (uint64_t)(2 * g + 2) * G);
    const double u1 = (double)(z1 >> 11) * (1.0 / 9007199254740992.0);
    const double u2 = (double)(z2 >> 11) * (1.0 / 9007199254740992.0);
    const double r = sqrt(-2.0 * log(1.0 - u1));
    const double t = 2.0 * 3.141592653589793238462643383279502884197 * u2;
    const double sc = 1.0 / (1.0 + m.kin[g]);
    x[g] = cmk<C>((R)(r * cos(t) * sc), (R)(r * sin(t) * sc));
  }
}

template <typename R>
void init_psi_batch(BandBlock<cx_t<R>> X, const PwMaps* mt, int k0, int nkc, int64_t max_npw, int nb, uint64_t seed,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_init_psi_batch<cx_t<R>>, dim3(blocks(max_npw, 4 * NT), nb, nkc), dim3(NT), 0, s, X, mt, k0,
                     seed);
  DEREK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ occupations
// One workgroup of 1024 threads.  kT > 0: Fermi-Dirac with mu from
// multisection (16 trial points per step, one per wavefront) on
// sum_k w_k 2 f(eps_kb) = nelec; kT = 0: the lowest nelec/2 bands filled.
// Writes occ, occw = w_k occ / volume, and report[REP_MU, REP_MTS, REP_EBAND].
constexpr int ONT = 1024;
constexpr int OW = ONT / 64;

__device__ __forceinline__ double fd_count(const double* eps, const double* wk, int nk, int nb, double mu, double kT,
                                           int lane) {
  double t = 0.0;
  const int n = nk * nb;
  for (int i = lane; i < n; i += 64) {
    const double x = fmin(200.0, fmax(-200.0, (eps[i] - mu) / kT));
    t += wk[i / nb] * 2.0 / (1.0 + exp(x));
  }
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
  return t;
}

__global__ __launch_bounds__(ONT) void k_occupations(const double* __restrict__ eps, const double* __restrict__ wk,
                                                     int nk, int nb, double nelec, double kT, double volume,
                                                     double* __restrict__ occ, double* __restrict__ occw,
                                                     double* __restrict__ rep) {
  __shared__ double sh[OW], s_lo, s_hi, s_mu;
  __shared__ int s_done;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int n = nk * nb;
  if (kT <= 0.0) {
    const int nocc = (int)llrint(nelec / 2.0);
    double eb = 0.0;
    for (int i = tid; i < n; i += ONT) {
      const int b = i % nb, k = i / nb;
      const double o = b < nocc ? 2.0 : 0.0;
      occ[i] = o;
      occw[i] = wk[k] * o / volume;
      eb += wk[k] * o * eps[i];
    }
    double mx = -1e300;
    if (nocc > 0)
      for (int k = tid; k < nk; k += ONT) mx = fmax(mx, eps[(int64_t)k * nb + nocc - 1]);
    eb = wsum(eb);
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_down(mx, o, 64));
    if (lane == 0) sh[wid] = eb;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int w = 0; w < OW; ++w) t += sh[w];
      rep[REP_EBAND] = t;
      rep[REP_MTS] = 0.0;
    }
    __syncthreads();
    if (lane == 0) sh[wid] = mx;
    __syncthreads();
    if (tid == 0) {
      double t = -1e300;
      for (int w = 0; w < OW; ++w) t = fmax(t, sh[w]);
      rep[REP_MU] = t;
    }
    return;
  }
  {
    // bracket [min eps - 50 kT, max eps + 50 kT]: strided min / max per
    // thread, then over the wavefront and the workgroup
    __shared__ double s_min[OW], s_max[OW];
    double lo = 1e300, hi = -1e300;
    for (int i = tid; i < n; i += ONT) {
      lo = fmin(lo, eps[i]);
      hi = fmax(hi, eps[i]);
    }
    for (int o = 32; o > 0; o >>= 1) {
      lo = fmin(lo, __shfl_xor(lo, o, 64));
      hi = fmax(hi, __shfl_xor(hi, o, 64));
    }
    if (lane == 0) {
      s_min[wid] = lo;
      s_max[wid] = hi;
    }
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < OW; ++w) {
        lo = fmin(lo, s_min[w]);
        hi = fmax(hi, s_max[w]);
      }
      s_lo = lo - 50 * kT;
      s_hi = hi + 50 * kT;
      s_done = 0;
    }
    __syncthreads();
  }
  // count(lo) < nelec <= count(hi); count increases with mu
  for (int it = 0; it < 64; ++it) {
    const double lo = s_lo, hi = s_hi;
    const double mu = lo + (hi - lo) * (double)(wid + 1) / (OW + 1);
    const double c = fd_count(eps, wk, nk, nb, mu, kT, lane);
    if (lane == 0) sh[wid] = (c < nelec) ? 1.0 : 0.0;
    __syncthreads();
    if (tid == 0) {
      // trial points with count < nelec form a prefix 0..p
      int p = -1;
      for (int w = 0; w < OW; ++w)
        if (sh[w] != 0.0) p = w;
      const double nlo = p >= 0 ? lo + (hi - lo) * (double)(p + 1) / (OW + 1) : lo;
      const double nhi = p + 1 < OW ? lo + (hi - lo) * (double)(p + 2) / (OW + 1) : hi;
      const bool stuck = !(nlo > lo) && !(nhi < hi);
      s_lo = fmax(lo, nlo);
      s_hi = fmin(hi, nhi);
      if (stuck || s_hi - s_lo <= 4.0 * DBL_EPSILON * fmax(fabs(s_lo), fabs(s_hi))) s_done = 1;
    }
    __syncthreads();
    if (s_done) break;
  }
  if (tid == 0) s_mu = 0.5 * (s_lo + s_hi);
  __syncthreads();
  const double mu = s_mu;
  double ent = 0.0, eb = 0.0;
  for (int i = tid; i < n; i += ONT) {
    const int k = i / nb;
    const double x = fmin(200.0, fmax(-200.0, (eps[i] - mu) / kT));
    const double f = 1.0 / (1.0 + exp(x));
    occ[i] = 2.0 * f;
    occw[i] = wk[k] * 2.0 * f / volume;
    eb += wk[k] * 2.0 * f * eps[i];
    const double fs = fmin(1.0, fmax(1e-300, f)), gs = fmin(1.0, fmax(1e-300, 1.0 - f));
    ent -= wk[k] * 2.0 * (f * log(fs) + (1.0 - f) * log(gs));
  }
  ent = wsum(ent);
  eb = wsum(eb);
  __syncthreads();
  if (lane == 0) sh[wid] = ent;
  __syncthreads();
  double te = 0.0;
  if (tid == 0)
    for (int w = 0; w < OW; ++w) te += sh[w];
  __syncthreads();
  if (lane == 0) sh[wid] = eb;
  __syncthreads();
  if (tid == 0) {
    double tb = 0.0;
    for (int w = 0; w < OW; ++w) tb += sh[w];
    rep[REP_MU] = mu;
    rep[REP_MTS] = -kT * te;
    rep[REP_EBAND] = tb;
  }
}

void occupations(const double* eps, const double* wk, int nk, int nb, double nelec, double kT, double volume,
                 double* occ, double* occw, double* rep, hipStream_t s) {
  hipLaunchKernelGGL(k_occupations, dim3(1), dim3(ONT), 0, s, eps, wk, nk, nb, nelec, kT, volume, occ, occw, rep);
  DEREK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ scalar reductions
// rep[slot] = f(scale * sum partial[0..n)), f = sqrt if root
__global__ __launch_bounds__(NT) void k_finish_sum(const double* __restrict__ partial, int n, double scale, int root,
                                                   double* __restrict__ rep, int slot) {
  __shared__ double sh[NT / 64];
  double t = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) t += partial[i];
  t = bsum(t, sh);
  if (threadIdx.x == 0) {
    t *= scale;
    rep[slot] = root ? sqrt(fmax(t, 0.0)) : t;
  }
}

void finish_sum(const double* partial, int n, double scale, bool root, double* rep, int slot, hipStream_t s) {
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(NT), 0, s, partial, n, scale, root ? 1 : 0, rep, slot);
  DEREK_HIP(hipGetLastError());
}

__global__ __launch_bounds__(NT) void k_finish_max(const double* __restrict__ x, int n, double* __restrict__ rep,
                                                   int slot) {
  __shared__ double sh[NT / 64];
  double t = 0.0;
  for (int i = threadIdx.x; i < n; i += NT) t = fmax(t, x[i]);
  for (int o = 32; o > 0; o >>= 1) t = fmax(t, __shfl_down(t, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < NT / 64; ++w) r = fmax(r, sh[w]);
    rep[slot] = r;
  }
}

void finish_max(const double* x, int n, double* rep, int slot, hipStream_t s) {
  hipLaunchKernelGGL(k_finish_max, dim3(1), dim3(NT), 0, s, x, n, rep, slot);
  DEREK_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ Pulay mixer
// History ring of H slots; st[0] = number of pushes so far.  Push: slot =
// st[0] % H takes rho_in and F = rho_out - rho_in.
__global__ void k_pulay_push(const int* __restrict__ st, int H, const double* __restrict__ rin,
                             const double* __restrict__ rout, double* __restrict__ hin, double* __restrict__ hres,
                             int64_t n) {
  const int slot = st[0] % H;
  double* hi = hin + (int64_t)slot * n;
  double* hr = hres + (int64_t)slot * n;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double a = rin[i];
    hi[i] = a;
    hr[i] = rout[i] - a;
  }
}

// partial[j*nblk + blk] = sum over the block's slice of F_slot . F_j (j < H)
__global__ __launch_bounds__(NT) void k_pulay_dots(const int* __restrict__ st, int H, const double* __restrict__ hres,
                                                   int64_t n, double* __restrict__ partial) {
  __shared__ double sh[NT / 64];
  const int slot = st[0] % H, j = blockIdx.y;
  const double* a = hres + (int64_t)slot * n;
  const double* b = hres + (int64_t)j * n;
  double t = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) t += a[i] * b[i];
  t = bsum(t, sh);
  if (threadIdx.x == 0) partial[(int64_t)j * gridDim.x + blockIdx.x] = t;
}

// A[slot][j] = A[j][slot] = sum of the partials (slots j already filled);
// then the coefficients, oldest slot first as in the reference ordering,
// from a = A^-1 1 / (1^T A^-1 1) with A scaled by its largest diagonal and
// 1e-12 added to the diagonal; st[0] += 1.
// The partial sums are added by the 64 lanes of the one wavefront (lane-strided
// blocks, then a butterfly; lane 0's total is used), the small system by lane 0.
__global__ __launch_bounds__(64) void k_pulay_solve(int* __restrict__ st, int H, const double* __restrict__ partial,
                                                   int nblk, double* __restrict__ A, double* __restrict__ coef) {
  const int lane = threadIdx.x;
  const int np = st[0];
  const int slot = np % H;
  const int cnt = np + 1 < H ? np + 1 : H;
  for (int j = 0; j < H; ++j) {
    double t = 0.0;
    for (int b = lane; b < nblk; b += 64) t += partial[(int64_t)j * nblk + b];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    const bool filled = j == slot || j < cnt;  // slots 0..cnt-1 hold data
    if (lane == 0 && filled) {
      A[slot * H + j] = t;
      A[j * H + slot] = t;
    }
  }
  if (lane != 0) return;
  int ord[16];
  for (int i = 0; i < cnt; ++i) ord[i] = (np + 1 - cnt + i) % H;
  double M[16][17];
  double c[16];
  double smax = 0.0;
  for (int i = 0; i < cnt; ++i) smax = fmax(smax, A[ord[i] * H + ord[i]]);
  for (int j = 0; j < H; ++j) coef[j] = 0.0;
  bool flat = !(smax > 0.0);
  if (!flat) {
    for (int i = 0; i < cnt; ++i) {
      for (int j = 0; j < cnt; ++j) M[i][j] = A[ord[i] * H + ord[j]] / smax + (i == j ? 1e-12 : 0.0);
      M[i][cnt] = 1.0;
    }
    for (int col = 0; col < cnt && !flat; ++col) {
      int piv = col;
      for (int r = col + 1; r < cnt; ++r)
        if (fabs(M[r][col]) > fabs(M[piv][col])) piv = r;
      if (piv != col)
        for (int j = 0; j <= cnt; ++j) {
          const double t = M[col][j];
          M[col][j] = M[piv][j];
          M[piv][j] = t;
        }
      const double pv = M[col][col];
      if (pv == 0.0) {
        flat = true;
        break;
      }
      for (int r = col + 1; r < cnt; ++r) {
        const double f = M[r][col] / pv;
        for (int j = col; j <= cnt; ++j) M[r][j] -= f * M[col][j];
      }
    }
    if (!flat) {
      for (int i = cnt - 1; i >= 0; --i) {
        double v = M[i][cnt];
        for (int j = i + 1; j < cnt; ++j) v -= M[i][j] * c[j];
        c[i] = v / M[i][i];
      }
      double sum = 0.0;
      for (int i = 0; i < cnt; ++i) sum += c[i];
      for (int i = 0; i < cnt; ++i) coef[ord[i]] = c[i] / sum;
    }
  }
  if (flat)
    for (int i = 0; i < cnt; ++i) coef[ord[i]] = 1.0 / cnt;
  st[0] = np + 1;
}

// out = sum_j coef[j] x_j  (x_j = xs + j*n, j < H; unused slots have coef 0)
__global__ void k_lincomb_dev(double* __restrict__ out, const double* __restrict__ xs, const double* __restrict__ coef,
                              int H, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int j = 0; j < H; ++j) {
      const double c = coef[j];
      if (c != 0.0) v += c * xs[(int64_t)j * n + i];
    }
    out[i] = v;
  }
}

int pulay_blocks(int64_t n) { return (int)std::min<int64_t>(256, ceil_div(n, NT)); }

void pulay_update(int* st, int H, const double* rin, const double* rout, double* hin, double* hres, double* A,
                  double* coef, double* partial, double* r_opt, double* f_opt, int64_t n, hipStream_t s) {
  DEREK_REQUIRE(H >= 1 && H <= 16, "Pulay history must be 1..16");
  hipLaunchKernelGGL(k_pulay_push, dim3(blocks(n)), dim3(NT), 0, s, st, H, rin, rout, hin, hres, n);
  DEREK_HIP(hipGetLastError());
  const int nblk = pulay_blocks(n);
  hipLaunchKernelGGL(k_pulay_dots, dim3(nblk, H), dim3(NT), 0, s, st, H, hres, n, partial);
  DEREK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_pulay_solve, dim3(1), dim3(64), 0, s, st, H, partial, nblk, A, coef);
  DEREK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_lincomb_dev, dim3(blocks(n)), dim3(NT), 0, s, r_opt, hin, coef, H, n);
  DEREK_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_lincomb_dev, dim3(blocks(n)), dim3(NT), 0, s, f_opt, hres, coef, H, n);
  DEREK_HIP(hipGetLastError());
}

#define INST(R)                                                                                                      \
  template void band_stats<R>(BandBlock<cx_t<R>>, BandBlock<cx_t<R>>, const PwMaps*, int, int, int64_t,            \
                              const double*, int, double*, double*, double*, hipStream_t);                         \
  template void tpa_batch<R>(BandBlock<cx_t<R>>, BandBlock<cx_t<R>>, BandBlock<cx_t<R>>, const PwMaps*, int, int,   \
                             int64_t, const double*, const double*, const double*, int, double, hipStream_t);       \
  template void init_psi_batch<R>(BandBlock<cx_t<R>>, const PwMaps*, int, int, int64_t, int, uint64_t, hipStream_t);
INST(double)
INST(float)
#undef INST

// ------------------------------------------------------------------ Davidson subspace blocks
// For Y = [X, R] with orthonormal Ritz vectors X: S(X,X) = I and
// H(X,X) = diag(eps) (per matrix kz of the batch, m2 x m2, ld m2), and the
// lower-left blocks S(R,X), H(R,X) mirrored from the computed upper-right
// ones.
__global__ void k_subspace_blocks(double2* __restrict__ S, double2* __restrict__ H, const double* __restrict__ eps,
                                  int m2, int nb) {
  const int kz = blockIdx.y;
  double2* s = S + (int64_t)kz * m2 * m2;
  double2* h = H + (int64_t)kz * m2 * m2;
  const double* e = eps + (int64_t)kz * nb;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nb * nb; q += gridDim.x * blockDim.x) {
    const int i = q % nb, j = q / nb;
    s[i + (int64_t)j * m2] = make_double2(i == j ? 1.0 : 0.0, 0.0);
    h[i + (int64_t)j * m2] = make_double2(i == j ? e[i] : 0.0, 0.0);
    // (nb + i, j) = conj((j, nb + i))
    const double2 a = s[j + (int64_t)(nb + i) * m2], b = h[j + (int64_t)(nb + i) * m2];
    s[nb + i + (int64_t)j * m2] = make_double2(a.x, -a.y);
    h[nb + i + (int64_t)j * m2] = make_double2(b.x, -b.y);
  }
}

void subspace_blocks(double2* S, double2* H, const double* eps, int m2, int nb, int nkc, hipStream_t s) {
  DEREK_REQUIRE(m2 == 2 * nb, "subspace blocks: m2 must be 2 nb");
  const int nblk = (int)std::min<int64_t>(64, ceil_div((int64_t)nb * nb, NT));
  hipLaunchKernelGGL(k_subspace_blocks, dim3(nblk, nkc), dim3(NT), 0, s, S, H, eps, m2, nb);
  DEREK_HIP(hipGetLastError());
}

}  // namespace derek
