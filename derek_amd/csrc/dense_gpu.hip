// FP64 dense subspace solvers on the GPU, one 1024-thread workgroup per
// matrix (blockIdx.x selects the matrix of a batch).
//
// k_heev: lowest nev eigenpairs of a Hermitian n x n matrix.
//   1. Householder reduction to real tridiagonal form: while the trailing
//      matrix exceeds LDS, one multi-workgroup launch per column on the matrix
//      in global memory (L2 resident); then the last <= 98 columns in LDS in
//      one workgroup (lower triangle, look-ahead reflector, two barriers per
//      column).  The reflectors overwrite the eliminated columns.
//   2. Phases that make the sub-diagonal real and non-negative (a prefix
//      product over the workgroup).
//   3. Sturm-sequence multisection for the lowest nev eigenvalues, one
//      wavefront each (64 points per step).
//   4. Inverse iteration for the eigenvectors of the tridiagonal matrix with
//      partial-pivoting LU; eigenvalues closer than 1e-5 |T| form a cluster
//      whose vectors are computed by one thread in order and re-orthogonalised
//      (modified Gram-Schmidt) against the earlier members.
//   5. Back-transformation by the reflectors, one wavefront per vector.
// k_chol_drop: Cholesky of a unit-diagonal Hermitian overlap that skips every
//   column whose remaining pivot is below tol (dependent on earlier columns).
// k_trinv: inverse of a lower-triangular matrix, one wavefront per column.
#include <cfloat>

#include "dense_gpu.hpp"
#include "launch.hpp"

namespace derek {

constexpr int DT = 1024;  // threads per workgroup
constexpr int DW = DT / 64;
// LDS of the matrix-resident tridiagonalisation (gfx950: 160 KiB per workgroup)
constexpr size_t kTridiagLds = 156 * 1024;

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 zconjmul(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// One DPP move of a double (both 32-bit halves) with the given control and
// row mask; lanes without a source (or outside the mask) read 0.
template <int CTRL, int ROWS = 0xf> __device__ __forceinline__ double dpp_move(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffull), CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over the 64 lanes of a wave (all lanes active), every lane gets it:
// quad butterflies, row shifts by 4 and 8 (lane 15 of each row holds the row
// sum), row broadcasts 15 and 31 (lane 63 holds the total), then a readlane.
// Cross-lane moves on the VALU instead of LDS permutes.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_move<0xb1>(x);        // quad_perm [1,0,3,2]
  x += dpp_move<0x4e>(x);        // quad_perm [2,3,0,1]
  x += dpp_move<0x114>(x);       // row_shr:4
  x += dpp_move<0x118>(x);       // row_shr:8
  x += dpp_move<0x142, 0xa>(x);  // row_bcast:15 into rows 1 and 3
  x += dpp_move<0x143, 0xc>(x);  // row_bcast:31 into rows 2 and 3
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), 63);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), 63);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum over the workgroup; every thread gets the result.
__device__ double block_sum_all(double x, double* red) {
  x = wave_sum(x);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = x;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < DW; ++w) t += red[w];
  return t;
}

__device__ double block_min_all(double x, double* red) {
  for (int o = 32; o > 0; o >>= 1) x = fmin(x, __shfl_down(x, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = x;
  __syncthreads();
  double t = red[0];
  for (int w = 1; w < DW; ++w) t = fmin(t, red[w]);
  return t;
}

__device__ double block_max_all(double x, double* red) {
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_down(x, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = x;
  __syncthreads();
  double t = red[0];
  for (int w = 1; w < DW; ++w) t = fmax(t, red[w]);
  return t;
}

// 1/q for q != 0: the hardware estimate refined by two Newton steps, fewer
// issue cycles than the division sequence (the Sturm recurrence below is
// bound by its wavefront's issue rate).  q is clamped to the finite range
// first, so an infinite q gives 0 as the division would.
__device__ __forceinline__ double recip_nr(double q) {
  const double qc = fmin(fmax(q, -DBL_MAX), DBL_MAX);
  double y = __builtin_amdgcn_rcp(qc);
  double e = fma(-qc, y, 1.0);
  y = fma(y, e, y);
  e = fma(-qc, y, 1.0);
  return fma(y, e, y);
}

// number of eigenvalues of the tridiagonal (d, e) smaller than x
__device__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  int c = 0;
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  if (q < 0.0) ++c;
  for (int i = 1; i < n; ++i) {
    q = d[i] - x - e2[i - 1] * recip_nr(q);
    if (fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++c;
  }
  return c;
}

// Solve (T - lambda I) x = b by LU with partial pivoting; T has diagonal d and
// off-diagonal e.  u0/u1/u2/l/pv are scratch of length n; b is overwritten by x.
// LU factorisation with partial pivoting of T - lambda I (T symmetric
// tridiagonal: diagonal d, off-diagonal e), pivots below pivmin raised to it;
// U has the three diagonals (u0 holds the reciprocals of the pivots, so the
// solves multiply), L the multipliers l with row swaps pv.  The arrays are
// separate LDS regions (restrict), so the loads of the next row are issued
// while the recurrence runs in registers.
__device__ void tri_factor(const double* __restrict__ d, const double* __restrict__ e, int n, double lambda,
                           double pivmin, double* __restrict__ u0, double* __restrict__ u1, double* __restrict__ u2,
                           double* __restrict__ l, double* __restrict__ pv) {
  double cd = d[0] - lambda, cs = n > 1 ? e[0] : 0.0, cs2 = 0.0;
  for (int i = 0; i + 1 < n; ++i) {
    const double c = e[i];
    const double a1 = d[i + 1] - lambda;
    const double b1 = i + 2 < n ? e[i + 1] : 0.0;
    if (fabs(cd) >= fabs(c)) {
      const double piv = fabs(cd) < pivmin ? copysign(pivmin, cd == 0.0 ? 1.0 : cd) : cd;
      const double rp = recip_nr(piv);
      const double li = c * rp;
      u0[i] = rp;
      u1[i] = cs;
      u2[i] = cs2;
      l[i] = li;
      pv[i] = 0.0;
      cd = a1 - li * cs;
      cs = b1 - li * cs2;
      cs2 = 0.0;
    } else {
      const double rc = recip_nr(c);
      const double li = cd * rc;
      u0[i] = rc;
      u1[i] = a1;
      u2[i] = b1;
      l[i] = li;
      pv[i] = 1.0;
      cd = cs - li * a1;
      cs = cs2 - li * b1;
      cs2 = 0.0;
    }
  }
  u0[n - 1] = recip_nr(fabs(cd) < pivmin ? copysign(pivmin, cd == 0.0 ? 1.0 : cd) : cd);
}

// b <- (T - lambda I)^-1 b from tri_factor's factors; the values carried from
// row to row stay in registers and the next row's operands are loaded one
// step ahead, so each step waits only on its own arithmetic.
__device__ void tri_solve(int n, const double* __restrict__ u0, const double* __restrict__ u1,
                          const double* __restrict__ u2, const double* __restrict__ l, const double* __restrict__ pv,
                          double* __restrict__ b) {
  double bi = b[0];
  if (n > 1) {
    double nb = b[1], nl = l[0], np = pv[0];
    for (int i = 0; i + 1 < n; ++i) {
      double bn = nb;
      const double li = nl, pi = np;
      if (i + 2 < n) {
        nb = b[i + 2];
        nl = l[i + 1];
        np = pv[i + 1];
      }
      if (pi != 0.0) {
        const double t = bi;
        bi = bn;
        bn = t;
      }
      bn -= li * bi;
      b[i] = bi;
      bi = bn;
    }
  }
  double b1 = bi * u0[n - 1];  // b[i+1]
  b[n - 1] = b1;
  if (n < 2) return;
  double b0 = (b[n - 2] - u1[n - 2] * b1) * u0[n - 2];  // b[i+1] after the step below
  b[n - 2] = b0;
  if (n < 3) return;
  double cb = b[n - 3], c1 = u1[n - 3], c2 = u2[n - 3], cr = u0[n - 3];
  for (int i = n - 3; i >= 0; --i) {
    const double xb = cb, x1 = c1, x2 = c2, xr = cr;
    if (i > 0) {
      cb = b[i - 1];
      c1 = u1[i - 1];
      c2 = u2[i - 1];
      cr = u0[i - 1];
    }
    const double bi2 = (xb - x1 * b0 - x2 * b1) * xr;
    b[i] = bi2;
    b1 = b0;
    b0 = bi2;
  }
}

// Per-matrix workspace layout (doubles), see heev_work_doubles.
struct HeevWork {
  double2* sub;  // [n]  T(k+1,k) before the phase change
  double2* ph;   // [n]  phases
  double2* vb;   // [2n] reflector vectors of the wide reduction, double-buffered by k&1
  double2* uu;   // [2n] A v of the wide reduction, double-buffered by k&1
  double* d;     // [n]  diagonal
  double* e;     // [n]  |sub|
  double* lam;   // [nev]
  double* scal;  // [8]  wide reduction: (skip, 1/|v|, v0) per buffer
  double* xt;    // [n * nev] tridiagonal eigenvectors
  int* skip;     // [n]  1 if column k needed no reflector
  int* cstart;   // [nev + 1]
  int* ncl;      // [1]
  __device__ __host__ HeevWork(double* w, int n, int nev) {
    sub = reinterpret_cast<double2*>(w);
    ph = sub + n;
    vb = ph + n;
    uu = vb + 2 * n;
    d = reinterpret_cast<double*>(uu + 2 * n);
    e = d + n;
    lam = e + n;
    scal = lam + nev;
    xt = scal + 8;
    skip = reinterpret_cast<int*>(xt + (size_t)n * nev);
    cstart = skip + n;
    ncl = cstart + nev + 1;
  }
};

int64_t heev_work_doubles(int n, int nev) {
  // rounded to 256 bytes so every matrix of a batch starts aligned
  const int64_t w = 14 * (int64_t)n + nev + 8 + (int64_t)n * nev + (n + nev + 2 + 1) / 2 + 8;
  return (w + 31) / 32 * 32;
}

// ---- 1. Householder tridiagonalisation + phases
// Columns 0 .. n-99 (n > 98) take one multi-workgroup launch each
// (k_wide_step: the rank-2 update of column k's reflector, fused with the
// next reflector and u = A v for it); the trailing block of at most 98
// columns is copied to LDS (lower triangle) and reduced by one workgroup
// per matrix (k_tridiag).

// e = |sub| and the phases ph[0] = 1, ph[k+1] = ph[k] f_k, f_k = sub_k/|sub_k|
// (1 where sub_k = 0), that make D^H T D real.  The phases are a prefix
// product of unit complex numbers: each thread takes two consecutive f_k
// (n <= 2048 = 2 DT), the threads' products are scanned across the lanes of
// each wavefront and then across the wavefronts (s_wave: DW entries).
__device__ void tri_phases(HeevWork& W, int n, double2* s_wave) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const double2 one = make_double2(1.0, 0.0);
  double2 f[2], q = one;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int k = 2 * tid + r;
    f[r] = one;
    if (k + 1 < n) {
      const double2 sk = W.sub[k];
      const double a = sqrt(sk.x * sk.x + sk.y * sk.y);
      W.e[k] = a;
      if (a > 0.0) f[r] = make_double2(sk.x / a, sk.y / a);
    }
    q = zmul(q, f[r]);
  }
  // inclusive scan of the thread products over the wavefront
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double2 t = make_double2(__shfl_up(q.x, o, 64), __shfl_up(q.y, o, 64));
    if (lane >= o) q = zmul(t, q);
  }
  if (lane == 63) s_wave[wid] = q;
  __syncthreads();
  double2 p = one;
  for (int w = 0; w < wid; ++w) p = zmul(p, s_wave[w]);
  double2 ex = make_double2(__shfl_up(q.x, 1, 64), __shfl_up(q.y, 1, 64));
  if (lane == 0) ex = one;
  p = zmul(p, ex);  // ph[2 tid]
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int k = 2 * tid + r;
    p = zmul(p, f[r]);
    if (k + 1 < n) W.ph[k + 1] = p;
  }
  if (tid == 0) {
    W.ph[0] = one;
    W.e[n - 1] = 0.0;
  }
}


// LDS pitch (in 16-byte elements) of the trailing block in k_tridiag: 2 mod 4,
// so that the lanes of a column-wise read (eight columns of one row, stride =
// pitch) and of the update's mirrored rows (stride = pitch + 1) fall on
// distinct banks
__host__ __device__ constexpr int tri_pitch(int nb) { return nb + (6 - nb % 4) % 4; }

// LDS bytes of k_tridiag for a trailing block of nb columns: the block, u and
// the per-column reflector scalars (at least the DW entries of the phase scan)
static size_t tri_lds(int nb) {
  return std::max(sizeof(double2) * ((size_t)tri_pitch(nb) * nb + 2 * (size_t)nb) + 8 * (size_t)nb,
                  sizeof(double2) * DW);
}

// 1/sqrt(x) for x > 0 within a few ulps: the hardware estimate refined by two
// Newton steps, a shorter dependent chain than a square root and a division
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-x * y, y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-x * y, y, 1.0);
  return fma(0.5 * y, e, y);
}

__device__ __forceinline__ double read_lane(double x, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Householder scalars of column k from xn2 = sum_{i>k} |A(i,k)|^2 and
// x0 = A(k+1,k) (both wave-uniform): alpha = -(x0/|x0|) ||x|| (the new
// sub-diagonal), v0 = x0 - alpha and iv = 1/||v|| with
// ||v||^2 = 2 ||x|| (||x|| + |x0|); skip when x is already reduced.  When rec
// is set the caller's lane records alpha and the skip flag.
__device__ __forceinline__ void tri_scalars(bool rec, int k, double xn2, double2 x0, HeevWork& W, int& skip,
                                            double& iv, double2& v0) {
  const double a2 = x0.x * x0.x + x0.y * x0.y;
  skip = (xn2 == 0.0 || xn2 - a2 <= 1e-300) ? 1 : 0;
  iv = 0.0;
  v0 = make_double2(0.0, 0.0);
  double2 alpha = x0;
  if (!skip) {
    const double xn = xn2 * rsqrt_nr(xn2);
    double ax0 = 0.0;
    double2 phs = make_double2(1.0, 0.0);
    if (a2 > 0.0) {
      const double ra = rsqrt_nr(a2);
      ax0 = a2 * ra;
      phs = make_double2(x0.x * ra, x0.y * ra);
    }
    alpha = make_double2(-phs.x * xn, -phs.y * xn);
    v0 = make_double2(x0.x + phs.x * xn, x0.y + phs.y * xn);
    iv = rsqrt_nr(2.0 * xn * (xn + ax0));
  }
  if (rec) {
    W.skip[k] = skip;
    W.sub[k] = alpha;
  }
}

// Reflector of column k from sum |A(i,k)|^2 (i > k) and x0 = A(k+1,k), both
// wave-uniform, for the multi-workgroup steps: every lane derives the
// scalars, lane 0 records them (with the scalars for the next launch).
__device__ __forceinline__ void wide_reflector_scalars(int lane, int k, double xn2, double2 x0, HeevWork& W,
                                                       int& skip, double& iv, double2& v0) {
  tri_scalars(lane == 0, k, xn2, x0, W, skip, iv, v0);
  if (lane == 0) {
    double* sc = W.scal + 4 * (k & 1);
    sc[0] = skip;
    sc[1] = iv;
    sc[2] = v0.x;
    sc[3] = v0.y;
  }
}

// Reflector of column c of the LDS block from its values below the diagonal
// (one wavefront); lane 0 records it and stores its scalars in sc.
__device__ __forceinline__ void tri_col_reflector(const double2* __restrict__ As, int lda, int nb, int k0, int c,
                                                  int lane, HeevWork& W, double* sc) {
  const double2* col = As + c * lda;
  double part = 0.0;
  for (int i = c + 1 + lane; i < nb; i += 64) part += col[i].x * col[i].x + col[i].y * col[i].y;
  part = wave_sum(part);
  int skip;
  double iv;
  double2 v0;
  tri_scalars(lane == 0, k0 + c, part, col[c + 1], W, skip, iv, v0);
  if (lane == 0) {
    sc[0] = skip;
    sc[1] = iv;
    sc[2] = v0.x;
    sc[3] = v0.y;
  }
}

// Householder reduction of the trailing block A[k0:, k0:] (nb = n - k0 rows
// and columns) in LDS, one workgroup per matrix: the whole matrix for k0 = 0,
// or the tail after the multi-workgroup steps 0..k0-1.  Only the lower
// triangle is kept; indices below are local to the block.  Column c keeps x
// while it is reduced; v = x iv (x_{c+1} replaced by v0) is formed where it
// is read.  Per column c, with its reflector's scalars already in LDS:
//   1. u = A v with A(i,j) = conj(A(j,i)) above the diagonal: each wavefront
//      takes eight rows (lane >> 3) times eight column classes (lane & 7,
//      columns c+1+cl, c+9+cl, ...), summed within the eight lanes by DPP;
//      the wavefronts' parts of v^H u go to LDS;
//   2. A -= v w^H + w v^H (w = 2u - 2 gamma v): wavefront 0 updates column
//      c+1 and forms its reflector (look-ahead) while the other wavefronts
//      update the lower triangle of rows and columns c+2.., each thread a row
//      a and its mirror row m-1-a so that the threads share the work evenly.
// Two barriers per column.  The reflectors are scaled to unit norm when the
// block is written back.
__global__ __launch_bounds__(DT) void k_tridiag(double2* __restrict__ Aall, int lda_g, int64_t strideA, int n,
                                               int k0, int nev, double* __restrict__ work, int64_t strideWork) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int nb = n - k0, lda = tri_pitch(nb);
  double2* As = reinterpret_cast<double2*>(smem);     // [lda * nb] A(i, j) = As[i + j lda], i >= j
  double2* us = As + (size_t)lda * nb;                // [nb] u = A v
  double2* rv0 = us + nb;                             // [nb] v0 iv of column c
  double* riv = reinterpret_cast<double*>(rv0 + nb);  // [nb] iv of column c (0 without a reflector)
  __shared__ double s_sc[2][4];                       // skip, iv, v0 of columns c (slot c & 1) and c + 1
  __shared__ double s_g[DW];                          // the wavefronts' parts of v^H u
  const int mat = blockIdx.x;
  double2* Ag = Aall + mat * strideA + k0 + (int64_t)k0 * lda_g;
  HeevWork W(work + mat * strideWork, n, nev);
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int nn = nb * nb;
  for (int q0 = tid; q0 < nn; q0 += 4 * DT) {
    double2 t[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = q0 + r * DT, i = q % nb, j = q / nb;
      t[r] = make_double2(0.0, 0.0);
      if (q < nn && i >= j) t[r] = Ag[i + (int64_t)j * lda_g];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = q0 + r * DT, i = q % nb, j = q / nb;
      if (q < nn && i >= j) As[i + j * lda] = t[r];
    }
  }
  __syncthreads();
  if (nb > 2 && wid == 0) tri_col_reflector(As, lda, nb, k0, 0, lane, W, s_sc[0]);
  __syncthreads();
  for (int c = 0; c + 2 < nb; ++c) {
    const int m = nb - c - 1;       // trailing rows and columns
    const bool more = c + 3 < nb;  // column c+1 takes a reflector too
    const double* sc = s_sc[c & 1];
    const bool skip = sc[0] != 0.0;
    const double iv = sc[1];
    const double2 v0 = make_double2(sc[2], sc[3]);
    if (tid == 0) {
      rv0[c] = make_double2(v0.x * iv, v0.y * iv);
      riv[c] = iv;
    }
    if (skip) {  // nothing changes; column c+1 is read as it stands
      if (more && wid == 0) tri_col_reflector(As, lda, nb, k0, c + 1, lane, W, s_sc[(c + 1) & 1]);
      __syncthreads();
      continue;
    }
    const double2* col = As + c * lda;
    auto vat = [=](int j) {
      const double2 x = j == c + 1 ? v0 : col[j];
      return make_double2(x.x * iv, x.y * iv);
    };
    // 1. u = A v and the wavefront's part of v^H u
    double gp = 0.0;
    for (int i0 = c + 1 + wid * 8; i0 < nb; i0 += DT / 8) {
      const int i = i0 + (lane >> 3);
      double sr = 0.0, si = 0.0, tr = 0.0, ti = 0.0;
      if (i < nb) {
        for (int j0 = c + 1 + (lane & 7); j0 < nb; j0 += 64) {
          double2 a[8], b[8];
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int j = j0 + 8 * q;
            a[q] = make_double2(0.0, 0.0);
            b[q] = make_double2(0.0, 0.0);
            if (j < nb) {
              const double2 e = As[j <= i ? i + j * lda : j + i * lda];
              a[q] = j <= i ? e : make_double2(e.x, -e.y);
              b[q] = vat(j);
            }
          }
#pragma unroll
          for (int q = 0; q < 8; q += 2) {
            sr += a[q].x * b[q].x - a[q].y * b[q].y;
            si += a[q].x * b[q].y + a[q].y * b[q].x;
            tr += a[q + 1].x * b[q + 1].x - a[q + 1].y * b[q + 1].y;
            ti += a[q + 1].x * b[q + 1].y + a[q + 1].y * b[q + 1].x;
          }
        }
      }
      sr += tr;
      si += ti;
      sr += dpp_move<0xb1>(sr);  // quad_perm [1,0,3,2]
      si += dpp_move<0xb1>(si);
      sr += dpp_move<0x4e>(sr);  // quad_perm [2,3,0,1]
      si += dpp_move<0x4e>(si);
      sr += dpp_move<0x114>(sr);  // row_shr:4: lanes 4..7 of each eight hold the row's sum
      si += dpp_move<0x114>(si);
      if ((lane & 7) == 4 && i < nb) {
        us[i] = make_double2(sr, si);
        const double2 vi = vat(i);
        gp += vi.x * sr + vi.y * si;  // Re(conj(v_i) u_i)
      }
    }
    gp = wave_sum(gp);
    if (lane == 0) s_g[wid] = gp;
    __syncthreads();
    double gam = 0.0;
#pragma unroll
    for (int w = 0; w < DW; ++w) gam += s_g[w];
    const double g2 = 2.0 * gam;
    // 2. the rank-2 update
    if (wid == 0) {
      const int j = c + 1;
      const double2 vj = vat(j), uj = us[j];
      const double2 wj = make_double2(2.0 * uj.x - g2 * vj.x, 2.0 * uj.y - g2 * vj.y);
      double part = 0.0;
      double2 x0 = make_double2(0.0, 0.0);
      for (int i = j + lane; i < nb; i += 64) {
        const double2 vi = vat(i), ui = us[i];
        const double2 wi = make_double2(2.0 * ui.x - g2 * vi.x, 2.0 * ui.y - g2 * vi.y);
        const double2 t1 = zmul(vi, make_double2(wj.x, -wj.y));
        const double2 t2 = zmul(wi, make_double2(vj.x, -vj.y));
        double2& r = As[i + j * lda];
        r = make_double2(r.x - t1.x - t2.x, r.y - t1.y - t2.y);
        if (i > j) part += r.x * r.x + r.y * r.y;
        if (i == j + 1) x0 = r;
      }
      if (more) {  // row c+2 of column c+1 lies in lane 1
        part = wave_sum(part);
        x0 = make_double2(read_lane(x0.x, 1), read_lane(x0.y, 1));
        int skip1;
        double iv1;
        double2 v01;
        tri_scalars(lane == 0, k0 + c + 1, part, x0, W, skip1, iv1, v01);
        if (lane == 0) {
          double* s = s_sc[(c + 1) & 1];
          s[0] = skip1;
          s[1] = iv1;
          s[2] = v01.x;
          s[3] = v01.y;
        }
      }
    } else {
      const int m2 = m - 1;  // rows and columns c+2 .. nb-1
      const int P = (m2 + 1) / 2, cg = (DT - 64) / P;
      const int t = tid - 64, p = t % P, g = t / P;
      if (g < cg) {
        const int ra = c + 2 + p, rb = c + 1 + m2 - p;  // rows a and m2-1-a of the block
        const double2 va = vat(ra), ua = us[ra], vb = vat(rb), ub = us[rb];
        const double2 wa = make_double2(2.0 * ua.x - g2 * va.x, 2.0 * ua.y - g2 * va.y);
        const double2 wb = make_double2(2.0 * ub.x - g2 * vb.x, 2.0 * ub.y - g2 * vb.y);
        const int ne = ra == rb ? p + 1 : m2 + 1;  // row a: p + 1 elements, row m2-1-a: m2 - p
        for (int e0 = g; e0 < ne; e0 += 4 * cg) {
          int idx[4];
          double2 r[4], vj[4], uj[4];
          bool fs[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int e = e0 + q * cg;
            fs[q] = e <= p;
            const int i = fs[q] ? ra : rb, j = c + 2 + (fs[q] ? e : e - p - 1);
            idx[q] = e < ne ? i + j * lda : -1;
            r[q] = vj[q] = uj[q] = make_double2(0.0, 0.0);
            if (e < ne) {
              r[q] = As[idx[q]];
              vj[q] = vat(j);
              uj[q] = us[j];
            }
          }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (idx[q] < 0) continue;
            const double2 vi = fs[q] ? va : vb, wi = fs[q] ? wa : wb;
            const double2 wj = make_double2(2.0 * uj[q].x - g2 * vj[q].x, 2.0 * uj[q].y - g2 * vj[q].y);
            const double2 t1 = zmul(vi, make_double2(wj.x, -wj.y));
            const double2 t2 = zmul(wi, make_double2(vj[q].x, -vj[q].y));
            As[idx[q]] = make_double2(r[q].x - t1.x - t2.x, r[q].y - t1.y - t2.y);
          }
        }
      }
    }
    __syncthreads();
  }
  __syncthreads();
  // the reflectors (strict lower triangle, unit norm) for the back-transformation
  for (int q0 = tid; q0 < nn; q0 += 4 * DT) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = q0 + r * DT, i = q % nb, j = q / nb;
      if (q < nn && i > j && j + 2 < nb) {
        const double2 x = As[i + j * lda];
        Ag[i + (int64_t)j * lda_g] = i == j + 1 ? rv0[j] : make_double2(x.x * riv[j], x.y * riv[j]);
      }
    }
  }
  // the diagonal (rows before k0 were final after their own step) and the
  // last sub-diagonal element
  for (int i = tid; i < n; i += DT)
    W.d[i] = i < k0 ? Aall[mat * strideA + i + (int64_t)i * lda_g].x : As[(i - k0) * (lda + 1)].x;
  if (tid == 0) {
    if (nb >= 2) {
      W.sub[n - 2] = As[(nb - 1) + (nb - 2) * lda];
      W.skip[n - 2] = 1;
    }
    W.skip[n - 1] = 1;
  }
  __syncthreads();
  // phases: T' = D^H T D real with off-diagonal |sub| (As is free now)
  tri_phases(W, n, As);
}

// ---- 1'. wide tridiagonalisation (matrices that do not fit LDS): one
// launch per column, the trailing matrix spread over many workgroups (one
// wavefront per column).  k_wide_first and k_wide_gemv form reflector 0 and
// u_0 = A22 v_0; launch k (k_wide_step) applies the rank-2 update of step k
// and forms reflector k+1 and u_{k+1}.  Reflector k is kept in vb (buffer
// k&1, scaled to unit norm) with its scalars in scal; A's column k receives
// it for the back-transformation.
constexpr int WT = 256;       // threads per workgroup
constexpr int WC = WT / 64;   // columns per workgroup


// column k's reflector from its current values (one wavefront)
__device__ __forceinline__ void wide_reflector(const double2* __restrict__ A, int lda, int k, int n, HeevWork& W,
                                               int lane) {
  const double2* col = A + (int64_t)k * lda;
  double part = 0.0;
  for (int i = k + 1 + lane; i < n; i += 64) part += col[i].x * col[i].x + col[i].y * col[i].y;
  part = wave_sum(part);
  const double2 x0 = col[k + 1];
  int skip;
  double iv;
  double2 v0;
  wide_reflector_scalars(lane, k, part, x0, W, skip, iv, v0);
  double2* v = W.vb + (size_t)(k & 1) * n;
  for (int i = k + 1 + lane; i < n; i += 64) {
    const double2 a = i == k + 1 ? v0 : col[i];
    v[i] = make_double2(a.x * iv, a.y * iv);
  }
}

__global__ __launch_bounds__(64) void k_wide_first(double2* __restrict__ Aall, int lda, int64_t strideA, int n,
                                                   int nev, double* __restrict__ work, int64_t strideWork) {
  const int mat = blockIdx.x;
  HeevWork W(work + mat * strideWork, n, nev);
  wide_reflector(Aall + mat * strideA, lda, 0, n, W, threadIdx.x);
}

// u_i = sum_j conj(A(j,i)) v_j for the trailing columns i > k; the first
// workgroup also stores the reflector into column k.
__global__ __launch_bounds__(WT) void k_wide_gemv(double2* __restrict__ Aall, int lda, int64_t strideA, int n,
                                                  int nev, int k, double* __restrict__ work, int64_t strideWork) {
  const int mat = blockIdx.y;
  HeevWork W(work + mat * strideWork, n, nev);
  if (W.scal[4 * (k & 1)] != 0.0) return;  // no reflector for this column
  double2* A = Aall + mat * strideA;
  const double2* v = W.vb + (size_t)(k & 1) * n;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (blockIdx.x == 0 && wid == 0)
    for (int i = k + 1 + lane; i < n; i += 64) A[i + (int64_t)k * lda] = v[i];
  const int i = k + 1 + blockIdx.x * WC + wid;
  if (i >= n) return;
  const double2* col = A + (int64_t)i * lda;
  double sr = 0.0, si = 0.0;
  for (int j = k + 1 + lane; j < n; j += 64) {
    const double2 t = zconjmul(col[j], v[j]);
    sr += t.x;
    si += t.y;
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  if (lane == 0) W.uu[i] = make_double2(sr, si);
}

// Step k of the wide reduction in one launch.  Reflector k (vb[k&1], scal)
// and u_k = A22 v_k (uu[k&1]) come from the previous launch; with
// w = 2u - 2 (v^H u) v every wavefront applies A22 -= v w^H + w v^H to its
// column j.  Every workgroup also updates column k+1 on the fly and forms
// reflector k+1 from it in LDS (the column itself is not written back: its
// rows below k+1 become that reflector, and the other workgroups read its
// old values); the wavefront of each column j > k+1 then forms
// u_{k+1}(j) = (updated column j)^H v_{k+1} into uu[(k+1)&1], so the next
// step needs no separate product launch.  Workgroup 0 records
// reflector k+1 (vb, scal, skip, sub), the updated diagonal entry of column
// k+1, and reflector k into column k for the back-transformation.  On the
// last wide step (`last`) column k+1 is updated in place like the others,
// for the in-LDS reduction that follows.
__global__ __launch_bounds__(WT) void k_wide_step(double2* __restrict__ Aall, int lda, int64_t strideA, int n,
                                                  int nev, int k, int last, double* __restrict__ work,
                                                  int64_t strideWork) {
  extern __shared__ __align__(16) unsigned char smem[];
  double2* vn = reinterpret_cast<double2*>(smem);  // [n] reflector k+1 (rows > k+1)
  const int mat = blockIdx.y;
  HeevWork W(work + mat * strideWork, n, nev);
  double2* A = Aall + mat * strideA;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // the wavefront's column j and the thread's rows of column k+1 are loaded
  // before anything is stored (the stores into A below would otherwise keep
  // these loads, a second round trip, behind them)
  constexpr int PF = 4;  // rows per lane of column j held from entry (n - k - 1 <= 256)
  const int j = k + 1 + blockIdx.x * WC + wid;
  const bool colj = j < n && (last || j != k + 1);
  double2 pre[PF];
#pragma unroll
  for (int q = 0; q < PF; ++q) {
    const int i = k + 1 + lane + 64 * q;
    pre[q] = colj && i < n ? A[i + (int64_t)j * lda] : make_double2(0.0, 0.0);
  }
  const int i1 = k + 2 + (int)threadIdx.x;  // this thread's first row of column k+1
  const double2 c1pre = !last && i1 < n ? A[i1 + (int64_t)(k + 1) * lda] : make_double2(0.0, 0.0);
  const bool skip = W.scal[4 * (k & 1)] != 0.0;  // column k needed no reflector
  const double2* v = W.vb + (size_t)(k & 1) * n;
  const double2* u = W.uu + (size_t)(k & 1) * n;
  // v and u are read whether or not column k took a reflector (the values are
  // then unused), so these loads need not wait for the flag
  double gp = 0.0;
  for (int i = k + 1 + lane; i < n; i += 64) gp += zconjmul(v[i], u[i]).x;
  const double g2 = skip ? 0.0 : 2.0 * wave_sum(gp);
  auto wof = [=](int i) {
    const double2 vi = v[i], ui = u[i];
    return make_double2(2.0 * ui.x - g2 * vi.x, 2.0 * ui.y - g2 * vi.y);
  };
  // a - v_i conj(w_j) - w_i conj(v_j)
  auto upd = [](double2 a, double2 vi, double2 wi, double2 vj, double2 wj) {
    const double2 t1 = zmul(vi, make_double2(wj.x, -wj.y));
    const double2 t2 = zmul(wi, make_double2(vj.x, -vj.y));
    return make_double2(a.x - t1.x - t2.x, a.y - t1.y - t2.y);
  };
  if (blockIdx.x == 0 && wid == 0 && !skip)
    for (int i = k + 1 + lane; i < n; i += 64) A[i + (int64_t)k * lda] = v[i];
  // reflector k+1 from column k+1 after step k (rows > k+1), formed by the
  // whole workgroup: the unscaled values in vn, the norm from per-wavefront
  // partial sums; the scale and the first element enter the products below
  __shared__ double red[WC];
  int skip1 = 1;
  double iv1 = 0.0;
  double2 v01 = make_double2(0.0, 0.0);
  if (!last) {
    const double2* c1 = A + (int64_t)(k + 1) * lda;
    const double2 v1 = skip ? make_double2(0.0, 0.0) : v[k + 1];
    const double2 w1 = skip ? make_double2(0.0, 0.0) : wof(k + 1);
    double part = 0.0;
    for (int i = i1; i < n; i += WT) {  // thread 0 holds row k+2
      const double2 ci = i == i1 ? c1pre : c1[i];
      const double2 c = skip ? ci : upd(ci, v[i], wof(i), v1, w1);
      vn[i] = c;
      part += c.x * c.x + c.y * c.y;
    }
    part = wave_sum(part);
    if (lane == 0) red[wid] = part;
    __syncthreads();
    part = 0.0;
    for (int q = 0; q < WC; ++q) part += red[q];
    // every thread derives the scalars; thread 0 of workgroup 0 records them
    wide_reflector_scalars(blockIdx.x == 0 && threadIdx.x == 0 ? 0 : 1, k + 1, part, vn[k + 2], W, skip1, iv1, v01);
    if (blockIdx.x == 0) {
      double2* vb1 = W.vb + (size_t)((k + 1) & 1) * n;
      for (int i = k + 2 + (int)threadIdx.x; i < n; i += WT) {
        const double2 c = i == k + 2 ? v01 : vn[i];
        vb1[i] = make_double2(c.x * iv1, c.y * iv1);
      }
      if (threadIdx.x == 0 && !skip) {
        double2* dkk = A + (k + 1) + (int64_t)(k + 1) * lda;
        *dkk = upd(*dkk, v1, w1, v1, w1);
      }
    }
  }
  if (!colj) return;
  double2* col = A + (int64_t)j * lda;
  const double2 vj = skip ? make_double2(0.0, 0.0) : v[j];
  const double2 wj = skip ? make_double2(0.0, 0.0) : wof(j);
  double sr = 0.0, si = 0.0;
  auto row = [&](int i, double2 a) {
    if (!skip) {
      a = upd(a, v[i], wof(i), vj, wj);
      col[i] = a;
    }
    if (!last && i > k + 1) {
      const double2 t = zconjmul(a, i == k + 2 ? v01 : vn[i]);
      sr += t.x;
      si += t.y;
    }
  };
#pragma unroll
  for (int q = 0; q < PF; ++q) {
    const int i = k + 1 + lane + 64 * q;
    if (i < n) row(i, pre[q]);
  }
  for (int i = k + 1 + lane + 64 * PF; i < n; i += 64) row(i, col[i]);
  if (!last) {
    sr = wave_sum(sr) * iv1;
    si = wave_sum(si) * iv1;
    if (lane == 0) W.uu[(size_t)((k + 1) & 1) * n + j] = make_double2(sr, si);
  }
}


// ---- 2. eigenvalues by multisection: one wavefront per eigenvalue; the 64
// lanes count eigenvalues below 64 interior points, so every step narrows the
// bracket 65-fold.  When the eigenvalues of a launch fit one round of the CUs,
// four wavefronts share an eigenvalue (256 points, 257-fold).  Then the
// clusters (one thread).
constexpr int BSW = 4;  // wavefronts (eigenvalues) per workgroup

// With `shared` set, the BSW wavefronts of a workgroup work on one eigenvalue
// together (64 BSW points per step, one wavefront per SIMD): for launches with
// few eigenvalues, where the CUs would otherwise idle.
__global__ __launch_bounds__(BSW * 64) void k_bisect(int n, int nev, int shared, double* __restrict__ work,
                                                    int64_t strideWork) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* dd = reinterpret_cast<double*>(smem);  // [n]
  double* e2 = dd + n;                            // [n]
  const int mat = blockIdx.y;
  HeevWork W(work + mat * strideWork, n, nev);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  double glo = 1e300, ghi = -1e300, emax = 0.0;
  for (int i = threadIdx.x; i < n; i += BSW * 64) {
    const double ei = W.e[i];
    dd[i] = W.d[i];
    e2[i] = ei * ei;
  }
  __syncthreads();
  // Gershgorin bounds (every wave computes them; they are cheap)
  for (int i = lane; i < n; i += 64) {
    const double r = (i > 0 ? sqrt(e2[i - 1]) : 0.0) + (i + 1 < n ? sqrt(e2[i]) : 0.0);
    glo = fmin(glo, dd[i] - r);
    ghi = fmax(ghi, dd[i] + r);
    if (i + 1 < n) emax = fmax(emax, e2[i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    glo = fmin(glo, __shfl_down(glo, o, 64));
    ghi = fmax(ghi, __shfl_down(ghi, o, 64));
    emax = fmax(emax, __shfl_down(emax, o, 64));
  }
  glo = __shfl(glo, 0, 64);
  ghi = __shfl(ghi, 0, 64);
  emax = __shfl(emax, 0, 64);
  const double tnorm = fmax(fabs(glo), fabs(ghi));
  const double pivmin = fmax(DBL_MIN * fmax(1.0, emax), 1e-300);
  const double span = ghi - glo;
  glo -= 2.0 * DBL_EPSILON * tnorm * n + 1e-300 + 1e-12 * span;
  ghi += 2.0 * DBL_EPSILON * tnorm * n + 1e-300 + 1e-12 * span;
  if (shared) {
    // invariant as below; point k (1..NP) sits at lo + (hi - lo) k / (NP + 1),
    // every thread recomputes the new bracket from the same formula
    __shared__ int s_p[BSW];
    constexpr int NP = BSW * 64;
    const int t = blockIdx.x;
    auto point = [](double lo, double hi, int k) {
      return fmin(fmax(lo + (hi - lo) * ((double)k / (double)(NP + 1)), lo), hi);
    };
    double lo = glo, hi = ghi;
    for (int it = 0; it < 64; ++it) {
      if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin) break;
      const int cnt = sturm_count(dd, e2, n, point(lo, hi, wid * 64 + lane + 1), pivmin);
      const uint64_t below = __ballot(cnt <= t);
      if (lane == 0) s_p[wid] = below ? wid * 64 + 64 - __clzll(below) : 0;
      __syncthreads();
      int p = 0;
#pragma unroll
      for (int w = 0; w < BSW; ++w) p = max(p, s_p[w]);
      __syncthreads();
      const double newlo = p > 0 ? point(lo, hi, p) : lo;
      const double newhi = p < NP ? point(lo, hi, p + 1) : hi;
      if (!(newlo > lo) && !(newhi < hi)) break;
      lo = newlo;
      hi = newhi;
    }
    if (threadIdx.x == 0) W.lam[t] = 0.5 * (lo + hi);
    return;
  }
  const int t = blockIdx.x * BSW + wid;
  if (t >= nev) return;
  // invariant: count(lo) <= t < count(hi); the points increase with the lane
  // and the Sturm count is monotone in x, so the lanes with count <= t form a
  // prefix 0..p: x_p is the new lo and x_(p+1) (or hi) the new hi
  double lo = glo, hi = ghi;
  for (int it = 0; it < 64; ++it) {
    if (hi - lo <= 2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin) break;
    const double x = fmin(fmax(lo + (hi - lo) * ((double)(lane + 1) / 65.0), lo), hi);
    const int cnt = sturm_count(dd, e2, n, x, pivmin);
    const uint64_t below = __ballot(cnt <= t);
    const int p = below ? 63 - __clzll(below) : -1;
    const double newlo = p >= 0 ? __shfl(x, p, 64) : lo;
    const double newhi = p < 63 ? __shfl(x, p + 1, 64) : hi;
    if (!(newlo > lo) && !(newhi < hi)) break;
    lo = newlo;
    hi = newhi;
  }
  if (lane == 0) W.lam[t] = 0.5 * (lo + hi);
}

// (d, e and the eigenvalues are staged in LDS, n + nev doubles, so that the
// one-thread scans run at LDS latency)
__global__ void k_clusters(int n, int nev, double* __restrict__ work, int64_t strideWork) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* tn = reinterpret_cast<double*>(smem);  // [n] |d_i| + 2|e_i|
  double* lam = tn + n;                           // [nev]
  const int mat = blockIdx.x;
  HeevWork W(work + mat * strideWork, n, nev);
  for (int i = threadIdx.x; i < n; i += blockDim.x) tn[i] = fabs(W.d[i]) + 2.0 * (i + 1 < n ? fabs(W.e[i]) : 0.0);
  for (int t = threadIdx.x; t < nev; t += blockDim.x) lam[t] = W.lam[t];
  __syncthreads();
  if (threadIdx.x != 0) return;
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i) tnorm = fmax(tnorm, tn[i]);
  // eigenvalues closer than 1e-5 |T| are computed in order and re-orthogonalised
  const double ortol = 1e-5 * fmax(tnorm, 1e-300);
  int nc = 0;
  for (int t = 0; t < nev; ++t)
    if (t == 0 || lam[t] - lam[t - 1] > ortol) W.cstart[nc++] = t;
  W.cstart[nc] = nev;
  *W.ncl = nc;
}

// ---- 3. inverse iteration: one thread per cluster, factors and vector in LDS
constexpr int IT = 64;  // threads per inverse-iteration workgroup

__global__ __launch_bounds__(IT) void k_inviter(int n, int nev, int cpw, double* __restrict__ work,
                                               int64_t strideWork) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* dd = reinterpret_cast<double*>(smem);  // [n]
  double* ee = dd + n;                            // [n]
  const int mat = blockIdx.y;
  HeevWork W(work + mat * strideWork, n, nev);
  for (int i = threadIdx.x; i < n; i += IT) {
    dd[i] = W.d[i];
    ee[i] = W.e[i];
  }
  __syncthreads();
  const int c = blockIdx.x * cpw + threadIdx.x;
  if (threadIdx.x >= cpw || c >= *W.ncl) return;
  double* mine = ee + n + (size_t)threadIdx.x * (6 * n + 1);  // odd stride: the lanes' arrays on distinct banks
  double* u0 = mine;
  double* u1 = u0 + n;
  double* u2 = u1 + n;
  double* lv = u2 + n;
  double* pv = lv + n;
  double* x = pv + n;
  double emax = 0.0, tnorm = 0.0;
  for (int i = 0; i < n; ++i) {
    emax = fmax(emax, ee[i] * ee[i]);
    tnorm = fmax(tnorm, fabs(dd[i]) + 2.0 * ee[i]);
  }
  const double pivmin = fmax(DBL_MIN * fmax(1.0, emax), 1e-300);
  const int t0 = W.cstart[c], t1 = W.cstart[c + 1];
  double prev = -1e300;
  for (int t = t0; t < t1; ++t) {
    double lmb = W.lam[t];
    const double sep = 10.0 * DBL_EPSILON * fmax(tnorm, 1e-300);
    if (t > t0 && lmb - prev < sep) lmb = prev + sep;
    prev = lmb;
    uint64_t h = 0x9E3779B97F4A7C15ull * (uint64_t)(t + 1);
    for (int i = 0; i < n; ++i) {
      h ^= h >> 31;
      h *= 0xBF58476D1CE4E5B9ull;
      h ^= h >> 29;
      x[i] = ((double)(h >> 11) * (1.0 / 9007199254740992.0)) - 0.5;
    }
    tri_factor(dd, ee, n, lmb, pivmin, u0, u1, u2, lv, pv);
    for (int it = 0; it < 3; ++it) {
      tri_solve(n, u0, u1, u2, lv, pv, x);
      for (int q = t0; q < t; ++q) {  // orthogonalise against the earlier cluster members
        const double* y = W.xt + (size_t)q * n;
        double sdot = 0.0;
        for (int i = 0; i < n; ++i) sdot += y[i] * x[i];
        for (int i = 0; i < n; ++i) x[i] -= sdot * y[i];
      }
      double nr = 0.0;
      for (int i = 0; i < n; ++i) nr += x[i] * x[i];
      const double inr = nr > 0.0 ? 1.0 / sqrt(nr) : 0.0;
      for (int i = 0; i < n; ++i) x[i] *= inr;
    }
    double* out = W.xt + (size_t)t * n;
    for (int i = 0; i < n; ++i) out[i] = x[i];
  }
}

// ---- 4. back-transformation, one wavefront per eigenvector
// waves per back-transform workgroup: four, so that the eigenvectors spread
// over many CUs (a wave's reflector steps are bound by its own FP64 issue: 16
// waves per workgroup put 80 vectors on five CUs, 130 us at n = 160)
constexpr int BW = 4;

__global__ __launch_bounds__(BW * 64) void k_backtransform(const double2* __restrict__ Aall, int lda,
                                                          int64_t strideA, int n, int nev,
                                                          double* __restrict__ work, int64_t strideWork,
                                                          double* __restrict__ Wall, int64_t strideW,
                                                          double2* __restrict__ Zall, int ldz, int64_t strideZ) {
  const int mat = blockIdx.y;
  const double2* A = Aall + mat * strideA;
  HeevWork W(work + mat * strideWork, n, nev);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int t = blockIdx.x * BW + wid;
  if (t >= nev) return;
  double2* z = Zall + mat * strideZ + (int64_t)t * ldz;
  const double* xs = W.xt + (size_t)t * n;
  for (int i = lane; i < n; i += 64) z[i] = make_double2(W.ph[i].x * xs[i], W.ph[i].y * xs[i]);
  if (lane == 0) Wall[mat * strideW + t] = W.lam[t];
  for (int k = n - 3; k >= 0; --k) {
    if (W.skip[k]) continue;
    const double2* vk = A + (int64_t)k * lda;  // reflector in rows k+1..n-1
    double sr = 0.0, si = 0.0;
    for (int i = k + 1 + lane; i < n; i += 64) {
      const double2 p = zconjmul(vk[i], z[i]);
      sr += p.x;
      si += p.y;
    }
    sr = 2.0 * wave_sum(sr);
    si = 2.0 * wave_sum(si);
    for (int i = k + 1 + lane; i < n; i += 64) {
      const double2 p = zmul(vk[i], make_double2(sr, si));
      z[i] = make_double2(z[i].x - p.x, z[i].y - p.y);
    }
  }
}

// The same with the eigenvector in registers (rows lane + 64 q, q < NQ,
// n <= 64 NQ) and the reflector columns staged in LDS in panels of BTP
// columns: the workgroup loads the next panel (coalesced over rows) while its
// waves run the current one, and the skip flags sit in LDS, so no step waits
// on global memory.  LDS: 2 BTP n complex + n ints.
constexpr int BTP = 8;

template <int NQ>
__global__ __launch_bounds__(BW * 64) void k_backtransform_reg(const double2* __restrict__ Aall, int lda,
                                                              int64_t strideA, int n, int nev,
                                                              double* __restrict__ work, int64_t strideWork,
                                                              double* __restrict__ Wall, int64_t strideW,
                                                              double2* __restrict__ Zall, int ldz, int64_t strideZ) {
  extern __shared__ __align__(16) unsigned char smem[];
  double2* pan = reinterpret_cast<double2*>(smem);   // [2][BTP][n]
  int* skip = reinterpret_cast<int*>(pan + 2 * BTP * n);  // [n]
  const int mat = blockIdx.y;
  const double2* A = Aall + mat * strideA;
  HeevWork W(work + mat * strideWork, n, nev);
  constexpr int NT = BW * 64;
  constexpr int PL = (BTP * 64 * NQ + NT - 1) / NT;  // panel elements per thread (n <= 64 NQ)
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += NT) skip[i] = W.skip[i];
  // panel p holds columns k = n-3 - BTP p - c, c < BTP (rows > k, zero above)
  auto fetch = [&](int p, double2* r) {
#pragma unroll
    for (int e = 0; e < PL; ++e) {
      const int q = tid + e * NT, c = q / n, i = q - c * n, k = n - 3 - BTP * p - c;
      r[e] = (c < BTP && k >= 0 && i > k) ? A[i + (int64_t)k * lda] : make_double2(0.0, 0.0);
    }
  };
  auto store = [&](int p, const double2* r) {
    double2* b = pan + (size_t)(p & 1) * BTP * n;
#pragma unroll
    for (int e = 0; e < PL; ++e) {
      const int q = tid + e * NT;
      if (q < BTP * n) b[q] = r[e];
    }
  };
  double2 pre[PL];
  fetch(0, pre);
  store(0, pre);
  __syncthreads();
  const int lane = tid & 63, wid = tid >> 6;
  const int t = blockIdx.x * BW + wid;
  const bool act = t < nev;
  double2 z[NQ];
  if (act) {
    const double* xs = W.xt + (size_t)t * n;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int i = lane + 64 * q;
      z[q] = i < n ? make_double2(W.ph[i].x * xs[i], W.ph[i].y * xs[i]) : make_double2(0.0, 0.0);
    }
    if (lane == 0) Wall[mat * strideW + t] = W.lam[t];
  }
  const int np = n >= 3 ? (n - 3) / BTP + 1 : 0;
  for (int p = 0; p < np; ++p) {
    if (p + 1 < np) fetch(p + 1, pre);
    const double2* b = pan + (size_t)(p & 1) * BTP * n;
    if (act) {
      for (int c = 0; c < BTP; ++c) {
        const int k = n - 3 - BTP * p - c;
        if (k < 0) break;
        if (skip[k]) continue;
        const double2* col = b + (size_t)c * n;
        double2 v[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int i = lane + 64 * q;
          v[q] = i < n ? col[i] : make_double2(0.0, 0.0);
        }
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const double2 pr = zconjmul(v[q], z[q]);
          sr += pr.x;
          si += pr.y;
        }
        sr = 2.0 * wave_sum(sr);
        si = 2.0 * wave_sum(si);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const double2 pr = zmul(v[q], make_double2(sr, si));
          z[q] = make_double2(z[q].x - pr.x, z[q].y - pr.y);
        }
      }
    }
    if (p + 1 < np) store(p + 1, pre);
    __syncthreads();
  }
  if (act) {
    double2* zo = Zall + mat * strideZ + (int64_t)t * ldz;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int i = lane + 64 * q;
      if (i < n) zo[i] = z[q];
    }
  }
}

void heev_gpu(double2* A, int lda, int64_t strideA, int n, int nev, double* Wv, int64_t strideW, double2* Z, int ldz,
              int64_t strideZ, double* work, int64_t strideWork, int batch, hipStream_t s) {
  DEREK_REQUIRE(n >= 1 && nev >= 1 && nev <= n, "heev: need 1 <= nev <= n");
  DEREK_REQUIRE(n <= 2048, "heev: n must be <= 2048");
  {
    // the largest trailing block the LDS kernel holds
    int nbl = std::min(n, 98);
    while (nbl > 1 && tri_lds(nbl) > kTridiagLds) --nbl;
    const int k0 = n - nbl;
    // steps 0..k0-1 with two multi-workgroup launches per column while the
    // trailing matrix is larger than that, then the rest in one workgroup
    if (k0 > 0) {
      hipLaunchKernelGGL(k_wide_first, dim3(batch), dim3(64), 0, s, A, lda, strideA, n, nev, work, strideWork);
      hipLaunchKernelGGL(k_wide_gemv, dim3((unsigned)ceil_div(n - 1, WC), batch), dim3(WT), 0, s, A, lda, strideA, n,
                         nev, 0, work, strideWork);
      for (int k = 0; k < k0; ++k) {
        dim3 grid((unsigned)ceil_div(n - k - 1, WC), batch);
        hipLaunchKernelGGL(k_wide_step, grid, dim3(WT), sizeof(double2) * (size_t)n, s, A, lda, strideA, n, nev, k,
                           k == k0 - 1 ? 1 : 0, work, strideWork);
      }
    }
    allow_dynamic_lds(reinterpret_cast<const void*>(k_tridiag), kTridiagLds);
    hipLaunchKernelGGL(k_tridiag, dim3(batch), dim3(DT), tri_lds(nbl), s, A, lda, strideA, n, k0, nev, work,
                       strideWork);
    DEREK_HIP(hipGetLastError());
  }
  {
    const size_t lds = sizeof(double) * 2 * (size_t)n;
    // one eigenvalue per workgroup while they fit one round of the CUs
    const int shared = (int64_t)nev * batch <= 256 ? 1 : 0;
    hipLaunchKernelGGL(k_bisect, dim3((unsigned)(shared ? nev : ceil_div(nev, BSW)), batch), dim3(BSW * 64), lds, s, n,
                       nev, shared, work, strideWork);
    DEREK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_clusters, dim3(batch), dim3(64), sizeof(double) * ((size_t)n + nev), s, n, nev, work,
                       strideWork);
    DEREK_HIP(hipGetLastError());
  }
  {
    // clusters per workgroup: LDS holds d, e and 6 n + 1 doubles per cluster
    // thread, up to 156 KiB (at n = 600 that is five clusters per workgroup,
    // so the 300 clusters take one round of the CUs instead of two)
    constexpr size_t kInvLds = 156 * 1024;
    const int cpw =
        std::max(1, std::min(IT, (int)((kInvLds / sizeof(double) - 2 * (size_t)n) / (6 * (size_t)n + 1))));
    const size_t lds = sizeof(double) * (2 * (size_t)n + (6 * (size_t)n + 1) * cpw);
    DEREK_REQUIRE(lds <= kInvLds, "heev: inverse-iteration scratch exceeds the LDS budget");
    if (lds > 65536) allow_dynamic_lds(reinterpret_cast<const void*>(k_inviter), kInvLds);
    dim3 grid((unsigned)ceil_div(nev, cpw), batch);
    hipLaunchKernelGGL(k_inviter, grid, dim3(IT), lds, s, n, nev, cpw, work, strideWork);
    DEREK_HIP(hipGetLastError());
  }
  {
    dim3 grid((unsigned)ceil_div(nev, BW), batch);
    const size_t lds = sizeof(double2) * 2 * BTP * (size_t)n + sizeof(int) * (size_t)n;
#define BT_(NQ)                                                                                                  \
  do {                                                                                                           \
    if (lds > 65536) allow_dynamic_lds(reinterpret_cast<const void*>(k_backtransform_reg<NQ>), lds);             \
    hipLaunchKernelGGL(k_backtransform_reg<NQ>, grid, dim3(BW * 64), lds, s, A, lda, strideA, n, nev, work,      \
                       strideWork, Wv, strideW, Z, ldz, strideZ);                                                \
  } while (0)
    if (n <= 64) BT_(1);
    else if (n <= 128) BT_(2);
    else if (n <= 192) BT_(3);
    else if (n <= 256) BT_(4);
    else if (n <= 384) BT_(6);
    else if (n <= 512) BT_(8);
    else if (n <= 624) BT_(10);  // LDS: two 8-column panels of n rows, 260 n bytes <= 160 KiB
    else
      hipLaunchKernelGGL(k_backtransform, grid, dim3(BW * 64), 0, s, A, lda, strideA, n, nev, work, strideWork, Wv,
                         strideW, Z, ldz, strideZ);
#undef BT_
    DEREK_HIP(hipGetLastError());
  }
}

// ---------------------------------------------------------------- generalized problem pieces
// All matrices m x m, column-major, ld m; matrix kz of a batch at kz*m*m
// (blockIdx.y or blockIdx.x selects it).

// dsc = 1/sqrt(diag S) (0 for a zero column), S <- D S D
__global__ void k_unit_diag(double2* __restrict__ Sall, int m, double* __restrict__ dscall) {
  extern __shared__ __align__(16) unsigned char smem[];
  double* sd = reinterpret_cast<double*>(smem);
  const int kz = blockIdx.y;
  double2* S = Sall + (int64_t)kz * m * m;
  double* dsc = dscall + (int64_t)kz * m;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    const double d = S[i + (int64_t)i * m].x;
    const double v = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
    sd[i] = v;
    dsc[i] = v;
  }
  __syncthreads();
  const int64_t mm = (int64_t)m * m;
  for (int64_t q = threadIdx.x; q < mm; q += blockDim.x) {
    const int r = (int)(q % m), c = (int)(q / m);
    const double f = sd[r] * sd[c];
    S[q] = make_double2(S[q].x * f, S[q].y * f);
  }
}

// Blocked right-looking Cholesky of the unit-diagonal S with dropping, on a
// work copy Wm (lower triangle used).  A panel of nb columns is staged in
// LDS and factored there: column j is kept if its updated pivot exceeds
// tol * S_jj; a kept column is L's column j and updates the rest of the
// panel; a dropped column is zeroed and contributes nothing.  The panel then
// updates the trailing lower triangle in one pass (one wavefront per column,
// rows over the lanes).  Output: Lf (lower triangular) = L on the kept rows
// and columns, the identity on the dropped ones, keep[j], na = #kept, ok =
// (na >= nev).
__global__ __launch_bounds__(DT) void k_chol_drop(const double2* __restrict__ Sall, int m, int nb, double tol,
                                                  int nev, double2* __restrict__ Wall, double2* __restrict__ Lall,
                                                  int* __restrict__ keepall, int* __restrict__ naall,
                                                  int* __restrict__ okall) {
  extern __shared__ __align__(16) unsigned char smem[];
  double2* P = reinterpret_cast<double2*>(smem);               // [nb][m]: P[jj*m + i]
  double* sd = reinterpret_cast<double*>(P + (size_t)nb * m);  // [m] diagonal of S
  int* kp = reinterpret_cast<int*>(sd + m);                    // [m] keep flags
  const int kz = blockIdx.x;
  const int64_t mm = (int64_t)m * m;
  const double2* S = Sall + kz * mm;
  double2* Wm = Wall + kz * mm;
  double2* Lf = Lall + kz * mm;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  for (int64_t q = tid; q < mm; q += DT) Wm[q] = S[q];
  for (int i = tid; i < m; i += DT) sd[i] = S[i + (int64_t)i * m].x;
  __syncthreads();
  int c = 0;
  for (int p0 = 0; p0 < m; p0 += nb) {
    const int pn = min(nb, m - p0), rows = m - p0;
    for (int q = tid; q < pn * rows; q += DT) {
      const int jj = q / rows, i = p0 + q % rows;
      P[jj * m + i] = Wm[i + (int64_t)(p0 + jj) * m];
    }
    __syncthreads();
    for (int jj = 0; jj < pn; ++jj) {
      const int j = p0 + jj;
      double2* pc = P + jj * m;
      const double d = pc[j].x;
      if (d > tol * sd[j]) {
        const double idj = rsqrt_nr(d), lj = d * idj;
        double2* lcol = Lf + (int64_t)j * m;
        for (int i = j + 1 + tid; i < m; i += DT) {
          const double2 l = make_double2(pc[i].x * idj, pc[i].y * idj);
          pc[i] = l;
          lcol[i] = l;
        }
        if (tid == 0) {
          lcol[j] = make_double2(lj, 0.0);
          kp[j] = 1;
        }
        __syncthreads();
        // rest of the panel: P(i,kk) -= l_i conj(l_(p0+kk)), i >= p0+kk
        for (int kk = jj + 1 + wid; kk < pn; kk += DW) {
          const double2 lk = pc[p0 + kk];
          double2* pk = P + kk * m;
          for (int i = p0 + kk + lane; i < m; i += 64) {
            const double2 li = pc[i];
            pk[i] = make_double2(pk[i].x - (li.x * lk.x + li.y * lk.y), pk[i].y - (li.y * lk.x - li.x * lk.y));
          }
        }
        __syncthreads();
        ++c;
      } else {
        // only column jj's own rows are written; the next column reads
        // column jj+1, and the trailing update waits for a barrier
        for (int i = j + tid; i < m; i += DT) pc[i] = make_double2(0.0, 0.0);
        if (tid == 0) kp[j] = 0;
      }
    }
    __syncthreads();
    // trailing update W(i,k) -= sum_jj P(i,jj) conj(P(k,jj)), i >= k >= p0+pn
    for (int k = p0 + pn + wid; k < m; k += DW) {
      double2* wk = Wm + (int64_t)k * m;
      for (int i = k + lane; i < m; i += 64) {
        double2 acc = wk[i];
        for (int jj = 0; jj < pn; ++jj) {
          const double2 li = P[jj * m + i], lk = P[jj * m + k];
          acc.x -= li.x * lk.x + li.y * lk.y;
          acc.y -= li.y * lk.x - li.x * lk.y;
        }
        wk[i] = acc;
      }
    }
    __syncthreads();
  }
  // upper triangle zero; dropped rows and columns -> identity
  for (int k = wid; k < m; k += DW) {
    double2* lcol = Lf + (int64_t)k * m;
    for (int i = lane; i < m; i += 64)
      if (i < k || !kp[i] || !kp[k]) lcol[i] = make_double2(i == k ? 1.0 : 0.0, 0.0);
  }
  for (int i = tid; i < m; i += DT) keepall[(int64_t)kz * m + i] = kp[i];
  if (tid == 0) {
    naall[kz] = c;
    okall[kz] = c >= nev ? 1 : 0;
  }
}

// Inverse of a lower-triangular m x m matrix, one wavefront per column j:
// forward substitution on e_j with the solution and the current column of L
// in registers (row i lives in lane i % 64, slot i / 64); the pivot and x_k
// are read across lanes with readlane.  The columns of L are L2 reads: D of
// them are kept in flight in a register ring (slot d holds column k0 + d and
// is refilled with column k0 + d + D once used), so a step waits for a load
// issued D steps earlier instead of one.  TR = rows per lane (m <= 64 TR).
constexpr int TR = 10;  // largest rows per lane: m <= 640

__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// rows i >= c of column c of L (zero elsewhere, and for c >= m)
template <int NR>
__device__ __forceinline__ void trinv_load(const double2* __restrict__ L, int m, int c, int lane, double2 (&col)[NR]) {
#pragma unroll
  for (int t = 0; t < NR; ++t) {
    const int i = lane + 64 * t;
    col[t] = (c < m && i >= c && i < m) ? L[i + (int64_t)c * m] : make_double2(0.0, 0.0);
  }
}

template <int NR, int D>
__global__ __launch_bounds__(64) void k_trinv(const double2* __restrict__ Lall, int m,
                                              double2* __restrict__ Linvall) {
  const int j = blockIdx.x, lane = threadIdx.x;
  const int64_t mm = (int64_t)m * m;
  const double2* L = Lall + blockIdx.y * mm;
  double2 x[NR], ring[D][NR];
#pragma unroll
  for (int t = 0; t < NR; ++t) {
    const int i = lane + 64 * t;
    x[t] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
#pragma unroll
  for (int d = 0; d < D; ++d) trinv_load<NR>(L, m, j + d, lane, ring[d]);
  for (int k0 = j; k0 < m; k0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int k = k0 + d;
      if (k >= m) break;  // uniform over the wavefront
      const int tk = k >> 6, lk = k & 63;
      double2 xo = make_double2(0.0, 0.0), dg = make_double2(1.0, 0.0);
#pragma unroll
      for (int t = 0; t < NR; ++t)
        if (t == tk) {
          xo = x[t];
          dg = ring[d][t];
        }
      xo = make_double2(readlane_d(xo.x, lk), readlane_d(xo.y, lk));
      dg = make_double2(readlane_d(dg.x, lk), readlane_d(dg.y, lk));
      const double den = 1.0 / (dg.x * dg.x + dg.y * dg.y);
      const double2 xk = make_double2((xo.x * dg.x + xo.y * dg.y) * den, (xo.y * dg.x - xo.x * dg.y) * den);
#pragma unroll
      for (int t = 0; t < NR; ++t) {
        const int i = lane + 64 * t;
        if (i > k) {
          const double2 u = zmul(ring[d][t], xk);
          x[t] = make_double2(x[t].x - u.x, x[t].y - u.y);
        } else if (i == k) {
          x[t] = xk;
        }
      }
      trinv_load<NR>(L, m, k + D, lane, ring[d]);
    }
  }
  double2* out = Linvall + blockIdx.y * mm + (int64_t)j * m;
#pragma unroll
  for (int t = 0; t < NR; ++t) {
    const int i = lane + 64 * t;
    if (i < m) out[i] = x[t];
  }
}

// B = A^H
__global__ void k_conj_transpose(const double2* __restrict__ A, int m, double2* __restrict__ B) {
  const int64_t mm = (int64_t)m * m;
  A += blockIdx.y * mm;
  B += blockIdx.y * mm;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < mm; q += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(q % m), c = (int)(q / m);
    const double2 a = A[c + (int64_t)r * m];
    B[q] = make_double2(a.x, -a.y);
  }
}

// Hc = (D H D) on the kept rows and columns, 0 elsewhere
__global__ void k_scale_mask(const double2* __restrict__ H, const double* __restrict__ dsc,
                             const int* __restrict__ keep, int m, double2* __restrict__ Hc) {
  const int64_t mm = (int64_t)m * m;
  const int kz = blockIdx.y;
  H += kz * mm;
  Hc += kz * mm;
  dsc += (int64_t)kz * m;
  keep += (int64_t)kz * m;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < mm; q += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(q % m), c = (int)(q / m);
    const double f = (keep[r] && keep[c]) ? dsc[r] * dsc[c] : 0.0;
    Hc[q] = make_double2(H[q].x * f, H[q].y * f);
  }
}

// Dropped diagonal entries of A <- 2 max_i sum_j |A_ij| + 1 (above every
// eigenvalue of the kept block), so the lowest eigenpairs are the kept ones.
// Row sums: R = min(m, DT) rows per pass over consecutive threads (coalesced
// column reads), the columns split over DT / R thread groups whose parts are
// added through LDS.
__global__ __launch_bounds__(DT) void k_big_diag(double2* __restrict__ Aall, const int* __restrict__ keepall, int m) {
  __shared__ double red[DW];
  __shared__ double part[DT];
  const int kz = blockIdx.x;
  double2* A = Aall + (int64_t)kz * m * m;
  const int* keep = keepall + (int64_t)kz * m;
  const int R = min(m, DT), cg = DT / R;
  const int t = threadIdx.x % R, g = threadIdx.x / R;
  double rmax = 0.0;
  for (int i0 = 0; i0 < m; i0 += R) {
    const int i = i0 + t;
    double r = 0.0;
    if (g < cg && i < m) {
#pragma unroll 4
      for (int j = g; j < m; j += cg) {
        const double2 a = A[i + (int64_t)j * m];
        r += sqrt(a.x * a.x + a.y * a.y);
      }
    }
    if (g < cg) part[g * R + t] = r;
    __syncthreads();
    if (threadIdx.x < R && i0 + (int)threadIdx.x < m) {
      double sum = 0.0;
      for (int q = 0; q < cg; ++q) sum += part[q * R + threadIdx.x];
      rmax = fmax(rmax, sum);
    }
    __syncthreads();
  }
  const double big = 2.0 * block_max_all(rmax, red) + 1.0;
  for (int i = threadIdx.x; i < m; i += DT)
    if (!keep[i]) A[i + (int64_t)i * m] = make_double2(big, 0.0);
}

// U (m x nev) = D C on the kept rows, 0 on the dropped ones; where ok == 0
// the subspace step is void: U = [I; 0] and w keeps its previous values.
__global__ void k_expand(const double2* __restrict__ Call, const double* __restrict__ dscall,
                         const int* __restrict__ keepall, const int* __restrict__ okall, int m, int nev,
                         const double* __restrict__ wnew, double* __restrict__ w, double2* __restrict__ Uall) {
  const int kz = blockIdx.y;
  const int ok = okall[kz];
  const double2* C = Call + (int64_t)kz * m * m;
  const double* dsc = dscall + (int64_t)kz * m;
  const int* keep = keepall + (int64_t)kz * m;
  double2* U = Uall + (int64_t)kz * m * nev;
  const int64_t n = (int64_t)m * nev;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(q % m), c = (int)(q / m);
    double2 v;
    if (ok) {
      const double f = keep[r] ? dsc[r] : 0.0;
      const double2 x = C[r + (int64_t)c * m];
      v = make_double2(x.x * f, x.y * f);
    } else {
      v = make_double2(r == c ? 1.0 : 0.0, 0.0);
    }
    U[q] = v;
  }
  if (blockIdx.x == 0 && ok)
    for (int c = threadIdx.x; c < nev; c += blockDim.x) w[(int64_t)kz * nev + c] = wnew[(int64_t)kz * nev + c];
}

// A <- (A + A^H)/2
__global__ void k_hermitize(double2* __restrict__ A, int m) {
  const int64_t mm = (int64_t)m * m;
  A += blockIdx.y * mm;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < mm; q += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(q % m), c = (int)(q / m);
    if (r > c) continue;
    const double2 a = A[q], b = A[c + (int64_t)r * m];
    const double2 h = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    A[q] = r == c ? make_double2(h.x, 0.0) : h;
    if (r != c) A[c + (int64_t)r * m] = make_double2(h.x, -h.y);
  }
}

static dim3 mgrid(int m, int batch) {
  return dim3((unsigned)std::min<int64_t>(256, ceil_div((int64_t)m * m, 256)), (unsigned)batch);
}

void hermitize_gpu(double2* A, int m, int batch, hipStream_t s) {
  hipLaunchKernelGGL(k_hermitize, mgrid(m, batch), dim3(256), 0, s, A, m);
  DEREK_HIP(hipGetLastError());
}

void unit_diag_gpu(double2* S, int m, double* dsc, int batch, hipStream_t s) {
  DEREK_REQUIRE(m <= 8192, "unit_diag: matrix too large");
  // one workgroup per matrix: S is scaled in place after its diagonal is read
  hipLaunchKernelGGL(k_unit_diag, dim3(1, batch), dim3(1024), sizeof(double) * m, s, S, m, dsc);
  DEREK_HIP(hipGetLastError());
}

void chol_drop_gpu(const double2* S, int m, double tol, int nev, double2* Wm, double2* Lf, int* keep, int* na,
                   int* ok, int batch, hipStream_t s) {
  constexpr size_t kLds = 60 * 1024;
  const size_t per_row = sizeof(double2) + sizeof(double) + sizeof(int);
  DEREK_REQUIRE(m >= 1 && (size_t)m * per_row <= kLds, "chol_drop: matrix too large");
  const int nb = (int)std::min<size_t>(16, (kLds - (sizeof(double) + sizeof(int)) * m) / (sizeof(double2) * m));
  const size_t lds = sizeof(double2) * (size_t)nb * m + (sizeof(double) + sizeof(int)) * m;
  hipLaunchKernelGGL(k_chol_drop, dim3(batch), dim3(DT), lds, s, S, m, nb, tol, nev, Wm, Lf, keep, na, ok);
  DEREK_HIP(hipGetLastError());
}

void trinv_gpu(const double2* L, int m, double2* Linv, int batch, hipStream_t s) {
  DEREK_REQUIRE(m <= 64 * TR, "trinv: matrix too large");
  const dim3 grid((unsigned)m, batch);
  if (m <= 64)
    hipLaunchKernelGGL((k_trinv<1, 8>), grid, dim3(64), 0, s, L, m, Linv);
  else if (m <= 192)
    hipLaunchKernelGGL((k_trinv<3, 8>), grid, dim3(64), 0, s, L, m, Linv);
  else
    hipLaunchKernelGGL((k_trinv<TR, 4>), grid, dim3(64), 0, s, L, m, Linv);
  DEREK_HIP(hipGetLastError());
}

void conj_transpose_gpu(const double2* A, int m, double2* B, int batch, hipStream_t s) {
  hipLaunchKernelGGL(k_conj_transpose, mgrid(m, batch), dim3(256), 0, s, A, m, B);
  DEREK_HIP(hipGetLastError());
}

void scale_mask_gpu(const double2* H, const double* dsc, const int* keep, int m, double2* Hc, int batch,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_scale_mask, mgrid(m, batch), dim3(256), 0, s, H, dsc, keep, m, Hc);
  DEREK_HIP(hipGetLastError());
}

void big_diag_gpu(double2* A, const int* keep, int m, int batch, hipStream_t s) {
  hipLaunchKernelGGL(k_big_diag, dim3(batch), dim3(DT), 0, s, A, keep, m);
  DEREK_HIP(hipGetLastError());
}

void expand_gpu(const double2* C, const double* dsc, const int* keep, const int* ok, int m, int nev,
                const double* wnew, double* w, double2* U, int batch, hipStream_t s) {
  const int64_t n = (int64_t)m * nev;
  hipLaunchKernelGGL(k_expand, dim3((unsigned)std::min<int64_t>(256, ceil_div(n, 256)), batch), dim3(256), 0, s, C,
                     dsc, keep, ok, m, nev, wnew, w, U);
  DEREK_HIP(hipGetLastError());
}

}  // namespace derek
