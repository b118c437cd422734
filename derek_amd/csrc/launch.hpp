// Host-side launchers of the gfx950 kernels (all stream-ordered, asynchronous).
#pragma once
#include "common.hpp"
#include "fft.hpp"

namespace derek {

// Twiddle table exp(-2*pi*i*t/n), t = 0..n-1, in the precision R.
template <typename R> struct Twiddles {
  int n = 0;
  cx_t<R>* d = nullptr;
};

// One 1D transform axis of a flat z-fastest grid: transform length n,
// element stride `stride` (1 for z), plan, twiddles.
template <typename R> struct Axis {
  int n;
  FftPlan plan;
  const cx_t<R>* tw;
};

// ---- FFT passes on batches of flat grids (nx, ny, nz), z fastest.
// Lines along z: nlines contiguous transforms of length ax.n.
template <typename R> void fft_z(cx_t<R>* data, int64_t nlines, const Axis<R>& ax, int sign, hipStream_t s);
// Strided axis: element (o, a, i) at o*n*stride + a*stride + i, i < stride.
template <typename R>
void fft_strided(cx_t<R>* data, int64_t nouter, int64_t stride, const Axis<R>& ax, int sign, hipStream_t s);
// Fused: inverse transform along the strided axis, multiply by vloc*vscale
// (vloc indexed a*stride + i, shared by all outer blocks), forward transform.
template <typename R>
void fft_strided_vmul(cx_t<R>* data, int64_t nouter, int64_t stride, const Axis<R>& ax, const R* vloc,
                      R vscale, hipStream_t s);
// Fused: inverse transform along the strided axis for outer = 0..nouter-1
// (bands) and rho[a*stride+i] = keep*rho + scale * sum_o occ[o] |.|^2.
template <typename R>
void fft_strided_density(const cx_t<R>* data, int64_t nouter, int64_t stride, const Axis<R>& ax,
                         const double* occ, double* rho, double scale, int keep, hipStream_t s);

// ---- plane-wave <-> grid
template <typename R> void zero(cx_t<R>* p, int64_t n, hipStream_t s);
template <typename R>
void scatter(const cx_t<R>* psi, int64_t ldpsi, int nb, const int32_t* idx, int64_t npw, cx_t<R>* grid,
             int64_t N, hipStream_t s);
// hpsi = grid[idx] + kin * psi
template <typename R>
void gather_kin(const cx_t<R>* grid, int64_t N, const int32_t* idx, const double* kin, const cx_t<R>* psi,
                int64_t ldpsi, cx_t<R>* hpsi, int64_t ldh, int nb, int64_t npw, hipStream_t s);

// ---- real-space fields (FP64)
void lda_xc(const double* rho, double* vxc, double* partial_exc, int64_t n, hipStream_t s);
int lda_partials(int64_t n);
void lda_eps(const double* rho, double* eps, int64_t n, hipStream_t s);
void real_to_cx(const double* in, double2* out, int64_t n, hipStream_t s);
void cx_real(const double2* in, double* out, double scale, int64_t n, hipStream_t s);
// g(G) *= fac(|G|^2): mode 0 Hartree 4*pi/G^2 * scale (G=0 -> 0);
// mode 1 Kerker G^2/(G^2+q0sq) * scale.
void gspace_scale(double2* g, const double* gsq, double q0sq, double scale, int mode, int64_t n, hipStream_t s);
// out = a + b + c (any may be null -> 0)
void add3(const double* a, const double* b, const double* c, double* out, int64_t n, hipStream_t s);
// partial dot products: partial[blk] = sum a*b over the block's slice
void dot_partials(const double* a, const double* b, double* partial, int64_t n, hipStream_t s);
int dot_nblocks(int64_t n);
// out = sum_i c[i] * x_i  (x_i = xs[i], up to 16 vectors, coefficients on host)
void lincomb(double* out, const double* const* xs, const double* c, int m, int64_t n, hipStream_t s);
void axpy(double* y, const double* x, double a, int64_t n, hipStream_t s);
void sub(double* out, const double* a, const double* b, int64_t n, hipStream_t s);

// ---- Davidson helpers (per band b: column b of leading dimension ld)
// partial[(b*nchunk + c)*2 + {0,1}] = sum over chunk c of {kin|x|^2, |hx - eps x|^2}
template <typename R>
void band_sums(const cx_t<R>* x, const cx_t<R>* hx, int64_t ld, const double* eps, const double* kin, int nb,
               int64_t npw, double* partial, int nchunk, hipStream_t s);
// w = TPA(kin/ekin_b) * (hx - eps_b x), for the active bands listed in act[0..nact)
template <typename R>
void tpa_residual(const cx_t<R>* x, const cx_t<R>* hx, int64_t ld, const double* eps, const double* ekin,
                  const int32_t* act, int nact, const double* kin, int64_t npw, cx_t<R>* w, int64_t ldw,
                  hipStream_t s);
// zero rows npw..ld-1 of nb columns
template <typename R> void zero_pad(cx_t<R>* x, int64_t ld, int64_t npw, int nb, hipStream_t s);
// psi(g, b) = (N + iN)/(1 + kin) from the counter-based generator
void init_psi(double2* psi, int64_t ld, int nb, int64_t npw, const double* kin, uint64_t seed, int k,
              hipStream_t s);
template <typename Rd, typename Rs>
void convert(cx_t<Rd>* dst, const cx_t<Rs>* src, int64_t n, hipStream_t s);
void convert_real_f(float* dst, const double* src, int64_t n, hipStream_t s);

// ---- tall-skinny complex GEMMs (MFMA)
// C (m x n, column-major, ldc) = X^H Y; X: npw x m (ldx), Y: npw x n (ldy).
// `work` must hold overlap_work_elems(m, n, npw) double2.
int64_t overlap_work_elems(int m, int n, int64_t npw);
void overlap_f64(const double2* X, int64_t ldx, const double2* Y, int64_t ldy, int m, int n, int64_t npw,
                 double2* C, int ldc, double2* work, bool hermitian, hipStream_t s);
// Z (npw x n, ldz) = X (npw x m, ldx) * U (m x n, column-major, ldu).
void rotate_f64(const double2* X, int64_t ldx, const double2* U, int ldu, int m, int n, int64_t npw, double2* Z,
                int64_t ldz, hipStream_t s);
// Mixed precision: X, Y complex64; bf16 MFMA with fp32 accumulation; C in FP64.
int64_t overlap_work_elems_bf16(int m, int n, int64_t npw);
void overlap_bf16(const float2* X, int64_t ldx, const float2* Y, int64_t ldy, int m, int n, int64_t npw,
                  double2* C, int ldc, float2* work, bool hermitian, hipStream_t s);
void rotate_f32(const float2* X, int64_t ldx, const float2* U, int ldu, int m, int n, int64_t npw, float2* Z,
                int64_t ldz, hipStream_t s);

}  // namespace derek
