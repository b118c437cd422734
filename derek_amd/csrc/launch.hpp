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

// Allows a kernel more than 64 KiB of dynamic LDS (set once per kernel).
void allow_dynamic_lds(const void* kern, size_t bytes);

// ---- FFT passes on batches of flat grids (nx, ny, nz), z fastest.
// Lines along z: nlines contiguous transforms of length ax.n.
template <typename R> void fft_z(cx_t<R>* data, int64_t nlines, const Axis<R>& ax, int sign, hipStream_t s);
// Strided axis: element (o, a, i) at o*n*stride + a*stride + i, i < stride.
template <typename R>
void fft_strided(cx_t<R>* data, int64_t nouter, int64_t stride, const Axis<R>& ax, int sign, hipStream_t s);
// ---- buffers
template <typename R> void zero(cx_t<R>* p, int64_t n, hipStream_t s);
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
void axpy(double* y, const double* x, double a, int64_t n, hipStream_t s);
void sub(double* out, const double* a, const double* b, int64_t n, hipStream_t s);
void wave_axpby(double2* z, double2 a, const double2* x, double2 b, const double2* y, int64_t n, hipStream_t s);

template <typename Rd, typename Rs>
void convert(cx_t<Rd>* dst, const cx_t<Rs>* src, int64_t n, hipStream_t s);
void convert_real_f(float* dst, const double* src, int64_t n, hipStream_t s);
// t = scale * v with x and z swapped: t[(z*ny + y)*nx + x]; t32 (nullable) its FP32 copy.
void transpose_xz(const double* v, double* t64, float* t32, int nx, int ny, int nz, double scale, hipStream_t s);

// ---- tall-skinny complex GEMMs (MFMA), batched: matrix kz of a batch is at
// X + kz*sx, Y + kz*sy, C + kz*sc ...; npwk (device, nullable) gives each
// matrix its own row count (<= npw), npw otherwise.
// C (m x n, column-major, ldc) = X^H Y; X: npw x m (ldx), Y: npw x n (ldy),
// accumulated on MFMA in R, chunk partials summed in FP64.
// `work` (work_cap double2, checked) must hold overlap_work_elems(m, n, npw, batch),
// or batch * max_chunks * m * n when max_chunks > 0 caps the row chunks;
// `rounds` (1, 2, 4): rounds of resident workgroups the row chunks fill at least;
// `tile` (64, 128): the C tile edge of the BF16 kernels (the others take 64).
// bf16_terms (complex64 data only): 0 multiplies on the FP32 matrix cores;
// 1..3 on the BF16 ones with every FP32 value split into that many bf16
// terms (1: plain BF16, 2: ~16 significant bits, 3: ~FP32), FP32 accumulation.
int64_t overlap_work_elems(int m, int n, int64_t npw, int batch);
// the SCF engine's partial-sum buffer: every nb x nb and 2nb x 2nb overlap
// over batches of kc (and the last nk mod kc) k-points
int64_t engine_overlap_work_elems(int nb, int nk, int kc, int64_t npw);
template <typename R>
void overlap(const cx_t<R>* X, int64_t ldx, int64_t sx, const cx_t<R>* Y, int64_t ldy, int64_t sy, int m, int n,
             int64_t npw, const int* npwk, double2* C, int ldc, int64_t sc, int batch, double2* work,
             int64_t work_cap, bool hermitian, hipStream_t s, int bf16_terms = 0, int64_t max_chunks = 0,
             int rounds = 2, int tile = 64);
// Z (npw x n, ldz) = X (npw x m, ldx) * U (m x n, FP64, column-major, ldu),
// U rounded to R; bf16_terms as for overlap.
template <typename R>
void rotate(const cx_t<R>* X, int64_t ldx, int64_t sx, const double2* U, int ldu, int64_t su, int m, int n,
            int64_t npw, const int* npwk, cx_t<R>* Z, int64_t ldz, int64_t sz, int batch, hipStream_t s,
            int bf16_terms = 0);

}  // namespace derek
