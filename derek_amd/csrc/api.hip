// Context, C-ABI entry points and the resident SCF step.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#include <chrono>

#include "../../include/derek_hip.h"
#include "basis.hpp"
#include "dense.hpp"
#include "dense_gpu.hpp"
#include "launch.hpp"
#include "pw_fft.hpp"

using namespace derek;
using cd = std::complex<double>;

namespace {

thread_local std::string g_last_error;

template <typename F> int32_t guard(F&& f) {
  try {
    f();
    return DEREK_OK;
  } catch (const ArgError& e) {
    g_last_error = e.what();
    return DEREK_ERR_ARG;
  } catch (const NumericError& e) {
    g_last_error = e.what();
    return DEREK_ERR_NUMERIC;
  } catch (const StateError& e) {
    g_last_error = e.what();
    return DEREK_ERR_STATE;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return DEREK_ERR_HIP;
  }
}

// Device buffer that grows on demand.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T* get(size_t n) {
    const size_t need = n * sizeof(T);
    if (need > bytes) {
      release();
      DEREK_HIP(hipMalloc(&p, need));
      bytes = need;
    }
    return reinterpret_cast<T*>(p);
  }
};

template <typename T> T* dmalloc(size_t n) {
  void* p = nullptr;
  DEREK_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
  return reinterpret_cast<T*>(p);
}

inline void dfree(void* p) {
  if (p) (void)hipFree(p);
}

int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

}  // namespace

// Per-kernel-class timers: HIP events recorded on the context's stream around
// each launch while enabled, resolved lazily; algorithmic bytes and flops of
// each launch are accumulated beside the times.
enum KernelClass {
  KC_STICKS_Z_INV = 0,
  KC_PLANES_Y_INV,
  KC_PLANES_X_VMUL,
  KC_PLANES_Y_FWD,
  KC_STICKS_Z_FWD,
  KC_PLANES_X_DENSITY,
  KC_OVERLAP,
  KC_ROTATE,
  KC_FFT_Z,
  KC_FFT_Y,
  KC_FFT_X,
  KC_COUNT
};

struct Profiler {
  struct Pending {
    int cls;
    hipEvent_t a, b;
    double bytes, flops;
  };
  bool on = false;
  std::vector<hipEvent_t> pool;
  std::vector<Pending> pending;
  int64_t count[KC_COUNT] = {};
  double ms[KC_COUNT] = {}, bytes[KC_COUNT] = {}, flops[KC_COUNT] = {};
  hipEvent_t take() {
    if (pool.empty()) {
      hipEvent_t e;
      DEREK_HIP(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  void resolve() {
    for (auto& p : pending) {
      DEREK_HIP(hipEventSynchronize(p.b));
      float t = 0.f;
      DEREK_HIP(hipEventElapsedTime(&t, p.a, p.b));
      count[p.cls] += 1;
      ms[p.cls] += t;
      bytes[p.cls] += p.bytes;
      flops[p.cls] += p.flops;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  void reset() {
    resolve();
    for (int i = 0; i < KC_COUNT; ++i) {
      count[i] = 0;
      ms[i] = bytes[i] = flops[i] = 0.0;
    }
  }
  ~Profiler() {
    for (auto& p : pending) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

struct KState {
  int64_t npw = 0, ld = 0;
  int32_t* idx = nullptr;
  double* kin = nullptr;
  double weight = 0.0;
  bool set = false;
  PwMaps maps;                  // sphere maps (device pointers into mapbuf)
  int32_t* mapbuf = nullptr;
  // SCF wavefunction blocks, 2*nbands columns of leading dimension ld:
  // Y = [X | W], HY = [HX | HW]; Y2/HY2 receive rotations.
  void* Y = nullptr;
  void* HY = nullptr;
  void* Y2 = nullptr;
  void* HY2 = nullptr;
};

struct derek_ctx {
  int dims[3] = {0, 0, 0};
  int64_t N = 0;
  int nk = 0, nb = 0, prec = DEREK_FP64;
  double volume = 0.0;
  int64_t max_npw = 0;
  hipStream_t s = nullptr;
  FftPlan plan[3];
  double2* tw64[3] = {nullptr, nullptr, nullptr};
  float2* tw32[3] = {nullptr, nullptr, nullptr};
  Axis<double> ax64[3];
  Axis<float> ax32[3];
  std::vector<KState> ks;
  double* gsq = nullptr;
  double* vext = nullptr;
  double* vloc = nullptr;
  float* vloc32 = nullptr;
  bool gsq_set = false, vloc_set = false, vext_set = false;
  int batch = 8;
  // pruned-FFT work buffers for one band batch (sized for the largest k)
  int64_t st_elems = 0, pl_elems = 0;
  double2 *S64 = nullptr, *P64 = nullptr;
  float2 *S32 = nullptr, *P32 = nullptr;
  template <typename R> cx_t<R>* sticks();
  template <typename R> cx_t<R>* planes();
  void ensure_pw_buffers();
  double2* cgrid = nullptr;
  DevBuf tA, tB, tC, tD, ovwork, partial;
  // SCF
  bool scf_ready = false;
  derek_scf_params P{};
  int iscf = 0;
  double *rho_in = nullptr, *rho_out = nullptr, *vh = nullptr, *vxc = nullptr, *vhxc = nullptr, *tmp1 = nullptr,
         *tmp2 = nullptr;
  std::vector<double*> h_rin, h_res;
  std::vector<int> order;  // history slots, oldest first
  std::vector<std::vector<double>> A;
  std::vector<double> eig, occ, res;
  double* d_eps = nullptr;
  double* d_ekin = nullptr;
  double* d_occw = nullptr;
  int32_t* d_act = nullptr;
  double2* d_S = nullptr;
  double2* d_H = nullptr;
  double2* d_U = nullptr;
  // GPU subspace-solver scratch (sized for m = 2 nbands)
  double* d_dsc = nullptr;
  double* d_w = nullptr;
  double* d_hwork = nullptr;
  int* d_sact = nullptr;
  int* d_na = nullptr;
  double2 *d_Lb = nullptr, *d_Lc = nullptr, *d_Hc = nullptr, *d_Linv = nullptr, *d_LinvH = nullptr, *d_T = nullptr,
          *d_A = nullptr, *d_Zs = nullptr, *d_Ca = nullptr;
  int dense_cap = 0;
  int sub_cap = 0;  // columns d_S / d_H / d_U can hold (sub_cap x sub_cap)
  void ensure_sub(int m);
  void ensure_dense(int m, int nev);
  void rr_gpu(int n, double* w_host);
  bool gen_gpu(int m, int nev, double* w_host);
  double fermi = 0.0, minus_ts = 0.0;
  double t_dense = 0.0;  // host seconds in the subspace solves of the current step
  bool have_rho_out = false;
  double* rho_target = nullptr;  // density accumulation target
  void set_rho_target(double* p) { rho_target = p; }
  double* rho_out_target() const { return rho_target; }
  Profiler prof;
  template <typename F> void timed(int cls, double bytes, double flops, F&& launch) {
    if (!prof.on) {
      launch();
      return;
    }
    hipEvent_t a = prof.take(), b = prof.take();
    DEREK_HIP(hipEventRecord(a, s));
    launch();
    DEREK_HIP(hipEventRecord(b, s));
    prof.pending.push_back({cls, a, b, bytes, flops});
    if (prof.pending.size() > 4096) prof.resolve();
  }

  ~derek_ctx();
  template <typename R> const Axis<R>* axes() const;
  template <typename R> const R* vlocal() const;

  template <typename R> void fft3d(cx_t<R>* g, int howmany, int dir);
  template <typename R>
  void apply_h(int k, const cx_t<R>* psi, int64_t ldp, cx_t<R>* hpsi, int64_t ldh, int nbands);
  template <typename R> void density(int k, const cx_t<R>* psi, int64_t ldp, int nbands, const double* occw, bool keep);
  void hartree(const double* rho, double* vh_out, double* eh);
  double lda(const double* rho, double* vxc_out);
  double dot(const double* a, const double* b, int64_t n);
  void kerker(const double* f, double* out);
  template <typename R> void overlap(const cx_t<R>* X, int64_t ldx, const cx_t<R>* Y, int64_t ldy, int m, int n,
                                     int64_t npw, double2* C, int ldc, bool herm);
  template <typename R> void rotate(const cx_t<R>* X, int64_t ldx, const double2* U, int ldu, int m, int n,
                                    int64_t npw, cx_t<R>* Z, int64_t ldz);
  template <typename R> void band_stats(int k, const cx_t<R>* X, const cx_t<R>* HX, const double* eps_h, double* ekin,
                                        double* rn);
  template <typename R> int davidson(int k, int n_iter, double res_tol = 0.0);
  template <typename R> void orthonormalise(int k);
  void mix();
  void require_basis(int k) const;
  void sync() { DEREK_HIP(hipStreamSynchronize(s)); }
};

template <> const Axis<double>* derek_ctx::axes<double>() const { return ax64; }
template <> const Axis<float>* derek_ctx::axes<float>() const { return ax32; }
template <> const double* derek_ctx::vlocal<double>() const { return vloc; }
template <> double2* derek_ctx::sticks<double>() { return S64; }
template <> float2* derek_ctx::sticks<float>() { return S32; }
template <> double2* derek_ctx::planes<double>() { return P64; }
template <> float2* derek_ctx::planes<float>() { return P32; }

// (Re)allocate the stick and plane buffers when a k-point needs more room.
void derek_ctx::ensure_pw_buffers() {
  int64_t st = 0, pl = 0;
  for (const KState& k : ks)
    if (k.set) {
      st = std::max<int64_t>(st, (int64_t)k.maps.nst * dims[2]);
      pl = std::max<int64_t>(pl, (int64_t)k.maps.nax * dims[1] * dims[2]);
    }
  st *= batch;
  pl *= batch;
  if (st > st_elems || pl > pl_elems) {
    for (void* p : {(void*)S64, (void*)P64, (void*)S32, (void*)P32}) dfree(p);
    S64 = dmalloc<double2>(st);
    P64 = dmalloc<double2>(pl);
    S32 = dmalloc<float2>(st);
    P32 = dmalloc<float2>(pl);
    st_elems = st;
    pl_elems = pl;
  }
}
template <> const float* derek_ctx::vlocal<float>() const { return vloc32; }

derek_ctx::~derek_ctx() {
  if (s) (void)hipStreamSynchronize(s);
  for (int a = 0; a < 3; ++a) {
    dfree(tw64[a]);
    dfree(tw32[a]);
  }
  for (auto& k : ks) {
    dfree(k.idx);
    dfree(k.kin);
    dfree(k.mapbuf);
    dfree(k.Y);
    dfree(k.HY);
    dfree(k.Y2);
    dfree(k.HY2);
  }
  for (void* p : {(void*)S64, (void*)P64, (void*)S32, (void*)P32}) dfree(p);
  for (void* p : {(void*)gsq, (void*)vext, (void*)vloc, (void*)vloc32, (void*)cgrid,
                  (void*)rho_in, (void*)rho_out, (void*)vh, (void*)vxc, (void*)vhxc, (void*)tmp1, (void*)tmp2,
                  (void*)d_eps, (void*)d_ekin, (void*)d_occw, (void*)d_act, (void*)d_S, (void*)d_H, (void*)d_U})
    dfree(p);
  for (void* p : {(void*)d_dsc, (void*)d_w, (void*)d_hwork, (void*)d_sact, (void*)d_na, (void*)d_Lb, (void*)d_Lc,
                  (void*)d_Hc, (void*)d_Linv, (void*)d_LinvH, (void*)d_T, (void*)d_A, (void*)d_Zs, (void*)d_Ca})
    dfree(p);
  for (double* p : h_rin) dfree(p);
  for (double* p : h_res) dfree(p);
  tA.release();
  tB.release();
  tC.release();
  tD.release();
  ovwork.release();
  partial.release();
  if (s) (void)hipStreamDestroy(s);
}

void derek_ctx::require_basis(int k) const {
  if (k < 0 || k >= nk) throw ArgError("k-point index " + std::to_string(k) + " out of range");
  if (!ks[k].set) throw StateError("basis of k-point " + std::to_string(k) + " not set");
}

// ------------------------------------------------------------------ grid ops
template <typename R> void derek_ctx::fft3d(cx_t<R>* g, int howmany, int dir) {
  const int sign = dir == 1 ? -1 : +1;
  const Axis<R>* ax = axes<R>();
  const double pass = 2.0 * (double)howmany * N * sizeof(cx_t<R>);
  timed(KC_FFT_Z, pass, 0, [&] { fft_z<R>(g, (int64_t)howmany * dims[0] * dims[1], ax[2], sign, s); });
  timed(KC_FFT_Y, pass, 0, [&] { fft_strided<R>(g, (int64_t)howmany * dims[0], dims[2], ax[1], sign, s); });
  timed(KC_FFT_X, pass, 0, [&] { fft_strided<R>(g, howmany, (int64_t)dims[1] * dims[2], ax[0], sign, s); });
}

// Algorithmic bytes per launch (compulsory reads + writes of each kernel's
// own inputs and outputs) are passed to the profiler with every launch.
template <typename R>
void derek_ctx::apply_h(int k, const cx_t<R>* psi, int64_t ldp, cx_t<R>* hpsi, int64_t ldh, int nbands) {
  const KState& K = ks[k];
  const PwMaps& m = K.maps;
  const Axis<R>* ax = axes<R>();
  cx_t<R>* S = sticks<R>();
  cx_t<R>* P = planes<R>();
  const R* v = vlocal<R>();
  const double sc = sizeof(cx_t<R>);
  const int64_t nyz = (int64_t)dims[1] * dims[2];
  for (int b0 = 0; b0 < nbands; b0 += batch) {
    const int nbb = std::min(batch, nbands - b0);
    const double pw_b = (double)nbb * K.npw * sc;
    const double st_b = (double)nbb * m.nst * dims[2] * sc;
    const double pl_b = (double)nbb * m.nax * nyz * sc;
    timed(KC_STICKS_Z_INV, pw_b + 12.0 * K.npw + st_b, 0,
          [&] { sticks_z_inv<R>(psi + b0 * ldp, ldp, nbb, m, ax, dims[2], S, s); });
    timed(KC_PLANES_Y_INV, st_b + pl_b, 0, [&] { planes_y<R>(S, P, nbb, m, ax, dims[1], dims[2], +1, nullptr, s); });
    timed(KC_PLANES_X_VMUL, 2 * pl_b + (double)N * sizeof(R), 0,
          [&] { planes_x_vmul<R>(P, nbb, m, ax, nyz, v, (R)(1.0 / (double)N), s); });
    timed(KC_PLANES_Y_FWD, st_b + pl_b, 0, [&] { planes_y<R>(nullptr, P, nbb, m, ax, dims[1], dims[2], -1, S, s); });
    timed(KC_STICKS_Z_FWD, st_b + 2 * pw_b + 20.0 * K.npw, 0, [&] {
      sticks_z_fwd<R>(S, nbb, m, ax, dims[2], psi + b0 * ldp, ldp, K.kin, hpsi + b0 * ldh, ldh, s);
    });
  }
}

template <typename R>
void derek_ctx::density(int k, const cx_t<R>* psi, int64_t ldp, int nbands, const double* occw, bool keep) {
  const KState& K = ks[k];
  const PwMaps& m = K.maps;
  const Axis<R>* ax = axes<R>();
  cx_t<R>* S = sticks<R>();
  cx_t<R>* P = planes<R>();
  const double sc = sizeof(cx_t<R>);
  const int64_t nyz = (int64_t)dims[1] * dims[2];
  for (int b0 = 0; b0 < nbands; b0 += batch) {
    const int nbb = std::min(batch, nbands - b0);
    const double pw_b = (double)nbb * K.npw * sc;
    const double st_b = (double)nbb * m.nst * dims[2] * sc;
    const double pl_b = (double)nbb * m.nax * nyz * sc;
    const int kp = (keep || b0 > 0) ? 1 : 0;
    timed(KC_STICKS_Z_INV, pw_b + 12.0 * K.npw + st_b, 0,
          [&] { sticks_z_inv<R>(psi + b0 * ldp, ldp, nbb, m, ax, dims[2], S, s); });
    timed(KC_PLANES_Y_INV, st_b + pl_b, 0, [&] { planes_y<R>(S, P, nbb, m, ax, dims[1], dims[2], +1, nullptr, s); });
    timed(KC_PLANES_X_DENSITY, pl_b + (1 + kp) * 8.0 * N, 0,
          [&] { planes_x_density<R>(P, nbb, m, ax, nyz, occw + b0, rho_out_target(), kp, s); });
  }
}

double derek_ctx::dot(const double* a, const double* b, int64_t n) {
  const int nbk = dot_nblocks(n);
  double* p = partial.get<double>(nbk);
  dot_partials(a, b, p, n, s);
  std::vector<double> h(nbk);
  DEREK_HIP(hipMemcpyAsync(h.data(), p, nbk * sizeof(double), hipMemcpyDeviceToHost, s));
  sync();
  double t = 0.0;
  for (double x : h) t += x;
  return t;
}

void derek_ctx::hartree(const double* rho, double* vh_out, double* eh) {
  if (!gsq_set) throw StateError("hartree: |G|^2 not set (derek_hip_set_gsq)");
  real_to_cx(rho, cgrid, N, s);
  fft3d<double>(cgrid, 1, +1);
  gspace_scale(cgrid, gsq, 0.0, 1.0 / (double)N, 0, N, s);
  fft3d<double>(cgrid, 1, -1);
  cx_real(cgrid, vh_out, 1.0, N, s);
  if (eh) *eh = 0.5 * volume / (double)N * dot(vh_out, rho, N);
}

double derek_ctx::lda(const double* rho, double* vxc_out) {
  const int np = lda_partials(N);
  double* p = partial.get<double>(np);
  lda_xc(rho, vxc_out, p, N, s);
  std::vector<double> h(np);
  DEREK_HIP(hipMemcpyAsync(h.data(), p, np * sizeof(double), hipMemcpyDeviceToHost, s));
  sync();
  double t = 0.0;
  for (double x : h) t += x;
  return volume / (double)N * t;
}

void derek_ctx::kerker(const double* f, double* out) {
  real_to_cx(f, cgrid, N, s);
  fft3d<double>(cgrid, 1, +1);
  gspace_scale(cgrid, gsq, P.mix_q0 * P.mix_q0, 1.0 / (double)N, 1, N, s);
  fft3d<double>(cgrid, 1, -1);
  cx_real(cgrid, out, 1.0, N, s);
}

template <>
void derek_ctx::overlap<double>(const double2* X, int64_t ldx, const double2* Y, int64_t ldy, int m, int n,
                                int64_t npw, double2* C, int ldc, bool herm) {
  double2* w = ovwork.get<double2>(overlap_work_elems(m, n, npw));
  timed(KC_OVERLAP, (double)(m + n) * npw * 16.0, 8.0 * npw * m * n,
        [&] { overlap_f64(X, ldx, Y, ldy, m, n, npw, C, ldc, w, herm, s); });
}
template <>
void derek_ctx::rotate<double>(const double2* X, int64_t ldx, const double2* U, int ldu, int m, int n, int64_t npw,
                               double2* Z, int64_t ldz) {
  timed(KC_ROTATE, (double)(m + n) * npw * 16.0, 8.0 * npw * m * n,
        [&] { rotate_f64(X, ldx, U, ldu, m, n, npw, Z, ldz, s); });
}

template <typename R>
void derek_ctx::band_stats(int k, const cx_t<R>* X, const cx_t<R>* HX, const double* eps_h, double* ekin,
                           double* rn) {
  const KState& K = ks[k];
  const int nchunk = (int)std::max<int64_t>(1, std::min<int64_t>(64, ceil_div(K.npw, 4096)));
  DEREK_HIP(hipMemcpyAsync(d_eps, eps_h, nb * sizeof(double), hipMemcpyHostToDevice, s));
  double* p = partial.get<double>((size_t)nb * nchunk * 2);
  band_sums<R>(X, HX, K.ld, d_eps, K.kin, nb, K.npw, p, nchunk, s);
  std::vector<double> h((size_t)nb * nchunk * 2);
  DEREK_HIP(hipMemcpyAsync(h.data(), p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  sync();
  for (int b = 0; b < nb; ++b) {
    double sk = 0.0, sr = 0.0;
    for (int c = 0; c < nchunk; ++c) {
      sk += h[((size_t)b * nchunk + c) * 2 + 0];
      sr += h[((size_t)b * nchunk + c) * 2 + 1];
    }
    ekin[b] = sk;
    rn[b] = std::sqrt(sr);
  }
}

// Davidson with restart: Rayleigh-Ritz on X, then n_iter steps on [X, K R].
template <typename R> int derek_ctx::davidson(int k, int n_iter, double res_tol) {
  KState& K = ks[k];
  const int64_t ld = K.ld, npw = K.npw;
  auto Y = reinterpret_cast<cx_t<R>*>(K.Y);
  auto HY = reinterpret_cast<cx_t<R>*>(K.HY);
  auto Y2 = reinterpret_cast<cx_t<R>*>(K.Y2);
  auto HY2 = reinterpret_cast<cx_t<R>*>(K.HY2);
  double* eps = eig.data() + (size_t)k * nb;
  double* rn = res.data() + (size_t)k * nb;
  std::vector<double> ekin(nb);
  apply_h<R>(k, Y, ld, HY, ld, nb);
  // Rayleigh-Ritz (GPU FP64 eigensolver)
  overlap<R>(Y, ld, HY, ld, nb, nb, npw, d_S, nb, false);
  {
    const auto t0 = std::chrono::steady_clock::now();
    rr_gpu(nb, eps);
    t_dense += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  rotate<R>(Y, ld, d_U, nb, nb, nb, npw, Y2, ld);
  rotate<R>(HY, ld, d_U, nb, nb, nb, npw, HY2, ld);
  std::swap(Y, Y2);
  std::swap(HY, HY2);
  band_stats<R>(k, Y, HY, eps, ekin.data(), rn);
  int it = 0;
  for (; it < n_iter; ++it) {
    double rmax = 0.0;
    for (int b = 0; b < nb; ++b) rmax = std::max(rmax, rn[b]);
    if (rmax < res_tol) break;
    std::vector<int32_t> act;
    for (int b = 0; b < nb; ++b)
      if (rn[b] > 1e-10) act.push_back(b);
    const int nact = (int)act.size();
    if (nact == 0) break;
    DEREK_HIP(hipMemcpyAsync(d_act, act.data(), nact * sizeof(int32_t), hipMemcpyHostToDevice, s));
    DEREK_HIP(hipMemcpyAsync(d_ekin, ekin.data(), nb * sizeof(double), hipMemcpyHostToDevice, s));
    DEREK_HIP(hipMemcpyAsync(d_eps, eps, nb * sizeof(double), hipMemcpyHostToDevice, s));
    tpa_residual<R>(Y, HY, ld, d_eps, d_ekin, d_act, nact, K.kin, npw, Y + (int64_t)nb * ld, ld, s);
    apply_h<R>(k, Y + (int64_t)nb * ld, ld, HY + (int64_t)nb * ld, ld, nact);
    const int m = nb + nact;
    overlap<R>(Y, ld, Y, ld, m, m, npw, d_S, m, true);
    overlap<R>(Y, ld, HY, ld, m, m, npw, d_H, m, false);
    std::vector<double> w(nb);
    const auto td0 = std::chrono::steady_clock::now();
    const bool ok = gen_gpu(m, nb, w.data());  // C (m x nb) in d_U
    t_dense += std::chrono::duration<double>(std::chrono::steady_clock::now() - td0).count();
    if (!ok) break;
    rotate<R>(Y, ld, d_U, m, m, nb, npw, Y2, ld);
    rotate<R>(HY, ld, d_U, m, m, nb, npw, HY2, ld);
    std::swap(Y, Y2);
    std::swap(HY, HY2);
    for (int b = 0; b < nb; ++b) eps[b] = w[b];
    band_stats<R>(k, Y, HY, eps, ekin.data(), rn);
  }
  K.Y = Y;
  K.HY = HY;
  K.Y2 = Y2;
  K.HY2 = HY2;
  return it;
}

void derek_ctx::ensure_sub(int m) {
  if (m <= sub_cap) return;
  dfree(d_S);
  dfree(d_H);
  dfree(d_U);
  d_S = dmalloc<double2>((size_t)m * m);
  d_H = dmalloc<double2>((size_t)m * m);
  d_U = dmalloc<double2>((size_t)m * m);
  sub_cap = m;
}

void derek_ctx::ensure_dense(int m, int nev) {
  if (m <= dense_cap) return;
  for (void* p : {(void*)d_dsc, (void*)d_w, (void*)d_hwork, (void*)d_sact, (void*)d_na, (void*)d_Lb, (void*)d_Lc,
                  (void*)d_Hc, (void*)d_Linv, (void*)d_LinvH, (void*)d_T, (void*)d_A, (void*)d_Zs, (void*)d_Ca})
    dfree(p);
  const size_t mm = (size_t)m * m;
  d_dsc = dmalloc<double>(m);
  d_w = dmalloc<double>(m);
  d_hwork = dmalloc<double>(heev_work_doubles(m, m));
  d_sact = dmalloc<int>(m);
  d_na = dmalloc<int>(1);
  for (double2** q : {&d_Lb, &d_Lc, &d_Hc, &d_Linv, &d_LinvH, &d_T, &d_A, &d_Zs, &d_Ca}) *q = dmalloc<double2>(mm);
  dense_cap = m;
  (void)nev;
}

// Eigenpairs of the Hermitian n x n matrix in d_S: vectors to d_U, values to w_host.
void derek_ctx::rr_gpu(int n, double* w_host) {
  ensure_dense(n, n);
  hermitize_gpu(d_S, n, s);
  heev_gpu(d_S, n, 0, n, n, d_w, 0, d_U, n, 0, d_hwork, 0, 1, s);
  DEREK_HIP(hipMemcpyAsync(w_host, d_w, n * sizeof(double), hipMemcpyDeviceToHost, s));
  sync();
}

// Lowest nev solutions of d_H c = w d_S c (m x m): C (m x nev) to d_U, values to
// w_host.  Columns numerically dependent on earlier ones are dropped first.
bool derek_ctx::gen_gpu(int m, int nev, double* w_host) {
  ensure_dense(m, nev);
  hermitize_gpu(d_S, m, s);
  hermitize_gpu(d_H, m, s);
  hegv_gpu_prepare(d_S, m, d_dsc, s);
  chol_drop_gpu(d_S, m, 1e-8, d_T, d_Lb, d_sact, d_na, s);
  int na = 0;
  DEREK_HIP(hipMemcpyAsync(&na, d_na, sizeof(int), hipMemcpyDeviceToHost, s));
  sync();
  if (na < nev) return false;
  chol_compact_gpu(d_Lb, m, d_sact, d_na, d_H, m, d_dsc, d_Lc, d_Hc, s);
  trinv_gpu(d_Lc, d_na, na, d_Linv, s);
  rotate_f64(d_Linv, na, d_Hc, na, na, na, na, d_T, na, s);  // T = L^-1 H
  conj_transpose_gpu(d_Linv, d_na, na, d_LinvH, s);
  rotate_f64(d_T, na, d_LinvH, na, na, na, na, d_A, na, s);  // A = L^-1 H L^-H
  hermitize_gpu(d_A, na, s);
  heev_gpu(d_A, na, 0, na, nev, d_w, 0, d_Zs, na, 0, d_hwork, 0, 1, s);
  rotate_f64(d_LinvH, na, d_Zs, na, na, nev, na, d_Ca, na, s);  // C_active = L^-H Z
  expand_rows_gpu(d_Ca, d_na, d_sact, d_dsc, m, nev, d_U, s);
  DEREK_HIP(hipMemcpyAsync(w_host, d_w, nev * sizeof(double), hipMemcpyDeviceToHost, s));
  sync();
  return true;
}

template <typename R> void derek_ctx::orthonormalise(int k) {
  KState& K = ks[k];
  auto Y = reinterpret_cast<cx_t<R>*>(K.Y);
  auto Y2 = reinterpret_cast<cx_t<R>*>(K.Y2);
  ensure_dense(nb, nb);
  overlap<R>(Y, K.ld, Y, K.ld, nb, nb, K.npw, d_S, nb, true);
  // S' = D S D = L L^H (unit diagonal), X <- X D L^-H
  hegv_gpu_prepare(d_S, nb, d_dsc, s);
  chol_drop_gpu(d_S, nb, 0.0, d_T, d_Lb, d_sact, d_na, s);
  int na = 0;
  DEREK_HIP(hipMemcpyAsync(&na, d_na, sizeof(int), hipMemcpyDeviceToHost, s));
  sync();
  if (na < nb) throw NumericError("initial wavefunctions are linearly dependent");
  chol_compact_gpu(d_Lb, nb, d_sact, d_na, d_S, nb, d_dsc, d_Lc, d_Hc, s);
  trinv_gpu(d_Lc, d_na, nb, d_Linv, s);
  conj_transpose_gpu(d_Linv, d_na, nb, d_LinvH, s);
  expand_rows_gpu(d_LinvH, d_na, d_sact, d_dsc, nb, nb, d_U, s);
  rotate<R>(Y, K.ld, d_U, nb, nb, nb, K.npw, Y2, K.ld);
  std::swap(K.Y, K.Y2);
  sync();
}

// Pulay (DIIS) mixing of rho with a Kerker-preconditioned step.
static std::vector<double> solve_pulay(const std::vector<std::vector<double>>& A, const std::vector<int>& ord) {
  const int n = (int)ord.size();
  double smax = 0.0;
  for (int i = 0; i < n; ++i) smax = std::max(smax, A[ord[i]][ord[i]]);
  std::vector<double> c(n, 1.0 / n);
  if (!(smax > 0.0)) return c;
  std::vector<double> M((size_t)n * (n + 1));
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) M[i * (n + 1) + j] = A[ord[i]][ord[j]] / smax + (i == j ? 1e-12 : 0.0);
    M[i * (n + 1) + n] = 1.0;
  }
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int r = col + 1; r < n; ++r)
      if (std::fabs(M[r * (n + 1) + col]) > std::fabs(M[piv * (n + 1) + col])) piv = r;
    if (piv != col)
      for (int j = 0; j <= n; ++j) std::swap(M[col * (n + 1) + j], M[piv * (n + 1) + j]);
    const double pv = M[col * (n + 1) + col];
    if (pv == 0.0) return std::vector<double>(n, 1.0 / n);
    for (int r = col + 1; r < n; ++r) {
      const double f = M[r * (n + 1) + col] / pv;
      for (int j = col; j <= n; ++j) M[r * (n + 1) + j] -= f * M[col * (n + 1) + j];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double v = M[i * (n + 1) + n];
    for (int j = i + 1; j < n; ++j) v -= M[i * (n + 1) + j] * c[j];
    c[i] = v / M[i * (n + 1) + i];
  }
  double sum = 0.0;
  for (double x : c) sum += x;
  for (double& x : c) x /= sum;
  return c;
}

void derek_ctx::mix() {
  const int m = P.history;
  int slot;
  if ((int)order.size() < m) {
    slot = (int)order.size();
  } else {
    slot = order.front();
    order.erase(order.begin());
  }
  DEREK_HIP(hipMemcpyAsync(h_rin[slot], rho_in, N * sizeof(double), hipMemcpyDeviceToDevice, s));
  sub(h_res[slot], rho_out, rho_in, N, s);
  order.push_back(slot);
  for (int j : order) {
    const double v = dot(h_res[slot], h_res[j], N);
    A[slot][j] = v;
    A[j][slot] = v;
  }
  std::vector<double> c = solve_pulay(A, order);
  std::vector<const double*> xr, xf;
  for (int j : order) {
    xr.push_back(h_rin[j]);
    xf.push_back(h_res[j]);
  }
  lincomb(rho_in, xr.data(), c.data(), (int)order.size(), N, s);
  lincomb(tmp1, xf.data(), c.data(), (int)order.size(), N, s);
  kerker(tmp1, tmp2);
  axpy(rho_in, tmp2, P.mix_beta, N, s);
}

static void occupations(const std::vector<double>& eig, const std::vector<KState>& ks, int nk, int nb, double nelec,
                        double kT, std::vector<double>& occ, double& mu, double& mts) {
  occ.assign((size_t)nk * nb, 0.0);
  mts = 0.0;
  if (kT <= 0.0) {
    const int nocc = (int)std::lround(nelec / 2.0);
    if (nocc > nb) throw ArgError("more occupied bands than bands");
    mu = -1e300;
    for (int k = 0; k < nk; ++k) {
      for (int b = 0; b < nocc; ++b) occ[(size_t)k * nb + b] = 2.0;
      if (nocc > 0) mu = std::max(mu, eig[(size_t)k * nb + nocc - 1]);
    }
    return;
  }
  double lo = 1e300, hi = -1e300;
  for (double e : eig) {
    lo = std::min(lo, e);
    hi = std::max(hi, e);
  }
  lo -= 50 * kT;
  hi += 50 * kT;
  auto count = [&](double m) {
    double t = 0.0;
    for (int k = 0; k < nk; ++k)
      for (int b = 0; b < nb; ++b) {
        const double x = std::min(200.0, std::max(-200.0, (eig[(size_t)k * nb + b] - m) / kT));
        t += ks[k].weight * 2.0 / (1.0 + std::exp(x));
      }
    return t;
  };
  for (int i = 0; i < 200; ++i) {
    const double m = 0.5 * (lo + hi);
    if (count(m) < nelec) lo = m;
    else hi = m;
  }
  mu = 0.5 * (lo + hi);
  double ent = 0.0;
  for (int k = 0; k < nk; ++k)
    for (int b = 0; b < nb; ++b) {
      const double x = std::min(200.0, std::max(-200.0, (eig[(size_t)k * nb + b] - mu) / kT));
      const double f = 1.0 / (1.0 + std::exp(x));
      occ[(size_t)k * nb + b] = 2.0 * f;
      const double fs = std::min(1.0, std::max(1e-300, f)), gs = std::min(1.0, std::max(1e-300, 1.0 - f));
      ent -= ks[k].weight * 2.0 * (f * std::log(fs) + (1.0 - f) * std::log(gs));
    }
  mts = -kT * ent;
}

template <typename R> static R* make_twiddles(int n) {
  std::vector<cx_t<R>> h(n);
  const long double pi = 3.141592653589793238462643383279502884197L;
  for (int t = 0; t < n; ++t) {
    const long double a = 2.0L * pi * (long double)t / (long double)n;
    h[t].x = (R)std::cos(a);
    h[t].y = (R)(-std::sin(a));
  }
  cx_t<R>* d = dmalloc<cx_t<R>>(n);
  DEREK_HIP(hipMemcpy(d, h.data(), n * sizeof(cx_t<R>), hipMemcpyHostToDevice));
  return reinterpret_cast<R*>(d);
}

template <typename R> static void alloc_k(derek_ctx* c, KState& K) {
  const size_t n = (size_t)2 * c->nb * K.ld;
  for (void** p : {&K.Y, &K.HY, &K.Y2, &K.HY2}) {
    dfree(*p);
    *p = dmalloc<cx_t<R>>(n);
  }
  zero<R>(reinterpret_cast<cx_t<R>*>(K.Y), n, c->s);
  zero<R>(reinterpret_cast<cx_t<R>*>(K.HY), n, c->s);
  zero<R>(reinterpret_cast<cx_t<R>*>(K.Y2), n, c->s);
  zero<R>(reinterpret_cast<cx_t<R>*>(K.HY2), n, c->s);
}

// ================================================================== C-ABI
extern "C" {

int32_t derek_hip_version(void) { return 1; }

int32_t derek_hip_last_error(char* buf, int64_t len) {
  if (!buf || len <= 0) return DEREK_ERR_ARG;
  const size_t n = std::min<size_t>((size_t)len - 1, g_last_error.size());
  std::memcpy(buf, g_last_error.data(), n);
  buf[n] = 0;
  return DEREK_OK;
}

int32_t derek_basis_grid_dims(const double cell[9], double ecut, double g, int32_t sd[3], int32_t fd[3]) {
  return guard([&] {
    DEREK_REQUIRE(cell && sd && fd, "null argument");
    DEREK_REQUIRE(ecut > 0.0 && g > 0.0, "cut-off and g_fine_scale must be positive");
    int a[3], b[3];
    basis::grid_dims(cell, ecut, g, a, b);
    for (int i = 0; i < 3; ++i) {
      sd[i] = a[i];
      fd[i] = b[i];
    }
  });
}

int32_t derek_basis_inv_cell(const double cell[9], double inv[9]) {
  return guard([&] {
    DEREK_REQUIRE(cell && inv, "null argument");
    basis::inv_cell(cell, inv);
  });
}

static void check_dims(const int32_t d[3]) {
  DEREK_REQUIRE(d != nullptr, "null dims");
  for (int i = 0; i < 3; ++i) DEREK_REQUIRE(d[i] > 0 && d[i] <= 4096, "grid dimension out of range");
}

int32_t derek_basis_g_map(const double inv[9], const int32_t dims[3], double* g, int32_t* m) {
  return guard([&] {
    DEREK_REQUIRE(inv, "null inv_cell");
    check_dims(dims);
    const int d[3] = {dims[0], dims[1], dims[2]};
    basis::g_map(inv, d, g, m);
  });
}

int32_t derek_basis_real_map(const double cell[9], const int32_t dims[3], double* r, double* frac) {
  return guard([&] {
    DEREK_REQUIRE(cell, "null cell");
    check_dims(dims);
    const int d[3] = {dims[0], dims[1], dims[2]};
    basis::real_map(cell, d, r, frac);
  });
}

int32_t derek_basis_kpoints(const int32_t mp[3], double* kfrac, int32_t* nk) {
  return guard([&] {
    DEREK_REQUIRE(mp && nk, "null argument");
    DEREK_REQUIRE(mp[0] > 0 && mp[1] > 0 && mp[2] > 0, "invalid number of kpoints");
    const int m[3] = {mp[0], mp[1], mp[2]};
    *nk = basis::kpoints(m, kfrac);
  });
}

int32_t derek_basis_sphere(const double inv[9], const int32_t dims[3], const double kfrac[3], double ecut,
                           int64_t cap, int64_t* npw, int32_t* fft_index, double* kin) {
  return guard([&] {
    DEREK_REQUIRE(inv && kfrac && npw, "null argument");
    check_dims(dims);
    const int d[3] = {dims[0], dims[1], dims[2]};
    *npw = basis::sphere(inv, d, kfrac, ecut, cap, fft_index, kin, nullptr);
  });
}

int32_t derek_basis_gsq(const double inv[9], const int32_t dims[3], double* gsq) {
  return guard([&] {
    DEREK_REQUIRE(inv && gsq, "null argument");
    check_dims(dims);
    const int d[3] = {dims[0], dims[1], dims[2]};
    basis::gsq(inv, d, gsq);
  });
}

int32_t derek_gaussian_pot(const double cell[9], const int32_t dims[3], int32_t nwells, const double* centres,
                           const double* heights, const double* widths, double cos_amp, double* v) {
  return guard([&] {
    DEREK_REQUIRE(cell && v, "null argument");
    DEREK_REQUIRE(nwells >= 0, "negative well count");
    DEREK_REQUIRE(nwells == 0 || (centres && heights && widths), "null well arrays");
    check_dims(dims);
    const int d[3] = {dims[0], dims[1], dims[2]};
    basis::gaussian_pot(cell, d, nwells, centres, heights, widths, cos_amp, v);
  });
}


int32_t derek_hip_init(derek_ctx** out, const derek_dims* d) {
  return guard([&] {
    DEREK_REQUIRE(out && d, "null argument");
    *out = nullptr;
    check_dims(d->ngrid);
    DEREK_REQUIRE(d->nk >= 1, "nk must be >= 1");
    DEREK_REQUIRE(d->nbands >= 1, "nbands must be >= 1");
    DEREK_REQUIRE(d->precision == DEREK_FP64 || d->precision == DEREK_MIXED, "unknown precision mode");
    DEREK_REQUIRE(d->volume > 0.0, "volume must be positive");
    DEREK_REQUIRE(d->max_npw >= 1, "max_npw must be >= 1");
    std::unique_ptr<derek_ctx> c(new derek_ctx());
    for (int a = 0; a < 3; ++a) {
      c->dims[a] = d->ngrid[a];
      c->plan[a] = make_plan(d->ngrid[a]);
      DEREK_REQUIRE(c->plan[a].ns >= 0, "FFT length " + std::to_string(d->ngrid[a]) +
                                            " has a prime factor outside {2,3,5,7,11,13}");
    }
    c->N = (int64_t)d->ngrid[0] * d->ngrid[1] * d->ngrid[2];
    c->nk = d->nk;
    c->nb = d->nbands;
    c->prec = d->precision;
    c->volume = d->volume;
    c->max_npw = d->max_npw;
    DEREK_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    for (int a = 0; a < 3; ++a) {
      c->tw64[a] = reinterpret_cast<double2*>(make_twiddles<double>(c->dims[a]));
      c->tw32[a] = reinterpret_cast<float2*>(make_twiddles<float>(c->dims[a]));
      c->ax64[a] = Axis<double>{c->dims[a], c->plan[a], c->tw64[a]};
      c->ax32[a] = Axis<float>{c->dims[a], c->plan[a], c->tw32[a]};
    }
    c->ks.resize(c->nk);
    c->gsq = dmalloc<double>(c->N);
    c->vext = dmalloc<double>(c->N);
    c->vloc = dmalloc<double>(c->N);
    c->vloc32 = dmalloc<float>(c->N);
    c->cgrid = dmalloc<double2>(c->N);
    // bands per launch of the pruned FFT pipeline (stick + plane buffers are
    // allocated for this many bands when the basis is set)
    c->batch = 16;
    if (const char* b = std::getenv("DEREK_BAND_BATCH")) c->batch = std::max(1, std::atoi(b));
    *out = c.release();
  });
}

int32_t derek_hip_destroy(derek_ctx* c) {
  return guard([&] { delete c; });
}

int32_t derek_hip_set_basis(derek_ctx* c, int32_t k, int64_t npw, const int32_t* idx, const double* kin,
                            double weight) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    DEREK_REQUIRE(k >= 0 && k < c->nk, "k-point index out of range");
    DEREK_REQUIRE(npw >= 1 && npw <= c->max_npw, "npw must be in 1..max_npw");
    DEREK_REQUIRE(npw <= c->N, "more plane waves than grid points");
    DEREK_REQUIRE(idx && kin, "null basis arrays");
    for (int64_t g = 0; g < npw; ++g)
      DEREK_REQUIRE(idx[g] >= 0 && idx[g] < c->N, "fft_index entry outside the grid");
    KState& K = c->ks[k];
    dfree(K.idx);
    dfree(K.kin);
    K.npw = npw;
    K.ld = round_up(npw, 64);
    K.idx = dmalloc<int32_t>(npw);
    K.kin = dmalloc<double>(npw);
    DEREK_HIP(hipMemcpy(K.idx, idx, npw * sizeof(int32_t), hipMemcpyHostToDevice));
    DEREK_HIP(hipMemcpy(K.kin, kin, npw * sizeof(double), hipMemcpyHostToDevice));
    // sphere maps for the pruned FFT pipeline
    const PwMapsHost h = build_pw_maps(idx, npw, c->dims[0], c->dims[1], c->dims[2]);
    dfree(K.mapbuf);
    const size_t nmap = 3 * (size_t)npw + 2 * (size_t)h.nst + 1 + 2 * (size_t)h.nax + 1;
    std::vector<int32_t> flat;
    flat.reserve(nmap);
    for (const std::vector<int32_t>* v : {&h.perm, &h.pw_z, &h.pw_st, &h.st_pw, &h.st_y, &h.pl_st, &h.ax_x})
      flat.insert(flat.end(), v->begin(), v->end());
    K.mapbuf = dmalloc<int32_t>(flat.size());
    DEREK_HIP(hipMemcpy(K.mapbuf, flat.data(), flat.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    int32_t* q = K.mapbuf;
    K.maps.npw = h.npw;
    K.maps.nst = h.nst;
    K.maps.nax = h.nax;
    K.maps.perm = q;
    q += npw;
    K.maps.pw_z = q;
    q += npw;
    K.maps.pw_st = q;
    q += npw;
    K.maps.st_pw = q;
    q += h.nst + 1;
    K.maps.st_y = q;
    q += h.nst;
    K.maps.pl_st = q;
    q += h.nax + 1;
    K.maps.ax_x = q;
    K.weight = weight;
    K.set = true;
    c->ensure_pw_buffers();
  });
}

int32_t derek_hip_set_gsq(derek_ctx* c, const double* g) {
  return guard([&] {
    DEREK_REQUIRE(c && g, "null argument");
    DEREK_HIP(hipMemcpy(c->gsq, g, c->N * sizeof(double), hipMemcpyHostToDevice));
    c->gsq_set = true;
  });
}

int32_t derek_hip_set_vloc(derek_ctx* c, const double* v) {
  return guard([&] {
    DEREK_REQUIRE(c && v, "null argument");
    DEREK_HIP(hipMemcpy(c->vloc, v, c->N * sizeof(double), hipMemcpyHostToDevice));
    convert_real_f(c->vloc32, c->vloc, c->N, c->s);
    c->sync();
    c->vloc_set = true;
  });
}

int32_t derek_hip_set_vext(derek_ctx* c, const double* v) {
  return guard([&] {
    DEREK_REQUIRE(c && v, "null argument");
    DEREK_HIP(hipMemcpy(c->vext, v, c->N * sizeof(double), hipMemcpyHostToDevice));
    c->vext_set = true;
  });
}

int32_t derek_hip_fft3d(derek_ctx* c, double* grid, int32_t howmany, int32_t dir) {
  return guard([&] {
    DEREK_REQUIRE(c && grid, "null argument");
    DEREK_REQUIRE(dir == 1 || dir == -1, "direction must be +1 or -1");
    DEREK_REQUIRE(howmany >= 1, "howmany must be >= 1");
    const size_t n = (size_t)howmany * c->N;
    double2* d = c->tA.get<double2>(n);
    DEREK_HIP(hipMemcpyAsync(d, grid, n * sizeof(double2), hipMemcpyHostToDevice, c->s));
    if (c->prec == DEREK_FP64) {
      c->fft3d<double>(d, howmany, dir);
    } else {
      float2* f = c->tB.get<float2>(n);
      convert<float, double>(f, d, (int64_t)n, c->s);
      c->fft3d<float>(f, howmany, dir);
      convert<double, float>(d, f, (int64_t)n, c->s);
    }
    DEREK_HIP(hipMemcpyAsync(grid, d, n * sizeof(double2), hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_apply_h(derek_ctx* c, int32_t k, int32_t nb, const double* psi, double* hpsi) {
  return guard([&] {
    DEREK_REQUIRE(c && psi && hpsi, "null argument");
    c->require_basis(k);
    DEREK_REQUIRE(nb >= 1, "nb must be >= 1");
    if (!c->vloc_set) throw StateError("apply_h: local potential not set (derek_hip_set_vloc)");
    const KState& K = c->ks[k];
    const size_t n = (size_t)nb * K.npw;
    double2* dp = c->tA.get<double2>(n);
    double2* dh = c->tB.get<double2>(n);
    DEREK_HIP(hipMemcpyAsync(dp, psi, n * sizeof(double2), hipMemcpyHostToDevice, c->s));
    if (c->prec == DEREK_FP64) {
      c->apply_h<double>(k, dp, K.npw, dh, K.npw, nb);
    } else {
      float2* fp = c->tC.get<float2>(n);
      float2* fh = c->tD.get<float2>(n);
      convert<float, double>(fp, dp, (int64_t)n, c->s);
      c->apply_h<float>(k, fp, K.npw, fh, K.npw, nb);
      convert<double, float>(dh, fh, (int64_t)n, c->s);
    }
    DEREK_HIP(hipMemcpyAsync(hpsi, dh, n * sizeof(double2), hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_overlap(derek_ctx* c, int64_t npw, int32_t m, int32_t n, const double* X, const double* Y,
                          double* S) {
  return guard([&] {
    DEREK_REQUIRE(c && X && Y && S, "null argument");
    DEREK_REQUIRE(npw >= 1 && m >= 1 && n >= 1, "empty overlap");
    // X == Y: Hermitian product, only the upper tiles are computed
    const bool herm = (X == Y) && (m == n);
    double2* dx = c->tA.get<double2>((size_t)m * npw);
    double2* dy = herm ? dx : c->tB.get<double2>((size_t)n * npw);
    double2* ds = c->tC.get<double2>((size_t)m * n);
    DEREK_HIP(hipMemcpyAsync(dx, X, (size_t)m * npw * sizeof(double2), hipMemcpyHostToDevice, c->s));
    if (!herm) DEREK_HIP(hipMemcpyAsync(dy, Y, (size_t)n * npw * sizeof(double2), hipMemcpyHostToDevice, c->s));
    c->overlap<double>(dx, npw, dy, npw, m, n, npw, ds, m, herm);
    DEREK_HIP(hipMemcpyAsync(S, ds, (size_t)m * n * sizeof(double2), hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_rotate(derek_ctx* c, int64_t npw, int32_t m, int32_t n, const double* X, const double* U,
                         double* Z) {
  return guard([&] {
    DEREK_REQUIRE(c && X && U && Z, "null argument");
    DEREK_REQUIRE(npw >= 1 && m >= 1 && n >= 1, "empty rotation");
    double2* dx = c->tA.get<double2>((size_t)m * npw);
    double2* du = c->tB.get<double2>((size_t)m * n);
    double2* dz = c->tC.get<double2>((size_t)n * npw);
    DEREK_HIP(hipMemcpyAsync(dx, X, (size_t)m * npw * sizeof(double2), hipMemcpyHostToDevice, c->s));
    DEREK_HIP(hipMemcpyAsync(du, U, (size_t)m * n * sizeof(double2), hipMemcpyHostToDevice, c->s));
    c->rotate<double>(dx, npw, du, m, m, n, npw, dz, npw);
    DEREK_HIP(hipMemcpyAsync(Z, dz, (size_t)n * npw * sizeof(double2), hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_density(derek_ctx* c, int32_t k, int32_t nb, const double* psi, const double* occ, double* rho,
                          int32_t accumulate) {
  return guard([&] {
    DEREK_REQUIRE(c && psi && occ && rho, "null argument");
    c->require_basis(k);
    DEREK_REQUIRE(nb >= 1, "nb must be >= 1");
    const KState& K = c->ks[k];
    const size_t n = (size_t)nb * K.npw;
    double2* dp = c->tA.get<double2>(n);
    double* dr = c->tB.get<double>(c->N);
    double* dw = c->tD.get<double>(nb);
    std::vector<double> w(nb);
    for (int b = 0; b < nb; ++b) w[b] = K.weight * occ[b] / c->volume;
    DEREK_HIP(hipMemcpyAsync(dp, psi, n * sizeof(double2), hipMemcpyHostToDevice, c->s));
    DEREK_HIP(hipMemcpyAsync(dw, w.data(), nb * sizeof(double), hipMemcpyHostToDevice, c->s));
    if (accumulate) DEREK_HIP(hipMemcpyAsync(dr, rho, c->N * sizeof(double), hipMemcpyHostToDevice, c->s));
    c->set_rho_target(dr);
    if (c->prec == DEREK_FP64) {
      c->density<double>(k, dp, K.npw, nb, dw, accumulate != 0);
    } else {
      float2* fp = c->tC.get<float2>(n);
      convert<float, double>(fp, dp, (int64_t)n, c->s);
      c->density<float>(k, fp, K.npw, nb, dw, accumulate != 0);
    }
    DEREK_HIP(hipMemcpyAsync(rho, dr, c->N * sizeof(double), hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_lda(derek_ctx* c, int64_t np, const double* rho, double* exc, double* vxc) {
  return guard([&] {
    DEREK_REQUIRE(c && rho && vxc, "null argument");
    DEREK_REQUIRE(np >= 0, "negative point count");
    if (np == 0) return;
    double* dr = c->tA.get<double>(np);
    double* dv = c->tB.get<double>(np);
    DEREK_HIP(hipMemcpyAsync(dr, rho, np * sizeof(double), hipMemcpyHostToDevice, c->s));
    lda_xc(dr, dv, nullptr, np, c->s);
    DEREK_HIP(hipMemcpyAsync(vxc, dv, np * sizeof(double), hipMemcpyDeviceToHost, c->s));
    if (exc) {
      double* de = c->tC.get<double>(np);
      lda_eps(dr, de, np, c->s);
      DEREK_HIP(hipMemcpyAsync(exc, de, np * sizeof(double), hipMemcpyDeviceToHost, c->s));
    }
    c->sync();
  });
}

int32_t derek_hip_hartree(derek_ctx* c, const double* rho, double* vh, double* eh) {
  return guard([&] {
    DEREK_REQUIRE(c && rho && vh, "null argument");
    double* dr = c->tA.get<double>(c->N);
    double* dv = c->tB.get<double>(c->N);
    DEREK_HIP(hipMemcpyAsync(dr, rho, c->N * sizeof(double), hipMemcpyHostToDevice, c->s));
    double e = 0.0;
    c->hartree(dr, dv, &e);
    DEREK_HIP(hipMemcpyAsync(vh, dv, c->N * sizeof(double), hipMemcpyDeviceToHost, c->s));
    c->sync();
    if (eh) *eh = e;
  });
}

// ------------------------------------------------------------------ resident SCF

int32_t derek_hip_scf_init(derek_ctx* c, const derek_scf_params* p) {
  return guard([&] {
    DEREK_REQUIRE(c && p, "null argument");
    for (int k = 0; k < c->nk; ++k) c->require_basis(k);
    if (!c->gsq_set) throw StateError("scf_init: |G|^2 not set");
    if (!c->vext_set) throw StateError("scf_init: external potential not set");
    DEREK_REQUIRE(p->nelec > 0.0, "nelec must be positive");
    DEREK_REQUIRE(p->kT >= 0.0, "kT must be >= 0");
    DEREK_REQUIRE(p->history >= 1 && p->history <= 16, "Pulay history must be 1..16");
    DEREK_REQUIRE(p->n_david_first >= 0 && p->n_david >= 0, "Davidson step counts must be >= 0");
    DEREK_REQUIRE(p->kT > 0.0 || std::lround(p->nelec / 2.0) <= c->nb, "more occupied bands than bands");
    if (c->prec != DEREK_FP64) throw ArgError("resident SCF: mixed precision is not available in this build");
    c->P = *p;
    c->iscf = 0;
    const int nb = c->nb;
    const int m2 = 2 * nb;
    c->ensure_sub(m2);
    for (double** q : {&c->d_eps, &c->d_ekin, &c->d_occw}) {
      dfree(*q);
      *q = dmalloc<double>(nb > 0 ? (size_t)c->nk * nb : 1);
    }
    dfree(c->d_act);
    c->d_act = dmalloc<int32_t>(nb);
    for (double** q : {&c->rho_in, &c->rho_out, &c->vh, &c->vxc, &c->vhxc, &c->tmp1, &c->tmp2}) {
      dfree(*q);
      *q = dmalloc<double>(c->N);
    }
    for (double* q : c->h_rin) dfree(q);
    for (double* q : c->h_res) dfree(q);
    c->h_rin.assign(p->history, nullptr);
    c->h_res.assign(p->history, nullptr);
    for (int i = 0; i < p->history; ++i) {
      c->h_rin[i] = dmalloc<double>(c->N);
      c->h_res[i] = dmalloc<double>(c->N);
    }
    c->order.clear();
    c->A.assign(p->history, std::vector<double>(p->history, 0.0));
    c->eig.assign((size_t)c->nk * nb, 0.0);
    c->occ.assign((size_t)c->nk * nb, 0.0);
    c->res.assign((size_t)c->nk * nb, 0.0);
    for (int k = 0; k < c->nk; ++k) {
      KState& K = c->ks[k];
      alloc_k<double>(c, K);
      init_psi(reinterpret_cast<double2*>(K.Y), K.ld, nb, K.npw, K.kin, p->seed, k, c->s);
      c->orthonormalise<double>(k);
    }
    std::vector<double> r(c->N, p->nelec / c->volume);
    DEREK_HIP(hipMemcpy(c->rho_in, r.data(), c->N * sizeof(double), hipMemcpyHostToDevice));
    c->have_rho_out = false;
    c->scf_ready = true;
  });
}

int32_t derek_hip_scf_step(derek_ctx* c, derek_scf_report* rep) {
  return guard([&] {
    DEREK_REQUIRE(c && rep, "null argument");
    if (!c->scf_ready) throw StateError("scf_step before scf_init");
    c->iscf += 1;
    const auto tstep0 = std::chrono::steady_clock::now();
    c->t_dense = 0.0;
    const int nb = c->nb;
    const double dv = c->volume / (double)c->N;
    // V_in = V_ext + V_H[rho_in] + V_xc[rho_in]
    c->hartree(c->rho_in, c->vh, nullptr);
    c->lda(c->rho_in, c->vxc);
    add3(c->vh, c->vxc, nullptr, c->vhxc, c->N, c->s);
    add3(c->vext, c->vhxc, nullptr, c->vloc, c->N, c->s);
    c->vloc_set = true;
    int ndav = 0;
    const int nit = c->iscf == 1 ? c->P.n_david_first : c->P.n_david;
    for (int k = 0; k < c->nk; ++k) {
      const int d = c->davidson<double>(k, nit);
      if (k == 0) ndav = d;
    }
    double mu = 0.0, mts = 0.0;
    occupations(c->eig, c->ks, c->nk, nb, c->P.nelec, c->P.kT, c->occ, mu, mts);
    c->fermi = mu;
    c->minus_ts = mts;
    // rho_out
    std::vector<double> w((size_t)c->nk * nb);
    for (int k = 0; k < c->nk; ++k)
      for (int b = 0; b < nb; ++b)
        w[(size_t)k * nb + b] = c->ks[k].weight * c->occ[(size_t)k * nb + b] / c->volume;
    DEREK_HIP(hipMemcpyAsync(c->d_occw, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice, c->s));
    c->set_rho_target(c->rho_out);
    for (int k = 0; k < c->nk; ++k)
      c->density<double>(k, reinterpret_cast<const double2*>(c->ks[k].Y), c->ks[k].ld, nb, c->d_occw + (size_t)k * nb,
                         k > 0);
    c->have_rho_out = true;
    // energies
    double eh = 0.0;
    c->hartree(c->rho_out, c->tmp1, &eh);
    const double exc = c->lda(c->rho_out, c->tmp2);
    double e_band = 0.0;
    for (int k = 0; k < c->nk; ++k)
      for (int b = 0; b < nb; ++b)
        e_band += c->ks[k].weight * c->occ[(size_t)k * nb + b] * c->eig[(size_t)k * nb + b];
    const double e_dc = dv * c->dot(c->vhxc, c->rho_out, c->N);
    sub(c->tmp1, c->rho_out, c->rho_in, c->N, c->s);
    const double drho = std::sqrt(dv * c->dot(c->tmp1, c->tmp1, c->N));
    double maxres = 0.0;
    for (double x : c->res) maxres = std::max(maxres, x);
    rep->iscf = c->iscf;
    rep->ndav = ndav;
    rep->e_band = e_band;
    rep->e_hartree = eh;
    rep->e_xc = exc;
    rep->e_dc = e_dc;
    rep->minus_ts = mts;
    rep->fermi = mu;
    rep->etot = e_band - e_dc + eh + exc + mts;
    rep->drho = drho;
    rep->max_res = maxres;
    rep->t_dense = c->t_dense;
    c->mix();
    c->sync();
    rep->t_step = std::chrono::duration<double>(std::chrono::steady_clock::now() - tstep0).count();
  });
}

int32_t derek_hip_heev(derek_ctx* c, int32_t n, int32_t nev, const double* A, double* w, double* Z) {
  return guard([&] {
    DEREK_REQUIRE(c && A && w && Z, "null argument");
    DEREK_REQUIRE(n >= 1 && n <= 2048 && nev >= 1 && nev <= n, "need 1 <= nev <= n <= 2048");
    c->ensure_dense(n, nev);
    double2* dA = c->tA.get<double2>((size_t)n * n);
    double2* dZ = c->tB.get<double2>((size_t)n * nev);
    DEREK_HIP(hipMemcpyAsync(dA, A, (size_t)n * n * sizeof(double2), hipMemcpyHostToDevice, c->s));
    hermitize_gpu(dA, n, c->s);
    heev_gpu(dA, n, 0, n, nev, c->d_w, 0, dZ, n, 0, c->d_hwork, 0, 1, c->s);
    DEREK_HIP(hipMemcpyAsync(w, c->d_w, nev * sizeof(double), hipMemcpyDeviceToHost, c->s));
    DEREK_HIP(hipMemcpyAsync(Z, dZ, (size_t)n * nev * sizeof(double2), hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_hegv(derek_ctx* c, int32_t m, int32_t nev, const double* S, const double* H, double* w, double* C) {
  return guard([&] {
    DEREK_REQUIRE(c && S && H && w && C, "null argument");
    DEREK_REQUIRE(m >= 1 && m <= 640 && nev >= 1 && nev <= m, "need 1 <= nev <= m <= 640");
    const int cap = std::max(m, 2 * c->nb);
    c->ensure_sub(cap);
    c->ensure_dense(cap, nev);
    DEREK_HIP(hipMemcpyAsync(c->d_S, S, (size_t)m * m * sizeof(double2), hipMemcpyHostToDevice, c->s));
    DEREK_HIP(hipMemcpyAsync(c->d_H, H, (size_t)m * m * sizeof(double2), hipMemcpyHostToDevice, c->s));
    if (!c->gen_gpu(m, nev, w)) throw NumericError("overlap has fewer than nev independent columns");
    DEREK_HIP(hipMemcpyAsync(C, c->d_U, (size_t)m * nev * sizeof(double2), hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_profile_enable(derek_ctx* c, int32_t on) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    c->prof.resolve();
    c->prof.on = on != 0;
  });
}

int32_t derek_hip_profile_reset(derek_ctx* c) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    c->prof.reset();
  });
}

int32_t derek_hip_profile_read(derek_ctx* c, int32_t cls, int64_t* launches, double* total_ms, double* bytes,
                               double* flops) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    DEREK_REQUIRE(cls >= 0 && cls < KC_COUNT, "unknown kernel class");
    c->prof.resolve();
    if (launches) *launches = c->prof.count[cls];
    if (total_ms) *total_ms = c->prof.ms[cls];
    if (bytes) *bytes = c->prof.bytes[cls];
    if (flops) *flops = c->prof.flops[cls];
  });
}

int32_t derek_hip_scf_polish(derek_ctx* c, int32_t max_iter, double res_tol, double* max_res) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    DEREK_REQUIRE(max_iter >= 0 && res_tol >= 0.0, "max_iter and res_tol must be >= 0");
    if (!c->scf_ready || c->iscf == 0) throw StateError("scf_polish needs a completed scf_step");
    for (int k = 0; k < c->nk; ++k) c->davidson<double>(k, max_iter, res_tol);
    double r = 0.0;
    for (double x : c->res) r = std::max(r, x);
    if (max_res) *max_res = r;
  });
}

int32_t derek_hip_get_eigenvalues(derek_ctx* c, double* eig) {
  return guard([&] {
    DEREK_REQUIRE(c && eig, "null argument");
    if (!c->scf_ready) throw StateError("no SCF state");
    std::copy(c->eig.begin(), c->eig.end(), eig);
  });
}

int32_t derek_hip_get_occupations(derek_ctx* c, double* occ) {
  return guard([&] {
    DEREK_REQUIRE(c && occ, "null argument");
    if (!c->scf_ready) throw StateError("no SCF state");
    std::copy(c->occ.begin(), c->occ.end(), occ);
  });
}

int32_t derek_hip_get_density(derek_ctx* c, double* rho) {
  return guard([&] {
    DEREK_REQUIRE(c && rho, "null argument");
    if (!c->have_rho_out) throw StateError("no density computed yet");
    DEREK_HIP(hipMemcpyAsync(rho, c->rho_out, c->N * sizeof(double), hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_get_psi(derek_ctx* c, int32_t k, double* psi) {
  return guard([&] {
    DEREK_REQUIRE(c && psi, "null argument");
    c->require_basis(k);
    if (!c->scf_ready) throw StateError("no SCF state");
    const KState& K = c->ks[k];
    DEREK_HIP(hipMemcpy2DAsync(psi, K.npw * sizeof(double2), K.Y, K.ld * sizeof(double2), K.npw * sizeof(double2),
                               c->nb, hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_put_psi(derek_ctx* c, int32_t k, const double* psi) {
  return guard([&] {
    DEREK_REQUIRE(c && psi, "null argument");
    c->require_basis(k);
    if (!c->scf_ready) throw StateError("no SCF state");
    const KState& K = c->ks[k];
    DEREK_HIP(hipMemcpy2DAsync(K.Y, K.ld * sizeof(double2), psi, K.npw * sizeof(double2), K.npw * sizeof(double2),
                               c->nb, hipMemcpyHostToDevice, c->s));
    c->sync();
  });
}

int32_t derek_hip_apply_h_resident(derek_ctx* c, int32_t k, int32_t nb) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    c->require_basis(k);
    if (!c->scf_ready) throw StateError("no SCF state");
    if (!c->vloc_set) throw StateError("local potential not set");
    DEREK_REQUIRE(nb >= 1 && nb <= c->nb, "nb must be in 1..nbands");
    KState& K = c->ks[k];
    c->apply_h<double>(k, reinterpret_cast<const double2*>(K.Y), K.ld, reinterpret_cast<double2*>(K.HY), K.ld, nb);
  });
}

int32_t derek_hip_sync(derek_ctx* c) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    c->sync();
  });
}

int32_t derek_hip_stream(derek_ctx* c, void** stream) {
  return guard([&] {
    DEREK_REQUIRE(c && stream, "null argument");
    *stream = (void*)c->s;
  });
}

}  // extern "C"
