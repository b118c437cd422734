// Batched SCF engine: the body of the reference's electronic_scf
// (Source/electronic.f90:86-175) run for all k-points together.
//
// Layout in HBM: the wavefunction blocks of every k-point live in four
// buffers Y, HY, Y2, HY2 of nk * 2nb columns with one leading dimension
// ldk = round_up(max npw, 64); k-point k's block starts at k*kstride and
// holds [X | W] (bands, then corrections).  Each k-point has its own
// plane-wave count; rows past it are never read.
//
// One SCF step: V_in = V_ext + V_H[rho_in] + V_xc[rho_in]; for each batch of
// kc k-points, block Davidson (Rayleigh-Ritz on X, then n_david steps on
// [X, K R]) with every launch covering the whole batch; occupations,
// rho_out, energies and the Pulay mixer on the device.  Nothing waits for
// the host until the report is read at the end of the step, so the step can
// be captured once in a hipGraph and replayed.
//
// Precision: DEREK_FP64 keeps complex128 everywhere.  DEREK_MIXED keeps the
// wavefunctions in complex64 (FP32 FFTs, H|psi>, density, FP32 MFMA
// overlaps and rotations) with the subspace solves, reductions and energies
// in FP64, until |dE| < fp32_switch_tol; the wavefunctions are then
// promoted and the remaining steps (at least refine_steps) run in FP64.
#include <chrono>
#include <cmath>

#include "ov_geom.hpp"
#include "scf.hpp"
#include "subspace.hpp"

namespace derek {

struct ScfEngine {
  derek_scf_params P{};
  int iscf = 0, m2 = 0, kc = 1;
  int64_t ldk = 0, kstride = 0;
  bool fp32 = false, graphs = false;
  // the wavefunctions may be off orthonormal (just promoted from complex64):
  // the next FP64 step forms the full subspace overlaps, which restores it
  bool full_overlaps = false;
  int refine = 2, refine_left = 0;
  double switch_tol = 1e-5, e_prev = 0.0, de_prev = 1e300;
  bool have_prev = false, have_rho_out = false;
  void *Y = nullptr, *HY = nullptr, *Y2 = nullptr, *HY2 = nullptr;
  double2 *S = nullptr, *H = nullptr, *U = nullptr, *ovw = nullptr;
  int64_t ovw_cap = 0;  // double2 elements of ovw
  DenseWork dw;
  double *eps = nullptr, *ekin = nullptr, *rn = nullptr, *occ = nullptr, *occw = nullptr, *stats_partial = nullptr;
  double *rho_in = nullptr, *rho_out = nullptr, *vh = nullptr, *vxc = nullptr, *vhxc = nullptr, *tmp1 = nullptr,
         *tmp2 = nullptr;
  int* mix_st = nullptr;
  double *hin = nullptr, *hres = nullptr, *mixA = nullptr, *coef = nullptr, *mix_partial = nullptr;
  double* rep = nullptr;
  double* rep_h = nullptr;  // pinned
  // One executable graph per step shape, with derek_ctx::graph_epoch at the
  // time it was made; a graph whose epoch differs from the current one may
  // address freed work buffers or launch stale shapes and is not replayed.
  struct Graph {
    int nit;
    bool fp32, full;
    hipGraphExec_t exec;
    uint64_t epoch;
  };
  std::vector<Graph> cache;
  void drop_graphs() {
    for (auto& g : cache)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    cache.clear();
  }
  std::vector<void*> owned;
  template <typename T> T* alloc(size_t n) {
    T* p = dmalloc<T>(n);
    owned.push_back(p);
    return p;
  }
  void free_psi() {
    for (void** q : {&Y, &HY, &Y2, &HY2}) {
      dfree(*q);
      *q = nullptr;
    }
  }
  ~ScfEngine() {
    drop_graphs();
    free_psi();
    for (void* p : owned) dfree(p);
    if (rep_h) (void)hipHostFree(rep_h);
  }
};

void EngineDeleter::operator()(ScfEngine* e) const { delete e; }

namespace {

ScfEngine& engine(derek_ctx* c) {
  if (!c->scf) throw StateError("no SCF state (derek_hip_scf_init)");
  return *c->scf;
}

double npw_sum(const derek_ctx* c, int k0, int nkc) {
  double t = 0.0;
  for (int k = k0; k < k0 + nkc; ++k) t += (double)c->ks[k].npw;
  return t;
}

template <typename R> void alloc_psi(derek_ctx* c, ScfEngine& E) {
  const size_t n = (size_t)c->nk * E.kstride;
  for (void** q : {&E.Y, &E.HY, &E.Y2, &E.HY2}) {
    *q = dmalloc<cx_t<R>>(n);
    zero<R>(reinterpret_cast<cx_t<R>*>(*q), (int64_t)n, c->s);
  }
}

// bf16 terms of the complex64 GEMMs (complex128 data always uses FP64 MFMA)
template <typename R> static int bf16_terms(const derek_ctx* c) {
  return std::is_same<R, float>::value ? c->gemm_bf16 : 0;
}

// C[kz] (m x n, ld m, packed at m*n) = X^H Y for the batch k0..k0+nkc-1
template <typename R>
void ov(derek_ctx* c, const cx_t<R>* X, const cx_t<R>* Yp, int m, int n, int k0, int nkc, double2* C, bool herm) {
  ScfEngine& E = *c->scf;
  const double npw = npw_sum(c, k0, nkc);
  c->timed(KC_OVERLAP, (double)(herm ? m : m + n) * npw * sizeof(cx_t<R>), 8.0 * npw * m * n, [&] {
    overlap<R>(X, E.ldk, E.kstride, Yp, E.ldk, E.kstride, m, n, c->max_npw, c->d_npw + k0, C, m, (int64_t)m * n, nkc,
               E.ovw, E.ovw_cap, herm, c->s, bf16_terms<R>(c), 0, c->ov_rounds, c->ov_tile);
  });
}

// The X^H R / R^H R block form pays when it multiplies at least 20 % fewer
// 64 x 64 tiles than the Hermitian 2nb x 2nb overlap (it needs one more
// launch per matrix): from about 128 bands on.
static bool block_overlaps(int nb) {
  const int64_t t = ceil_div(nb, 64), T = ceil_div(2 * nb, 64);
  return 5 * (t * t + t * (t + 1) / 2) < 4 * (T * (T + 1) / 2);
}

// Block (rows of column block ca, columns of block cb) of the 2nb x 2nb
// subspace matrix C[kz] (ld 2nb): the nb x nb product X[:, ca:ca+nb]^H
// Yp[:, cb:cb+nb] for the batch k0..k0+nkc-1.
template <typename R>
void ov_block(derek_ctx* c, const cx_t<R>* X, int ca, const cx_t<R>* Yp, int cb, int k0, int nkc, double2* C,
              bool herm) {
  ScfEngine& E = *c->scf;
  const int nb = c->nb, m2 = E.m2;
  const double npw = npw_sum(c, k0, nkc);
  c->timed(KC_OVERLAP, (double)(herm ? nb : 2 * nb) * npw * sizeof(cx_t<R>), 8.0 * npw * nb * nb, [&] {
    overlap<R>(X + (int64_t)ca * E.ldk, E.ldk, E.kstride, Yp + (int64_t)cb * E.ldk, E.ldk, E.kstride, nb, nb,
               c->max_npw, c->d_npw + k0, C + ca + (int64_t)cb * m2, m2, (int64_t)m2 * m2, nkc, E.ovw, E.ovw_cap,
               herm, c->s, bf16_terms<R>(c), 0, c->ov_rounds, c->ov_tile);
  });
}

// Z[kz] (n columns) = X[kz] (m columns) U[kz] (m x n, ld m, stride su)
template <typename R>
void rot(derek_ctx* c, const cx_t<R>* X, const double2* Uc, int m, int n, int64_t su, int k0, int nkc, cx_t<R>* Z) {
  ScfEngine& E = *c->scf;
  const double npw = npw_sum(c, k0, nkc);
  c->timed(KC_ROTATE, (double)(m + n) * npw * sizeof(cx_t<R>), 8.0 * npw * m * n, [&] {
    rotate<R>(X, E.ldk, E.kstride, Uc, m, su, m, n, c->max_npw, c->d_npw + k0, Z, E.ldk, E.kstride, nkc, c->s,
              bf16_terms<R>(c));
  });
}

// Block Davidson for the k-points k0..k0+nkc-1 (reference algorithm: the
// oracle's davidson, oracle/scf.py); returns the steps done.  res_tol > 0
// stops early once every residual of the batch is below it (host check).
template <typename R> int davidson_chunk(derek_ctx* c, int k0, int nkc, int nit, double res_tol) {
  using C = cx_t<R>;
  ScfEngine& E = *c->scf;
  const int nb = c->nb, m2 = E.m2;
  const int64_t ld = E.ldk, ks = E.kstride;
  hipStream_t s = c->s;
  C* Ybase = reinterpret_cast<C*>(E.Y) + k0 * ks;
  C* Y = Ybase;
  C* HY = reinterpret_cast<C*>(E.HY) + k0 * ks;
  C* Y2 = reinterpret_cast<C*>(E.Y2) + k0 * ks;
  C* HY2 = reinterpret_cast<C*>(E.HY2) + k0 * ks;
  double2* S = E.S + (int64_t)k0 * m2 * m2;
  double2* H = E.H + (int64_t)k0 * m2 * m2;
  double2* U = E.U + (int64_t)k0 * m2 * m2;
  double* eps = E.eps + (int64_t)k0 * nb;
  int swaps = 0;
  auto pb = [&](C* p, int col) { return PwBlock<R>{p + (int64_t)col * ld, ld, ks}; };
  auto bb = [&](C* p, int col) { return BandBlock<C>{p + (int64_t)col * ld, ld, ks}; };
  auto stats = [&] {
    c->timed(KC_OTHER, 0, 0, [&] {
      band_stats<R>(bb(Y, 0), bb(HY, 0), c->d_maps, k0, nkc, c->max_npw, E.eps, nb, E.stats_partial, E.ekin, E.rn,
                    s);
    });
  };
  // Rayleigh-Ritz on X.  Complex64 blocks drift from orthonormality with
  // every rotation, so the FP32 phase solves X^H H X c = w X^H X c (which
  // re-orthonormalises X) where FP64 uses the standard problem.
  constexpr bool f32 = sizeof(R) == 4;
  // dependent-column threshold of the Cholesky: FP32 overlaps carry ~1e-6 errors
  const double drop = f32 ? 1e-5 : 1e-8;
  c->apply_h<R>(k0, nkc, pb(Y, 0), pb(HY, 0), nb);
  // Y^H (H Y) is Hermitian: only its upper tiles are multiplied, the lower
  // triangle is their conjugate mirror
  ov<R>(c, Y, HY, nb, nb, k0, nkc, H, true);
  if (f32) {
    ov<R>(c, Y, Y, nb, nb, k0, nkc, S, true);
    c->timed(KC_DENSE, 0, 0, [&] { gen_solve(E.dw, S, H, nb, nb, nkc, drop, eps, U, s); });
  } else {
    c->timed(KC_DENSE, 0, 0, [&] { rr_solve(E.dw, H, nb, nkc, eps, U, s); });
  }
  rot<R>(c, Y, U, nb, nb, (int64_t)nb * nb, k0, nkc, Y2);
  rot<R>(c, HY, U, nb, nb, (int64_t)nb * nb, k0, nkc, HY2);
  std::swap(Y, Y2);
  std::swap(HY, HY2);
  ++swaps;
  stats();
  int it = 0;
  for (; it < nit; ++it) {
    if (res_tol > 0.0) {
      std::vector<double> h((size_t)nkc * nb);
      DEREK_HIP(hipMemcpyAsync(h.data(), E.rn + (int64_t)k0 * nb, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                               s));
      c->sync();
      double rmax = 0.0;
      for (double x : h) rmax = std::max(rmax, x);
      if (rmax < res_tol) break;
    }
    c->timed(KC_OTHER, 0, 0, [&] {
      tpa_batch<R>(bb(Y, 0), bb(HY, 0), bb(Y, nb), c->d_maps, k0, nkc, c->max_npw, E.eps, E.ekin, E.rn, nb, 1e-10,
                   s);
    });
    c->apply_h<R>(k0, nkc, pb(Y, nb), pb(HY, nb), nb);
    if (f32 || E.full_overlaps || !block_overlaps(nb)) {
      ov<R>(c, Y, Y, m2, m2, k0, nkc, S, true);
      ov<R>(c, Y, HY, m2, m2, k0, nkc, H, true);
    } else {
      // X holds orthonormal Ritz vectors: S(X,X) = I, H(X,X) = diag(eps);
      // only X^H R, R^H R, X^H HR and R^H HR are multiplied
      ov_block<R>(c, Y, 0, Y, nb, k0, nkc, S, false);
      ov_block<R>(c, Y, nb, Y, nb, k0, nkc, S, true);
      ov_block<R>(c, Y, 0, HY, nb, k0, nkc, H, false);
      ov_block<R>(c, Y, nb, HY, nb, k0, nkc, H, true);
      c->timed(KC_OTHER, 0, 0, [&] { subspace_blocks(S, H, eps, m2, nb, nkc, s); });
    }
    c->timed(KC_DENSE, 0, 0, [&] { gen_solve(E.dw, S, H, m2, nb, nkc, drop, eps, U, s); });
    rot<R>(c, Y, U, m2, nb, (int64_t)m2 * nb, k0, nkc, Y2);
    rot<R>(c, HY, U, m2, nb, (int64_t)m2 * nb, k0, nkc, HY2);
    std::swap(Y, Y2);
    std::swap(HY, HY2);
    ++swaps;
    stats();
  }
  if (swaps & 1) {
    // keep X of every k-point in the Y buffer between calls
    const size_t pitch = (size_t)ks * sizeof(C);
    DEREK_HIP(hipMemcpy2DAsync(Ybase, pitch, Y, pitch, (size_t)nb * ld * sizeof(C), nkc, hipMemcpyDeviceToDevice, s));
  }
  return it;
}

void potential(derek_ctx* c) {
  ScfEngine& E = *c->scf;
  c->hartree(E.rho_in, E.vh, nullptr, 0);
  c->lda(E.rho_in, E.vxc, nullptr, 0);
  add3(E.vh, E.vxc, nullptr, E.vhxc, c->N, c->s);
  add3(c->vext, E.vhxc, nullptr, c->vloc, c->N, c->s);
  c->vloc_changed(E.fp32);
  c->vloc_set = true;
}

template <typename R> void enqueue_step(derek_ctx* c, int nit) {
  ScfEngine& E = *c->scf;
  const int nk = c->nk, nb = c->nb;
  const int64_t N = c->N;
  const double dv = c->volume / (double)N;
  hipStream_t s = c->s;
  potential(c);
  for (int k0 = 0; k0 < nk; k0 += E.kc) davidson_chunk<R>(c, k0, std::min(E.kc, nk - k0), nit, 0.0);
  E.full_overlaps = false;
  occupations(E.eps, c->d_wk, nk, nb, E.P.nelec, E.P.kT, c->volume, E.occ, E.occw, E.rep, s);
  c->density<R>(0, nk, PwBlock<R>{reinterpret_cast<cx_t<R>*>(E.Y), E.ldk, E.kstride}, nb, E.occw, nb, E.rho_out,
                false);
  c->hartree(E.rho_out, E.tmp1, E.rep, REP_EH);
  c->lda(E.rho_out, E.tmp2, E.rep, REP_EXC);
  c->dot_into(E.vhxc, E.rho_out, N, dv, false, E.rep, REP_EDC);
  sub(E.tmp1, E.rho_out, E.rho_in, N, s);
  c->dot_into(E.tmp1, E.tmp1, N, dv, true, E.rep, REP_DRHO);
  finish_max(E.rn, nk * nb, E.rep, REP_MAXRES, s);
  // rho_in <- sum a_i rho_in,i + beta K[sum a_i F_i]  (E.tmp1 = sum a_i F_i)
  pulay_update(E.mix_st, E.P.history, E.rho_in, E.rho_out, E.hin, E.hres, E.mixA, E.coef, E.mix_partial, E.rho_in,
               E.tmp1, N, s);
  c->kerker(E.tmp1, E.tmp2, E.P.mix_q0);
  axpy(E.rho_in, E.tmp2, E.P.mix_beta, N, s);
}

// Runs one step's work: eagerly, or from a graph captured at the second
// step with the same shape (the first one allocates and sets attributes).
void run_step(derek_ctx* c, int nit) {
  ScfEngine& E = *c->scf;
  auto enqueue = [&] {
    if (E.fp32) enqueue_step<float>(c, nit);
    else enqueue_step<double>(c, nit);
  };
  if (!E.graphs) {
    enqueue();
    return;
  }
  ScfEngine::Graph* g = nullptr;
  const bool full = E.full_overlaps;
  for (auto& e : E.cache)
    if (e.nit == nit && e.fp32 == E.fp32 && e.full == full) g = &e;
  if (g && g->exec && g->epoch != c->graph_epoch) {
    // a work buffer was reallocated or freed, or a launch setting changed,
    // since this graph was made: make it again
    (void)hipGraphExecDestroy(g->exec);
    g->exec = nullptr;
    g->epoch = ~(uint64_t)0;
  }
  // Eager step: the first use of a shape, or its first use since its graph
  // went stale.  It allocates what the shape needs, so the capture at the
  // shape's next use sees no allocation.
  auto eager = [&] {
    enqueue();
    if (!g) {
      E.cache.push_back({nit, E.fp32, full, nullptr, 0});
      g = &E.cache.back();
    }
    g->epoch = c->graph_epoch;
  };
  if (!g || (!g->exec && g->epoch == ~(uint64_t)0)) {
    eager();
    return;
  }
  if (!g->exec) {
    const uint64_t e0 = c->graph_epoch;
    c->prof.paused = true;
    hipGraph_t graph = nullptr;
    DEREK_HIP(hipStreamBeginCapture(c->s, hipStreamCaptureModeThreadLocal));
    try {
      enqueue();
    } catch (...) {
      hipGraph_t junk = nullptr;
      (void)hipStreamEndCapture(c->s, &junk);
      if (junk) (void)hipGraphDestroy(junk);
      c->prof.paused = false;
      throw;
    }
    DEREK_HIP(hipStreamEndCapture(c->s, &graph));
    c->prof.paused = false;
    if (c->graph_epoch != e0) {
      // a work buffer was (re)allocated while the step was recorded
      // (nothing ran): run the step eagerly, capture at the next use
      (void)hipGraphDestroy(graph);
      eager();
      return;
    }
    const hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    DEREK_HIP(e);
    g->epoch = e0;
  }
  DEREK_HIP(hipGraphLaunch(g->exec, c->s));
}

void switch_to_fp64(derek_ctx* c) {
  ScfEngine& E = *c->scf;
  const size_t n = (size_t)c->nk * E.kstride;
  double2* y64 = dmalloc<double2>(n);
  convert<double, float>(y64, reinterpret_cast<const float2*>(E.Y), (int64_t)n, c->s);
  c->sync();
  E.free_psi();
  E.Y = y64;
  for (void** q : {&E.HY, &E.Y2, &E.HY2}) {
    *q = dmalloc<double2>(n);
    zero<double>(reinterpret_cast<double2*>(*q), (int64_t)n, c->s);
  }
  E.fp32 = false;
  E.full_overlaps = true;
  c->S32.release();
  c->P32.release();
}

// Times H|psi> for the tunable launch shapes, starting from the current
// ones: bands per workgroup on the current (fused or separate) y/x pass,
// the other pass at the current bands per workgroup, then half and twice
// the items per launch.  The timed sample is at most about 256 items (the
// first k-points' leading bands), so timings are noisy: another shape
// replaces the current one only when it is more than 3 % faster.
template <typename R> void autotune(derek_ctx* c) {
  ScfEngine& E = *c->scf;
  constexpr int kTuneItems = 256;
  constexpr float kGain = 0.97f;
  const int nbt = std::min(c->nb, kTuneItems);
  const int nkc = std::min({E.kc, c->nk, (int)ceil_div(kTuneItems, nbt)});
  cx_t<R>* Y = reinterpret_cast<cx_t<R>*>(E.Y);
  cx_t<R>* HY = reinterpret_cast<cx_t<R>*>(E.HY);
  const PwBlock<R> x{Y, E.ldk, E.kstride}, hx{HY, E.ldk, E.kstride};
  const bool prof = c->prof.on;
  c->prof.on = false;
  hipEvent_t a, b;
  DEREK_HIP(hipEventCreate(&a));
  DEREK_HIP(hipEventCreate(&b));
  auto time_it = [&] {
    float best = 1e30f;
    c->apply_h<R>(0, nkc, x, hx, nbt);
    for (int r = 0; r < 2; ++r) {
      DEREK_HIP(hipEventRecord(a, c->s));
      c->apply_h<R>(0, nkc, x, hx, nbt);
      DEREK_HIP(hipEventRecord(b, c->s));
      DEREK_HIP(hipEventSynchronize(b));
      float t = 0.f;
      DEREK_HIP(hipEventElapsedTime(&t, a, b));
      best = std::min(best, t);
    }
    return best;
  };
  // bytes of stick + plane buffers per item; the sweep stays within 8 GiB
  const double item_bytes =
      (double)sizeof(cx_t<R>) * ((double)c->nst_max * c->dims[2] + (double)c->nax_max * c->dims[1] * c->dims[2]);
  const int cap_mem = (int)std::max(1.0, std::floor(8.0 * 1073741824.0 / item_bytes));
  c->item_cap = std::max(1, std::min(c->item_cap, cap_mem));
  // bring the clocks up before the first timed shape: at least 15 ms of work
  {
    const auto w0 = std::chrono::steady_clock::now();
    for (int r = 0; r < 100; ++r) {
      c->apply_h<R>(0, nkc, x, hx, nbt);
      c->sync();
      if (std::chrono::steady_clock::now() - w0 > std::chrono::milliseconds(15)) break;
    }
  }
  float best = time_it();
  auto try_shape = [&](int bpw, int fz, int cap) {
    const int b0 = c->bpw, f0 = c->fused, c0 = c->item_cap;
    c->bpw = bpw;
    c->fused = fz;
    c->item_cap = cap;
    const float t = time_it();
    if (t < kGain * best) {
      best = t;
      return;
    }
    c->bpw = b0;
    c->fused = f0;
    c->item_cap = c0;
  };
  // bit 0 of `fused` (the H|psi> pass) is tuned; the density bit is kept
  const int dens = c->fused & derek_ctx::kFusedDensity;
  const int fz0 = c->fused & derek_ctx::kFusedHpsi;
  const int bpw0 = c->bpw;
  // whether this grid has the fused pass at all
  c->fused = derek_ctx::kFusedHpsi | dens;
  const bool can_fuse = c->use_slab<R>(derek_ctx::kFusedHpsi);
  c->fused = fz0 | dens;
  for (int q : {1, 2, 4, 8})
    if (q != bpw0) try_shape(q, fz0 | dens, c->item_cap);
  if (can_fuse) try_shape(c->bpw, (fz0 ^ derek_ctx::kFusedHpsi) | dens, c->item_cap);
  const int all = nkc * nbt, cap0 = c->item_cap;
  for (int q : {cap0 / 2, cap0 * 2})
    if (q >= 1 && q <= cap_mem && q <= std::max(all, cap0)) try_shape(c->bpw, c->fused, q);
  // FFT tile: the other number of sticks per z-pass workgroup
  {
    const int zprev = c->zsl;
    const int z0 = zprev ? zprev : (sizeof(R) == 8 ? 16 : 32);
    c->zsl = z0 == 16 ? 32 : 16;
    const float t = time_it();
    if (t < kGain * best) best = t;
    else c->zsl = zprev;
  }
  // One launch per pass over every (k-point, band) item of the first batch
  // (the whole SCF batch), against launches of item_cap items, timed on that
  // whole batch; the pipeline buffers may then take up to 24 GiB.
  {
    const int full = std::min(E.kc, c->nk) * c->nb;
    const int cap_all = (int)std::max(1.0, std::floor(24.0 * 1073741824.0 / item_bytes));
    if (full > c->item_cap && full <= cap_all) {
      const int nkf = std::min(E.kc, c->nk), nbf = c->nb;
      auto time_full = [&] {
        float tb = 1e30f;
        c->apply_h<R>(0, nkf, x, hx, nbf);
        for (int r = 0; r < 2; ++r) {
          DEREK_HIP(hipEventRecord(a, c->s));
          c->apply_h<R>(0, nkf, x, hx, nbf);
          DEREK_HIP(hipEventRecord(b, c->s));
          DEREK_HIP(hipEventSynchronize(b));
          float t = 0.f;
          DEREK_HIP(hipEventElapsedTime(&t, a, b));
          tb = std::min(tb, t);
        }
        return tb;
      };
      const int c0 = c->item_cap;
      const float t0 = time_full();
      c->item_cap = full;
      const float t1 = time_full();
      if (!(t1 < kGain * t0)) c->item_cap = c0;
    }
  }
  // GEMM tiles and chunks: the C tile edge of the BF16 overlaps and the
  // rounds of resident workgroups their split-K row chunks fill, timed on
  // the Davidson step's 2nb x 2nb Hermitian overlap of the first batch
  {
    const int nkg = std::min(E.kc, c->nk);
    cx_t<R>* Yb = reinterpret_cast<cx_t<R>*>(E.Y);
    auto time_ov = [&] {
      float tb = 1e30f;
      ov<R>(c, Yb, Yb, E.m2, E.m2, 0, nkg, E.S, true);
      for (int r = 0; r < 2; ++r) {
        DEREK_HIP(hipEventRecord(a, c->s));
        ov<R>(c, Yb, Yb, E.m2, E.m2, 0, nkg, E.S, true);
        DEREK_HIP(hipEventRecord(b, c->s));
        DEREK_HIP(hipEventSynchronize(b));
        float t = 0.f;
        DEREK_HIP(hipEventElapsedTime(&t, a, b));
        tb = std::min(tb, t);
      }
      return tb;
    };
    float bestg = time_ov();
    // tile edge first (BF16 overlaps only), then the rounds at that edge
    if (bf16_terms<R>(c) > 0) {
      const int t0 = c->ov_tile;
      c->ov_tile = t0 == 64 ? 128 : 64;
      const float t = time_ov();
      if (t < kGain * bestg) bestg = t;
      else c->ov_tile = t0;
    }
    const int r0 = c->ov_rounds;
    for (int q : kOvRounds) {
      if (q == r0) continue;
      const int prev = c->ov_rounds;
      c->ov_rounds = q;
      const float t = time_ov();
      if (t < kGain * bestg) bestg = t;
      else c->ov_rounds = prev;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  c->prof.on = prof;
  // drop the sweep's largest buffers; the chosen size is allocated on use
  c->S64.release();
  c->P64.release();
  c->S32.release();
  c->P32.release();
}

}  // namespace

void scf_init(derek_ctx* c, const derek_scf_params& p) {
  c->require_all_bases();
  if (!c->gsq_set) throw StateError("scf_init: |G|^2 not set");
  if (!c->vext_set) throw StateError("scf_init: external potential not set");
  DEREK_REQUIRE(p.nelec > 0.0, "nelec must be positive");
  DEREK_REQUIRE(p.kT >= 0.0, "kT must be >= 0");
  DEREK_REQUIRE(p.history >= 1 && p.history <= 16, "Pulay history must be 1..16");
  DEREK_REQUIRE(p.n_david_first >= 0 && p.n_david >= 0, "Davidson step counts must be >= 0");
  // fixed occupations fill nelec/2 bands with two electrons each; Fermi-Dirac
  // occupations need room for nelec in 2 * nbands * sum_k w_k
  if (p.kT == 0.0) {
    DEREK_REQUIRE(std::fabs(p.nelec / 2.0 - std::nearbyint(p.nelec / 2.0)) < 1e-9,
                  "fixed occupations (kT = 0) need an even number of electrons");
    DEREK_REQUIRE(std::lround(p.nelec / 2.0) <= c->nb, "more occupied bands than bands");
  } else {
    double wsum = 0.0;
    for (const KState& K : c->ks) wsum += K.weight;
    DEREK_REQUIRE(p.nelec <= 2.0 * c->nb * wsum * (1.0 + 1e-12),
                  "nelec exceeds what the bands hold (2 * nbands * sum of k-point weights)");
  }
  DEREK_REQUIRE(p.fp32_switch_tol >= 0.0 && p.refine_steps >= 0 && p.kbatch >= 0, "negative tuning parameter");
  DEREK_REQUIRE(2 * c->nb <= 640, "2 * nbands must be <= 640");
  if (c->s) c->sync();
  c->scf.reset();
  c->scf.reset(new ScfEngine());
  ScfEngine& E = *c->scf;
  E.P = p;
  const int nk = c->nk, nb = c->nb;
  E.m2 = 2 * nb;
  int64_t npw_max = 0;
  for (const KState& K : c->ks) npw_max = std::max(npw_max, K.npw);
  E.ldk = round_up(npw_max, 64);
  E.kstride = (int64_t)E.m2 * E.ldk;
  E.kc = p.kbatch > 0 ? std::min(p.kbatch, nk) : nk;
  E.graphs = p.use_graph != 0;
  E.fp32 = c->prec == DEREK_MIXED;
  E.switch_tol = p.fp32_switch_tol > 0.0 ? p.fp32_switch_tol : 1e-5;
  // two FP64 steps: the first restores orthonormality (full overlaps), the
  // stop test then compares two consecutive FP64 energies
  E.refine = p.refine_steps > 0 ? p.refine_steps : 2;
  E.refine_left = E.fp32 ? E.refine : 0;
  if (E.fp32) alloc_psi<float>(c, E);
  else alloc_psi<double>(c, E);
  const size_t mm = (size_t)E.m2 * E.m2;
  E.S = E.alloc<double2>(mm * nk);
  E.H = E.alloc<double2>(mm * nk);
  E.U = E.alloc<double2>(mm * nk);
  E.dw.ensure(E.m2, E.m2, E.kc);
  // partial sums of every overlap shape the engine launches: nb x nb and
  // 2nb x 2nb, for full batches and the last, shorter one
  E.ovw_cap = engine_overlap_work_elems(nb, nk, E.kc, c->max_npw);
  E.ovw = E.alloc<double2>(E.ovw_cap);
  for (double** q : {&E.eps, &E.ekin, &E.rn, &E.occ, &E.occw}) *q = E.alloc<double>((size_t)nk * nb);
  E.stats_partial = E.alloc<double>((size_t)E.kc * nb * band_stats_chunks(c->max_npw) * 2);
  for (double** q : {&E.rho_in, &E.rho_out, &E.vh, &E.vxc, &E.vhxc, &E.tmp1, &E.tmp2}) *q = E.alloc<double>(c->N);
  const int Hh = p.history;
  E.hin = E.alloc<double>((size_t)Hh * c->N);
  E.hres = E.alloc<double>((size_t)Hh * c->N);
  E.mixA = E.alloc<double>((size_t)Hh * Hh);
  E.coef = E.alloc<double>(Hh);
  E.mix_partial = E.alloc<double>((size_t)Hh * pulay_blocks(c->N));
  E.mix_st = E.alloc<int>(1);
  E.rep = E.alloc<double>(REP_COUNT);
  DEREK_HIP(hipHostMalloc(reinterpret_cast<void**>(&E.rep_h), REP_COUNT * sizeof(double), hipHostMallocDefault));
  {
    std::vector<double> zeros((size_t)Hh * Hh + Hh, 0.0);
    DEREK_HIP(hipMemcpyAsync(E.mixA, zeros.data(), (size_t)Hh * Hh * sizeof(double), hipMemcpyHostToDevice, c->s));
    DEREK_HIP(hipMemcpyAsync(E.coef, zeros.data(), Hh * sizeof(double), hipMemcpyHostToDevice, c->s));
    const int z = 0;
    DEREK_HIP(hipMemcpyAsync(E.mix_st, &z, sizeof(int), hipMemcpyHostToDevice, c->s));
    c->sync();
  }
  // random wavefunctions, orthonormalised per k-point: X <- X D L^-H
  auto init = [&](auto tag) {
    using R = decltype(tag);
    using C = cx_t<R>;
    C* Y = reinterpret_cast<C*>(E.Y);
    C* Y2 = reinterpret_cast<C*>(E.Y2);
    init_psi_batch<R>(BandBlock<C>{Y, E.ldk, E.kstride}, c->d_maps, 0, nk, c->max_npw, nb, p.seed, c->s);
    for (int k0 = 0; k0 < nk; k0 += E.kc) {
      const int nkc = std::min(E.kc, nk - k0);
      C* Yk = Y + k0 * E.kstride;
      C* Y2k = Y2 + k0 * E.kstride;
      double2* S = E.S + (int64_t)k0 * mm;
      double2* U = E.U + (int64_t)k0 * mm;
      ov<R>(c, Yk, Yk, nb, nb, k0, nkc, S, true);
      orth_coeffs(E.dw, S, nb, nkc, U, c->s);
      std::vector<int> ok(nkc);
      DEREK_HIP(hipMemcpyAsync(ok.data(), E.dw.ok, nkc * sizeof(int), hipMemcpyDeviceToHost, c->s));
      c->sync();
      for (int v : ok)
        if (!v) throw NumericError("initial wavefunctions are linearly dependent");
      rot<R>(c, Yk, U, nb, nb, (int64_t)nb * nb, k0, nkc, Y2k);
      const size_t pitch = (size_t)E.kstride * sizeof(C);
      DEREK_HIP(hipMemcpy2DAsync(Yk, pitch, Y2k, pitch, (size_t)nb * E.ldk * sizeof(C), nkc, hipMemcpyDeviceToDevice,
                                 c->s));
    }
  };
  if (E.fp32) init(float{});
  else init(double{});
  std::vector<double> r(c->N, p.nelec / c->volume);
  DEREK_HIP(hipMemcpy(E.rho_in, r.data(), c->N * sizeof(double), hipMemcpyHostToDevice));
  if (p.autotune) {
    add3(c->vext, nullptr, nullptr, c->vloc, c->N, c->s);
    c->vloc_changed(true);
    if (E.fp32) autotune<float>(c);
    else autotune<double>(c);
  }
  c->sync();
}

void scf_step(derek_ctx* c, derek_scf_report& rep) {
  ScfEngine& E = engine(c);
  const auto t0 = std::chrono::steady_clock::now();
  E.iscf += 1;
  const int nit = E.iscf == 1 ? E.P.n_david_first : E.P.n_david;
  double dense0 = 0.0;
  if (c->prof.on) {
    c->prof.resolve();
    dense0 = c->prof.ms[KC_DENSE];
  }
  const bool fp32_step = E.fp32;
  run_step(c, nit);
  DEREK_HIP(hipMemcpyAsync(E.rep_h, E.rep, REP_COUNT * sizeof(double), hipMemcpyDeviceToHost, c->s));
  c->sync();
  E.have_rho_out = true;
  const double* h = E.rep_h;
  rep = derek_scf_report{};
  rep.iscf = E.iscf;
  rep.ndav = nit;
  rep.e_band = h[REP_EBAND];
  rep.e_hartree = h[REP_EH];
  rep.e_xc = h[REP_EXC];
  rep.e_dc = h[REP_EDC];
  rep.minus_ts = h[REP_MTS];
  rep.fermi = h[REP_MU];
  rep.drho = h[REP_DRHO];
  rep.max_res = h[REP_MAXRES];
  rep.etot = rep.e_band - rep.e_dc + rep.e_hartree + rep.e_xc + rep.minus_ts;
  rep.fp32 = fp32_step ? 1 : 0;
  if (fp32_step) {
    E.refine_left = E.refine;
    if (E.have_prev) {
      // leave the FP32 phase at |dE| < switch_tol, at its noise floor
      // (2e-7 |E|, a few complex64 roundings of the energy), or once |dE|
      // grows again after the step has reached 1e-3 Ha
      const double de = std::fabs(rep.etot - E.e_prev);
      const bool floor = de < std::max(E.switch_tol, 2e-7 * std::fabs(rep.etot));
      const bool stall = E.de_prev < 1e-3 && de > E.de_prev;
      if (floor || stall) switch_to_fp64(c);
      E.de_prev = de;
    }
  } else if (E.refine_left > 0) {
    E.refine_left -= 1;
  }
  rep.refine_left = E.refine_left;
  E.e_prev = rep.etot;
  E.have_prev = true;
  if (c->prof.on) {
    c->prof.resolve();
    rep.t_dense = (c->prof.ms[KC_DENSE] - dense0) * 1e-3;
  }
  rep.t_step = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double scf_polish(derek_ctx* c, int max_iter, double res_tol) {
  ScfEngine& E = engine(c);
  if (E.iscf == 0) throw StateError("scf_polish needs a completed scf_step");
  for (int k0 = 0; k0 < c->nk; k0 += E.kc) {
    const int nkc = std::min(E.kc, c->nk - k0);
    if (E.fp32) davidson_chunk<float>(c, k0, nkc, max_iter, res_tol);
    else davidson_chunk<double>(c, k0, nkc, max_iter, res_tol);
  }
  std::vector<double> h((size_t)c->nk * c->nb);
  DEREK_HIP(hipMemcpyAsync(h.data(), E.rn, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->s));
  c->sync();
  double r = 0.0;
  for (double x : h) r = std::max(r, x);
  return r;
}

// Eigenpairs at a fixed potential: Davidson on every k-point at the total
// potential set with set_vloc (no density or potential update), until every
// residual is below res_tol or max_iter iterations; the non-self-consistent
// part of the spectral tasks.
double scf_nscf(derek_ctx* c, int max_iter, double res_tol) {
  ScfEngine& E = engine(c);
  if (!c->vloc_set) throw StateError("non-self-consistent run needs the potential (set_vloc)");
  c->vloc_changed(E.fp32);
  for (int k0 = 0; k0 < c->nk; k0 += E.kc) {
    const int nkc = std::min(E.kc, c->nk - k0);
    if (E.fp32) davidson_chunk<float>(c, k0, nkc, max_iter, res_tol);
    else davidson_chunk<double>(c, k0, nkc, max_iter, res_tol);
  }
  std::vector<double> h((size_t)c->nk * c->nb);
  DEREK_HIP(hipMemcpyAsync(h.data(), E.rn, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->s));
  c->sync();
  double r = 0.0;
  for (double x : h) r = std::max(r, x);
  return r;
}

void scf_get(derek_ctx* c, int what, double* out) {
  ScfEngine& E = engine(c);
  if (what == 2) {
    if (!E.have_rho_out) throw StateError("no density computed yet");
    DEREK_HIP(hipMemcpyAsync(out, E.rho_out, c->N * sizeof(double), hipMemcpyDeviceToHost, c->s));
  } else if (what == 3) {
    if (!c->vloc_set) throw StateError("no potential set yet");
    DEREK_HIP(hipMemcpyAsync(out, c->vloc, c->N * sizeof(double), hipMemcpyDeviceToHost, c->s));
  } else {
    const double* src = what == 0 ? E.eps : (what == 1 ? E.occ : E.rn);
    DEREK_HIP(hipMemcpyAsync(out, src, (size_t)c->nk * c->nb * sizeof(double), hipMemcpyDeviceToHost, c->s));
  }
  c->sync();
}

void scf_psi(derek_ctx* c, int k, double* psi, bool put) {
  ScfEngine& E = engine(c);
  const int64_t npw = c->ks[k].npw;
  const size_t n = (size_t)c->nb * E.ldk;
  if (!E.fp32) {
    double2* y = reinterpret_cast<double2*>(E.Y) + k * E.kstride;
    if (put)
      DEREK_HIP(hipMemcpy2DAsync(y, E.ldk * sizeof(double2), psi, npw * sizeof(double2), npw * sizeof(double2), c->nb,
                                 hipMemcpyHostToDevice, c->s));
    else
      DEREK_HIP(hipMemcpy2DAsync(psi, npw * sizeof(double2), y, E.ldk * sizeof(double2), npw * sizeof(double2), c->nb,
                                 hipMemcpyDeviceToHost, c->s));
  } else {
    float2* y = reinterpret_cast<float2*>(E.Y) + k * E.kstride;
    double2* t = c->tA.get<double2>(n);
    if (put) {
      convert<double, float>(t, y, (int64_t)n, c->s);  // keeps the padding rows
      DEREK_HIP(hipMemcpy2DAsync(t, E.ldk * sizeof(double2), psi, npw * sizeof(double2), npw * sizeof(double2), c->nb,
                                 hipMemcpyHostToDevice, c->s));
      convert<float, double>(y, t, (int64_t)n, c->s);
    } else {
      convert<double, float>(t, y, (int64_t)n, c->s);
      DEREK_HIP(hipMemcpy2DAsync(psi, npw * sizeof(double2), t, E.ldk * sizeof(double2), npw * sizeof(double2), c->nb,
                                 hipMemcpyDeviceToHost, c->s));
    }
  }
  c->sync();
}

void scf_apply_h_resident(derek_ctx* c, int k, int nb) {
  ScfEngine& E = engine(c);
  // k = -1: every k-point in one batch
  const int k0 = k < 0 ? 0 : k, nkc = k < 0 ? c->nk : 1;
  if (E.fp32) {
    float2* y = reinterpret_cast<float2*>(E.Y) + k0 * E.kstride;
    float2* hy = reinterpret_cast<float2*>(E.HY) + k0 * E.kstride;
    c->apply_h<float>(k0, nkc, PwBlock<float>{y, E.ldk, E.kstride}, PwBlock<float>{hy, E.ldk, E.kstride}, nb);
  } else {
    double2* y = reinterpret_cast<double2*>(E.Y) + k0 * E.kstride;
    double2* hy = reinterpret_cast<double2*>(E.HY) + k0 * E.kstride;
    c->apply_h<double>(k0, nkc, PwBlock<double>{y, E.ldk, E.kstride}, PwBlock<double>{hy, E.ldk, E.kstride}, nb);
  }
}

void scf_tuning(derek_ctx* c, int32_t* out) {
  if (c->scf) {
    out[0] = c->scf->kc;
    out[1] = c->scf->graphs ? 1 : 0;
  } else {
    out[0] = 0;
    out[1] = 0;
  }
}

}  // namespace derek
