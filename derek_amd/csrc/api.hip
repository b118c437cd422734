// Context methods and the C-ABI entry points (the SCF engine is in scf.hip).
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

#include "basis.hpp"
#include "context.hpp"
#include "info.hpp"
#include "interchange.hpp"
#include "scf.hpp"
#include "subspace.hpp"

using namespace derek;

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
  } catch (const IoError& e) {
    g_last_error = e.what();
    return DEREK_ERR_IO;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return DEREK_ERR_HIP;
  }
}

}  // namespace

// ------------------------------------------------------------------ profiler
namespace derek {
hipEvent_t Profiler::take() {
  if (pool.empty()) {
    hipEvent_t e;
    DEREK_HIP(hipEventCreate(&e));
    return e;
  }
  hipEvent_t e = pool.back();
  pool.pop_back();
  return e;
}
void Profiler::resolve() {
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
void Profiler::reset() {
  resolve();
  for (int i = 0; i < KC_COUNT; ++i) {
    count[i] = 0;
    ms[i] = bytes[i] = flops[i] = 0.0;
  }
}
Profiler::~Profiler() {
  for (auto& p : pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : pool) (void)hipEventDestroy(e);
}
}  // namespace derek

// ------------------------------------------------------------------ context
derek_ctx::derek_ctx() {
  for (derek::DevBuf* b : {&S64, &P64, &S32, &P32, &rho_part, &tA, &tB, &tC, &tD, &ovwork, &partial})
    b->epoch = &graph_epoch;
}

derek_ctx::~derek_ctx() {
  if (s) (void)hipStreamSynchronize(s);
  scf.reset();
  for (int a = 0; a < 3; ++a) {
    dfree(tw64[a]);
    dfree(tw32[a]);
  }
  for (auto& k : ks) {
    dfree(k.idx);
    dfree(k.kin);
    dfree(k.mapbuf);
  }
  for (void* p : {(void*)d_maps, (void*)d_npw, (void*)d_wk, (void*)gsq, (void*)vext, (void*)vloc, (void*)vloc32,
                  (void*)vt64, (void*)vt32, (void*)cgrid})
    dfree(p);
  S64.release();
  P64.release();
  S32.release();
  P32.release();
  rho_part.release();
  tA.release();
  tB.release();
  tC.release();
  tD.release();
  ovwork.release();
  partial.release();
  if (s) (void)hipStreamDestroy(s);
}

template <> const Axis<double>* derek_ctx::axes<double>() const { return ax64; }
template <> const Axis<float>* derek_ctx::axes<float>() const { return ax32; }
template <> const double* derek_ctx::vlocal<double>() const { return vloc; }
template <> const float* derek_ctx::vlocal<float>() const { return vloc32; }
template <> const double* derek_ctx::vtrans<double>() const { return vt64; }
template <> const float* derek_ctx::vtrans<float>() const { return vt32; }
template <typename R> bool derek_ctx::use_slab(int part) const {
  return (fused & part) && slab_supported<R>(dims[0], dims[1], dims[2], nax_max, nst_max);
}
template bool derek_ctx::use_slab<double>(int) const;
template bool derek_ctx::use_slab<float>(int) const;

void derek_ctx::vloc_changed(bool f32) {
  if (f32) convert_real_f(vloc32, vloc, N, s);
  transpose_xz(vloc, vt64, f32 ? vt32 : nullptr, dims[0], dims[1], dims[2], 1.0 / (double)N, s);
}
template <> double2* derek_ctx::sticks<double>(int items) {
  return S64.get<double2>((size_t)items * nst_max * dims[2]);
}
template <> float2* derek_ctx::sticks<float>(int items) { return S32.get<float2>((size_t)items * nst_max * dims[2]); }
template <> double2* derek_ctx::planes<double>(int items) {
  return P64.get<double2>((size_t)items * nax_max * dims[1] * dims[2]);
}
template <> float2* derek_ctx::planes<float>(int items) {
  return P32.get<float2>((size_t)items * nax_max * dims[1] * dims[2]);
}

void derek_ctx::require_basis(int k) const {
  if (k < 0 || k >= nk) throw ArgError("k-point index " + std::to_string(k) + " out of range");
  if (!ks[k].set) throw StateError("basis of k-point " + std::to_string(k) + " not set");
}

void derek_ctx::require_all_bases() const {
  for (int k = 0; k < nk; ++k) require_basis(k);
}

// ------------------------------------------------------------------ pipelines
// Splits (k0..k0+nkc) x (0..nbands) into launches of at most item_cap items:
// all k-points of a slice together, the bands in groups.
template <typename F> static void for_item_launches(const derek_ctx* c, int k0, int nkc, int nbands, F&& f) {
  const int kstep = std::min(nkc, c->item_cap);
  for (int kk = 0; kk < nkc; kk += kstep) {
    const int kn = std::min(kstep, nkc - kk);
    const int nbb = std::max(1, std::min(nbands, c->item_cap / kn));
    for (int b0 = 0; b0 < nbands; b0 += nbb) f(k0 + kk, kk, kn, b0, std::min(nbb, nbands - b0));
  }
}

// Algorithmic bytes per launch (compulsory reads + writes of each kernel's
// own inputs and outputs) are passed to the profiler with every launch.
template <typename R>
void derek_ctx::apply_h(int k0, int nkc, PwBlock<R> psi, PwBlock<R> hpsi, int nbands) {
  const Axis<R>* ax = axes<R>();
  const R* v = vlocal<R>();
  const double sc = sizeof(cx_t<R>);
  const int64_t nyz = (int64_t)dims[1] * dims[2];
  const bool fz = use_slab<R>(kFusedHpsi);
  for_item_launches(this, k0, nkc, nbands, [&](int ka, int kk, int kn, int b0, int nbb) {
    PwBatch B{d_maps, ka, kn, nbb, nst_max, nax_max, zsl};
    cx_t<R>* S = sticks<R>(kn * nbb);
    double npw = 0, nst = 0, nax = 0;
    for (int k = ka; k < ka + kn; ++k) {
      npw += (double)ks[k].npw;
      nst += ks[k].maps.nst;
      nax += ks[k].maps.nax;
    }
    const double pw_b = nbb * npw * sc, st_b = nbb * nst * dims[2] * sc, pl_b = nbb * nax * nyz * sc;
    PwBlock<R> pi{psi.p + kk * psi.ks + (int64_t)b0 * psi.ld, psi.ld, psi.ks};
    PwBlock<R> ho{hpsi.p + kk * hpsi.ks + (int64_t)b0 * hpsi.ld, hpsi.ld, hpsi.ks};
    timed(KC_STICKS_Z_INV, pw_b + 12.0 * npw + st_b, 0, [&] { sticks_z_inv<R>(B, pi, ax, dims[2], S, fz, s); });
    if (fz) {
      // bytes: sticks in and out once, V once per launch; flops: 5 n log2 n
      // per line (y lines on the active planes, x lines on every y) and the
      // V multiply, per item
      const double fy = 5.0 * dims[1] * std::log2((double)dims[1]), fx = 5.0 * dims[0] * std::log2((double)dims[0]);
      const double fl = nbb * (2.0 * nax * dims[2] * fy + 2.0 * kn * (double)dims[1] * dims[2] * fx) + 2.0 * nbb * kn * N;
      timed(KC_SLAB_VMUL, 2 * st_b + (double)N * sizeof(R), fl,
            [&] { slab_vmul<R>(B, S, ax, dims[2], vtrans<R>(), bpw, s); });
    } else {
      cx_t<R>* P = planes<R>(kn * nbb);
      timed(KC_PLANES_Y_INV, st_b + pl_b, 0, [&] { planes_y<R>(B, S, P, ax, dims[1], dims[2], +1, bpw, s); });
      timed(KC_PLANES_X_VMUL, 2 * pl_b + (double)N * sizeof(R), 0,
            [&] { planes_x_vmul<R>(B, P, ax, nyz, v, (R)(1.0 / (double)N), bpw, s); });
      timed(KC_PLANES_Y_FWD, st_b + pl_b, 0, [&] { planes_y<R>(B, S, P, ax, dims[1], dims[2], -1, bpw, s); });
    }
    timed(KC_STICKS_Z_FWD, st_b + 2 * pw_b + 20.0 * npw, 0,
          [&] { sticks_z_fwd<R>(B, S, ax, dims[2], pi, ho, fz, s); });
  });
}

template <typename R>
void derek_ctx::density(int k0, int nkc, PwBlock<R> psi, int nbands, const double* occw, int occ_ks, double* rho,
                        bool keep) {
  const Axis<R>* ax = axes<R>();
  const double sc = sizeof(cx_t<R>);
  const int64_t nyz = (int64_t)dims[1] * dims[2];
  const bool fz = use_slab<R>(kFusedDensity);
  bool first = true;
  for_item_launches(this, k0, nkc, nbands, [&](int ka, int kk, int kn, int b0, int nbb) {
    PwBatch B{d_maps, ka, kn, nbb, nst_max, nax_max, zsl};
    cx_t<R>* S = sticks<R>(kn * nbb);
    double npw = 0, nst = 0, nax = 0;
    for (int k = ka; k < ka + kn; ++k) {
      npw += (double)ks[k].npw;
      nst += ks[k].maps.nst;
      nax += ks[k].maps.nax;
    }
    const double pw_b = nbb * npw * sc, st_b = nbb * nst * dims[2] * sc, pl_b = nbb * nax * nyz * sc;
    const int kp = (keep || !first) ? 1 : 0;
    first = false;
    PwBlock<R> pi{psi.p + kk * psi.ks + (int64_t)b0 * psi.ld, psi.ld, psi.ks};
    const double* ow = occw + (int64_t)kk * occ_ks + b0;
    timed(KC_STICKS_Z_INV, pw_b + 12.0 * npw + st_b, 0, [&] { sticks_z_inv<R>(B, pi, ax, dims[2], S, fz, s); });
    if (fz) {
      const int groups = slab_density_groups<R>(B, ax, dims[2]);
      double* part = rho_part.get<double>((size_t)kn * groups * N);
      // sticks in once, the partial planes out and back, rho read (keep) and written
      timed(KC_SLAB_DENSITY, st_b + 2.0 * 8.0 * kn * groups * N + (1 + kp) * 8.0 * N, 0,
            [&] { slab_density<R>(B, S, ax, dims[2], ow, occ_ks, part, groups, rho, kp, s); });
    } else {
      cx_t<R>* P = planes<R>(kn * nbb);
      timed(KC_PLANES_Y_INV, st_b + pl_b, 0, [&] { planes_y<R>(B, S, P, ax, dims[1], dims[2], +1, bpw, s); });
      timed(KC_PLANES_X_DENSITY, pl_b + (1 + kp) * 8.0 * N, 0,
            [&] { planes_x_density<R>(B, P, ax, nyz, ow, occ_ks, rho, kp, s); });
    }
  });
}

template void derek_ctx::apply_h<double>(int, int, PwBlock<double>, PwBlock<double>, int);
template void derek_ctx::apply_h<float>(int, int, PwBlock<float>, PwBlock<float>, int);
template void derek_ctx::density<double>(int, int, PwBlock<double>, int, const double*, int, double*, bool);
template void derek_ctx::density<float>(int, int, PwBlock<float>, int, const double*, int, double*, bool);

// ------------------------------------------------------------------ grid ops
template <typename R> void derek_ctx::fft3d(cx_t<R>* g, int howmany, int dir) {
  const int sign = dir == 1 ? -1 : +1;
  const Axis<R>* ax = axes<R>();
  const double pass = 2.0 * (double)howmany * N * sizeof(cx_t<R>);
  timed(KC_FFT_Z, pass, 0, [&] { fft_z<R>(g, (int64_t)howmany * dims[0] * dims[1], ax[2], sign, s); });
  timed(KC_FFT_Y, pass, 0, [&] { fft_strided<R>(g, (int64_t)howmany * dims[0], dims[2], ax[1], sign, s); });
  timed(KC_FFT_X, pass, 0, [&] { fft_strided<R>(g, howmany, (int64_t)dims[1] * dims[2], ax[0], sign, s); });
}
template void derek_ctx::fft3d<double>(double2*, int, int);
template void derek_ctx::fft3d<float>(float2*, int, int);

void derek_ctx::dot_into(const double* a, const double* b, int64_t n, double scale, bool root, double* rep, int slot) {
  const int nbk = dot_nblocks(n);
  double* p = partial.get<double>(1024);
  dot_partials(a, b, p, n, s);
  finish_sum(p, nbk, scale, root, rep, slot, s);
}

void derek_ctx::hartree(const double* rho, double* vh_out, double* rep, int slot) {
  if (!gsq_set) throw StateError("hartree: |G|^2 not set (derek_hip_set_gsq)");
  real_to_cx(rho, cgrid, N, s);
  fft3d<double>(cgrid, 1, +1);
  gspace_scale(cgrid, gsq, 0.0, 1.0 / (double)N, 0, N, s);
  fft3d<double>(cgrid, 1, -1);
  cx_real(cgrid, vh_out, 1.0, N, s);
  if (rep) dot_into(vh_out, rho, N, 0.5 * volume / (double)N, false, rep, slot);
}

void derek_ctx::lda(const double* rho, double* vxc_out, double* rep, int slot) {
  double* p = rep ? partial.get<double>(1024) : nullptr;
  lda_xc(rho, vxc_out, p, N, s);
  if (rep) finish_sum(p, lda_partials(N), volume / (double)N, false, rep, slot, s);
}

void derek_ctx::kerker(const double* f, double* out, double q0) {
  real_to_cx(f, cgrid, N, s);
  fft3d<double>(cgrid, 1, +1);
  gspace_scale(cgrid, gsq, q0 * q0, 1.0 / (double)N, 1, N, s);
  fft3d<double>(cgrid, 1, -1);
  cx_real(cgrid, out, 1.0, N, s);
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

// ================================================================== C-ABI
extern "C" {

int32_t derek_hip_version(void) { return 2; }

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

int32_t derek_basis_std_index(const int32_t fine_dims[3], const int32_t std_dims[3], int64_t npw,
                              const int32_t* fft_index, int64_t* std_index) {
  return guard([&] {
    DEREK_REQUIRE(npw >= 0 && (npw == 0 || (fft_index && std_index)), "null argument");
    check_dims(fine_dims);
    check_dims(std_dims);
    const int f[3] = {fine_dims[0], fine_dims[1], fine_dims[2]};
    const int d[3] = {std_dims[0], std_dims[1], std_dims[2]};
    const int64_t ng = (int64_t)f[0] * f[1] * f[2];
    for (int64_t p = 0; p < npw; ++p) DEREK_REQUIRE(fft_index[p] >= 0 && fft_index[p] < ng, "FFT bin outside the grid");
    basis::std_index(f, d, npw, fft_index, std_index);
  });
}

// ---- .info input (host only)
struct derek_info {
  info::Params p;
};

static void copy_text(const std::string& t, char* buf, int64_t len) {
  DEREK_REQUIRE(buf && len > 0, "null buffer");
  const size_t n = std::min<size_t>((size_t)len - 1, t.size());
  std::memcpy(buf, t.data(), n);
  buf[n] = '\0';
}

int32_t derek_info_read(const char* path, const char* safety, derek_info** out) {
  return guard([&] {
    DEREK_REQUIRE(path && out, "null argument");
    *out = nullptr;
    auto h = std::make_unique<derek_info>();
    h->p = info::read(path, safety && *safety ? safety : "Normal");
    *out = h.release();
  });
}

int32_t derek_info_free(derek_info* h) {
  delete h;
  return DEREK_OK;
}

int32_t derek_info_num(const derek_info* h, const char* key, double* v) {
  return guard([&] {
    DEREK_REQUIRE(h && key && v, "null argument");
    auto it = h->p.num.find(key);
    if (it == h->p.num.end()) throw ArgError(std::string("no numeric key ") + key);
    for (size_t i = 0; i < it->second.size(); ++i) v[i] = it->second[i];
  });
}

int32_t derek_info_str(const derek_info* h, const char* key, char* buf, int64_t len) {
  return guard([&] {
    DEREK_REQUIRE(h && key, "null argument");
    auto it = h->p.str.find(key);
    if (it == h->p.str.end()) throw ArgError(std::string("no text key ") + key);
    copy_text(it->second, buf, len);
  });
}

int32_t derek_info_present(const derek_info* h, const char* key, int32_t* present) {
  return guard([&] {
    DEREK_REQUIRE(h && key && present, "null argument");
    *present = std::find(h->p.present.begin(), h->p.present.end(), std::string(key)) != h->p.present.end();
  });
}

int32_t derek_info_cell(const derek_info* h, double cell[9]) {
  return guard([&] {
    DEREK_REQUIRE(h && cell, "null argument");
    if (!h->p.cell_declared) throw ArgError(" No lattice provided");
    for (int i = 0; i < 9; ++i) cell[i] = h->p.cell[i];
  });
}

int32_t derek_info_wells(const derek_info* h, int32_t* n, double* centres, double* heights, double* widths) {
  return guard([&] {
    DEREK_REQUIRE(h && n, "null argument");
    const int32_t m = (int32_t)h->p.wells.size();
    if (*n > 0 && centres && heights && widths) {
      DEREK_REQUIRE(*n >= m, "well arrays too small");
      for (int32_t i = 0; i < m; ++i) {
        for (int a = 0; a < 3; ++a) centres[3 * i + a] = h->p.wells[i].c[a];
        heights[i] = h->p.wells[i].height;
        widths[i] = h->p.wells[i].width;
      }
    }
    *n = m;
  });
}

int32_t derek_info_log(const derek_info* h, char* buf, int64_t len) {
  return guard([&] {
    DEREK_REQUIRE(h, "null argument");
    copy_text(h->p.log, buf, len);
  });
}

int32_t derek_info_preset_pot(const derek_info* h, const int32_t fine_dims[3], double* v) {
  return guard([&] {
    DEREK_REQUIRE(h && v, "null argument");
    check_dims(fine_dims);
    const int f[3] = {fine_dims[0], fine_dims[1], fine_dims[2]};
    info::preset_potential(h->p, f, v);
  });
}

int32_t derek_info_help(const char* mode, const char* term, char* buf, int64_t len) {
  return guard([&] {
    DEREK_REQUIRE(mode, "null argument");
    const std::string m = mode, t = term ? term : "";
    if (m == "help") copy_text(info::help_key(t), buf, len);
    else if (m == "search") copy_text(info::search(t), buf, len);
    else if (m == "list") copy_text(info::list(), buf, len);
    else throw ArgError("help mode must be help, search or list");
  });
}

// ---- interchange files (host only)
int32_t derek_io_pot_write(const char* path, const int32_t std_dims[3], int64_t nfine, const double* nc_pot,
                           int32_t allocated) {
  return guard([&] {
    DEREK_REQUIRE(path, "null path");
    pot_write(path, std_dims, nfine, nc_pot, allocated != 0);
  });
}

int32_t derek_io_pot_read(const char* path, const int32_t std_dims[3], int64_t nfine, double* nc_pot,
                          int32_t* allocated) {
  return guard([&] {
    DEREK_REQUIRE(path, "null path");
    const bool a = pot_read(path, std_dims, nfine, nc_pot);
    if (allocated) *allocated = a ? 1 : 0;
  });
}

int32_t derek_io_den_write(const char* path, const int32_t std_dims[3], int64_t nfine, const double* charge,
                           const double* spin, int32_t allocated) {
  return guard([&] {
    DEREK_REQUIRE(path, "null path");
    den_write(path, std_dims, nfine, charge, spin, allocated != 0);
  });
}

int32_t derek_io_den_read(const char* path, const int32_t std_dims[3], int64_t nfine, double* charge, double* spin,
                          int32_t* allocated) {
  return guard([&] {
    DEREK_REQUIRE(path, "null path");
    const bool a = den_read(path, std_dims, nfine, charge, spin);
    if (allocated) *allocated = a ? 1 : 0;
  });
}

int32_t derek_io_wvfn_write(const char* path, const int32_t std_dims[3], int32_t nk, int32_t nb,
                            const int64_t* npw, const int64_t* std_index, const double* psi, int32_t allocated) {
  return guard([&] {
    DEREK_REQUIRE(path, "null path");
    wvfn_write(path, std_dims, nk, nb, npw, std_index, psi, allocated != 0);
  });
}

int32_t derek_io_wvfn_read(const char* path, const int32_t std_dims[3], int32_t nk, int32_t nb, const int64_t* npw,
                           const int64_t* std_index, double* psi, int32_t* allocated) {
  return guard([&] {
    DEREK_REQUIRE(path, "null path");
    const bool a = wvfn_read(path, std_dims, nk, nb, npw, std_index, psi);
    if (allocated) *allocated = a ? 1 : 0;
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
    DEREK_REQUIRE(d->nk >= 1 && d->nk <= 65535, "nk must be in 1..65535");
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
    // complex64 GEMMs: BF16 matrix cores with two bf16 terms per value
    // (about 16 significant bits) by default; derek_hip_set_gemm changes it
    c->gemm_bf16 = c->prec == DEREK_MIXED ? 2 : 0;
    c->volume = d->volume;
    c->max_npw = d->max_npw;
    DEREK_HIP(hipGetDevice(&c->device));
    DEREK_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
    noalloc_setup(c->device);
    for (int a = 0; a < 3; ++a) {
      c->tw64[a] = reinterpret_cast<double2*>(make_twiddles<double>(c->dims[a]));
      c->tw32[a] = reinterpret_cast<float2*>(make_twiddles<float>(c->dims[a]));
      c->ax64[a] = Axis<double>{c->dims[a], c->plan[a], c->tw64[a]};
      c->ax32[a] = Axis<float>{c->dims[a], c->plan[a], c->tw32[a]};
    }
    c->ks.resize(c->nk);
    c->d_maps = dmalloc<PwMaps>(c->nk);
    c->d_npw = dmalloc<int>(c->nk);
    c->d_wk = dmalloc<double>(c->nk);
    c->gsq = dmalloc<double>(c->N);
    c->vext = dmalloc<double>(c->N);
    c->vloc = dmalloc<double>(c->N);
    c->vloc32 = dmalloc<float>(c->N);
    c->vt64 = dmalloc<double>(c->N);
    c->vt32 = dmalloc<float>(c->N);
    c->cgrid = dmalloc<double2>(c->N);
    // (k-point, band) items per pipeline launch; the SCF autotuner may change it
    if (const char* b = std::getenv("DEREK_BAND_BATCH")) c->item_cap = std::max(1, std::atoi(b));
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
    if (c->scf) throw StateError("set_basis after scf_init");
    for (int64_t g = 0; g < npw; ++g)
      DEREK_REQUIRE(idx[g] >= 0 && idx[g] < c->N, "fft_index entry outside the grid");
    KState& K = c->ks[k];
    dfree(K.idx);
    dfree(K.kin);
    K.npw = npw;
    K.idx = dmalloc<int32_t>(npw);
    K.kin = dmalloc<double>(npw);
    DEREK_HIP(hipMemcpy(K.idx, idx, npw * sizeof(int32_t), hipMemcpyHostToDevice));
    DEREK_HIP(hipMemcpy(K.kin, kin, npw * sizeof(double), hipMemcpyHostToDevice));
    // sphere maps for the pruned FFT pipeline
    const PwMapsHost h = build_pw_maps(idx, npw, c->dims[0], c->dims[1], c->dims[2]);
    dfree(K.mapbuf);
    std::vector<int32_t> flat;
    for (const std::vector<int32_t>* v : {&h.perm, &h.pw_z, &h.pw_st, &h.st_pw, &h.st_y, &h.pl_st, &h.ax_x, &h.pl_y})
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
    q += h.nax;
    K.maps.pl_y = q;
    K.maps.kin = K.kin;
    K.weight = weight;
    K.set = true;
    const int n32 = (int)npw;
    DEREK_HIP(hipMemcpy(c->d_maps + k, &K.maps, sizeof(PwMaps), hipMemcpyHostToDevice));
    DEREK_HIP(hipMemcpy(c->d_npw + k, &n32, sizeof(int), hipMemcpyHostToDevice));
    DEREK_HIP(hipMemcpy(c->d_wk + k, &weight, sizeof(double), hipMemcpyHostToDevice));
    c->nst_max = 0;
    c->nax_max = 0;
    for (const KState& s : c->ks)
      if (s.set) {
        c->nst_max = std::max(c->nst_max, s.maps.nst);
        c->nax_max = std::max(c->nax_max, s.maps.nax);
      }
    // stick capacity per item: a multiple of 16, so that z-major stick rows
    // start on 128-byte lines
    c->nst_max = (int)(ceil_div(c->nst_max, 16) * 16);
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
    c->vloc_changed(true);
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

int32_t derek_hip_fft1d(derek_ctx* c, double* data, int32_t n, int64_t howmany, int32_t dir) {
  return guard([&] {
    DEREK_REQUIRE(c && data, "null argument");
    DEREK_REQUIRE(dir == 1 || dir == -1, "direction must be +1 or -1");
    DEREK_REQUIRE(howmany >= 1, "howmany must be >= 1");
    int a = -1;
    for (int q = 2; q >= 0; --q)
      if (c->dims[q] == n) a = q;
    DEREK_REQUIRE(a >= 0, "fft1d length " + std::to_string(n) + " is not one of the context's grid lengths");
    const size_t len = (size_t)howmany * n;
    const int sign = dir == 1 ? -1 : +1;
    double2* d = c->tA.get<double2>(len);
    DEREK_HIP(hipMemcpyAsync(d, data, len * sizeof(double2), hipMemcpyHostToDevice, c->s));
    if (c->prec == DEREK_FP64) {
      c->timed(KC_FFT_Z, 2.0 * len * sizeof(double2), 0, [&] { fft_z<double>(d, howmany, c->ax64[a], sign, c->s); });
    } else {
      float2* f = c->tB.get<float2>(len);
      convert<float, double>(f, d, (int64_t)len, c->s);
      c->timed(KC_FFT_Z, 2.0 * len * sizeof(float2), 0, [&] { fft_z<float>(f, howmany, c->ax32[a], sign, c->s); });
      convert<double, float>(d, f, (int64_t)len, c->s);
    }
    DEREK_HIP(hipMemcpyAsync(data, d, len * sizeof(double2), hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_wave_axpby(derek_ctx* c, int64_t n, const double a[2], const double* x, const double b[2],
                             const double* y, double* z) {
  return guard([&] {
    DEREK_REQUIRE(c && a && b && x && z, "null argument");
    DEREK_REQUIRE(n >= 1, "n must be >= 1");
    const bool use_y = b[0] != 0.0 || b[1] != 0.0;
    DEREK_REQUIRE(!use_y || y, "y is needed when b != 0");
    double2* dx = c->tA.get<double2>(n);
    double2* dy = use_y ? c->tB.get<double2>(n) : nullptr;
    double2* dz = c->tC.get<double2>(n);
    DEREK_HIP(hipMemcpyAsync(dx, x, n * sizeof(double2), hipMemcpyHostToDevice, c->s));
    if (use_y) DEREK_HIP(hipMemcpyAsync(dy, y, n * sizeof(double2), hipMemcpyHostToDevice, c->s));
    c->timed(KC_OTHER, (use_y ? 3.0 : 2.0) * n * sizeof(double2), 0, [&] {
      wave_axpby(dz, make_double2(a[0], a[1]), dx, make_double2(b[0], b[1]), dy, n, c->s);
    });
    DEREK_HIP(hipMemcpyAsync(z, dz, n * sizeof(double2), hipMemcpyDeviceToHost, c->s));
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
      c->apply_h<double>(k, 1, PwBlock<double>{dp, K.npw, 0}, PwBlock<double>{dh, K.npw, 0}, nb);
    } else {
      float2* fp = c->tC.get<float2>(n);
      float2* fh = c->tD.get<float2>(n);
      convert<float, double>(fp, dp, (int64_t)n, c->s);
      c->apply_h<float>(k, 1, PwBlock<float>{fp, K.npw, 0}, PwBlock<float>{fh, K.npw, 0}, nb);
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
    const int64_t wcap = overlap_work_elems(m, n, npw, 1);
    double2* w = c->ovwork.get<double2>(wcap);
    if (c->prec == DEREK_FP64) {
      c->timed(KC_OVERLAP, (double)(m + n) * npw * 16.0, 8.0 * npw * m * n, [&] {
        overlap<double>(dx, npw, 0, dy, npw, 0, m, n, npw, nullptr, ds, m, 0, 1, w, wcap, herm, c->s);
      });
    } else {
      // complex64 operands on the f32 MFMA path, as in the SCF's FP32 phase
      float2* fx = c->tD.get<float2>((size_t)(herm ? m : m + n) * npw);
      float2* fy = herm ? fx : fx + (size_t)m * npw;
      convert<float, double>(fx, dx, (int64_t)m * npw, c->s);
      if (!herm) convert<float, double>(fy, dy, (int64_t)n * npw, c->s);
      c->timed(KC_OVERLAP, (double)(m + n) * npw * 8.0, 8.0 * npw * m * n, [&] {
        overlap<float>(fx, npw, 0, fy, npw, 0, m, n, npw, nullptr, ds, m, 0, 1, w, wcap, herm, c->s, c->gemm_bf16, 0,
                       c->ov_rounds, c->ov_tile);
      });
    }
    DEREK_HIP(hipMemcpyAsync(S, ds, (size_t)m * n * sizeof(double2), hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_overlap_noalloc(derek_ctx* c, int64_t npw, int32_t m, int32_t n, const double* X, const double* Y,
                                  double* S) {
  return guard([&] {
    DEREK_REQUIRE(c && X && Y && S, "null argument");
    DEREK_REQUIRE(npw >= 1 && m >= 1 && n >= 1, "empty overlap");
    DEREK_REQUIRE(c->prec == DEREK_FP64, "overlap_noalloc multiplies complex128 operands: it needs a DEREK_FP64 context");
    overlap_noalloc(c, npw, m, n, reinterpret_cast<const double2*>(X), reinterpret_cast<const double2*>(Y),
                    reinterpret_cast<double2*>(S));
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
    if (c->prec == DEREK_FP64) {
      c->timed(KC_ROTATE, (double)(m + n) * npw * 16.0, 8.0 * npw * m * n,
               [&] { rotate<double>(dx, npw, 0, du, m, 0, m, n, npw, nullptr, dz, npw, 0, 1, c->s); });
    } else {
      float2* fx = c->tD.get<float2>((size_t)(m + n) * npw);
      float2* fz = fx + (size_t)m * npw;
      convert<float, double>(fx, dx, (int64_t)m * npw, c->s);
      c->timed(KC_ROTATE, (double)(m + n) * npw * 8.0, 8.0 * npw * m * n,
               [&] { rotate<float>(fx, npw, 0, du, m, 0, m, n, npw, nullptr, fz, npw, 0, 1, c->s, c->gemm_bf16); });
      convert<double, float>(dz, fz, (int64_t)n * npw, c->s);
    }
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
    if (c->prec == DEREK_FP64) {
      c->density<double>(k, 1, PwBlock<double>{dp, K.npw, 0}, nb, dw, nb, dr, accumulate != 0);
    } else {
      float2* fp = c->tC.get<float2>(n);
      convert<float, double>(fp, dp, (int64_t)n, c->s);
      c->density<float>(k, 1, PwBlock<float>{fp, K.npw, 0}, nb, dw, nb, dr, accumulate != 0);
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
    double* de = c->tD.get<double>(REP_COUNT);
    DEREK_HIP(hipMemcpyAsync(dr, rho, c->N * sizeof(double), hipMemcpyHostToDevice, c->s));
    c->hartree(dr, dv, de, REP_EH);
    double e = 0.0;
    DEREK_HIP(hipMemcpyAsync(vh, dv, c->N * sizeof(double), hipMemcpyDeviceToHost, c->s));
    DEREK_HIP(hipMemcpyAsync(&e, de + REP_EH, sizeof(double), hipMemcpyDeviceToHost, c->s));
    c->sync();
    if (eh) *eh = e;
  });
}

int32_t derek_hip_heev(derek_ctx* c, int32_t n, int32_t nev, const double* A, double* w, double* Z) {
  return guard([&] {
    DEREK_REQUIRE(c && A && w && Z, "null argument");
    DEREK_REQUIRE(n >= 1 && n <= 1000 && nev >= 1 && nev <= n, "need 1 <= nev <= n <= 1000");
    double2* dA = c->tA.get<double2>((size_t)n * n);
    double2* dZ = c->tB.get<double2>((size_t)n * nev);
    double* dw = c->tC.get<double>(n);
    double* work = c->tD.get<double>(heev_work_doubles(n, nev));
    DEREK_HIP(hipMemcpyAsync(dA, A, (size_t)n * n * sizeof(double2), hipMemcpyHostToDevice, c->s));
    hermitize_gpu(dA, n, 1, c->s);
    heev_gpu(dA, n, 0, n, nev, dw, 0, dZ, n, 0, work, 0, 1, c->s);
    DEREK_HIP(hipMemcpyAsync(w, dw, nev * sizeof(double), hipMemcpyDeviceToHost, c->s));
    DEREK_HIP(hipMemcpyAsync(Z, dZ, (size_t)n * nev * sizeof(double2), hipMemcpyDeviceToHost, c->s));
    c->sync();
  });
}

int32_t derek_hip_hegv(derek_ctx* c, int32_t m, int32_t nev, const double* S, const double* H, double* w, double* C) {
  return guard([&] {
    DEREK_REQUIRE(c && S && H && w && C, "null argument");
    DEREK_REQUIRE(m >= 1 && m <= 640 && nev >= 1 && nev <= m, "need 1 <= nev <= m <= 640");
    static thread_local DenseWork dw;
    dw.ensure(m, nev, 1);
    double2* dS = c->tA.get<double2>((size_t)m * m);
    double2* dH = c->tB.get<double2>((size_t)m * m);
    double2* dU = c->tC.get<double2>((size_t)m * nev);
    double* dw_ = c->tD.get<double>(nev);
    DEREK_HIP(hipMemcpyAsync(dS, S, (size_t)m * m * sizeof(double2), hipMemcpyHostToDevice, c->s));
    DEREK_HIP(hipMemcpyAsync(dH, H, (size_t)m * m * sizeof(double2), hipMemcpyHostToDevice, c->s));
    gen_solve(dw, dS, dH, m, nev, 1, 1e-8, dw_, dU, c->s);
    int ok = 0;
    DEREK_HIP(hipMemcpyAsync(&ok, dw.ok, sizeof(int), hipMemcpyDeviceToHost, c->s));
    DEREK_HIP(hipMemcpyAsync(w, dw_, nev * sizeof(double), hipMemcpyDeviceToHost, c->s));
    DEREK_HIP(hipMemcpyAsync(C, dU, (size_t)m * nev * sizeof(double2), hipMemcpyDeviceToHost, c->s));
    c->sync();
    if (!ok) throw NumericError("overlap has fewer than nev independent columns");
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

int32_t derek_hip_scf_init(derek_ctx* c, const derek_scf_params* p) {
  return guard([&] {
    DEREK_REQUIRE(c && p, "null argument");
    scf_init(c, *p);
  });
}

int32_t derek_hip_scf_step(derek_ctx* c, derek_scf_report* rep) {
  return guard([&] {
    DEREK_REQUIRE(c && rep, "null argument");
    scf_step(c, *rep);
  });
}

int32_t derek_hip_nscf(derek_ctx* c, int32_t max_iter, double res_tol, double* max_res) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    DEREK_REQUIRE(max_iter >= 1 && res_tol >= 0.0, "max_iter must be >= 1 and res_tol >= 0");
    const double r = scf_nscf(c, max_iter, res_tol);
    if (max_res) *max_res = r;
  });
}

int32_t derek_hip_scf_polish(derek_ctx* c, int32_t max_iter, double res_tol, double* max_res) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    DEREK_REQUIRE(max_iter >= 0 && res_tol >= 0.0, "max_iter and res_tol must be >= 0");
    const double r = scf_polish(c, max_iter, res_tol);
    if (max_res) *max_res = r;
  });
}

int32_t derek_hip_get_eigenvalues(derek_ctx* c, double* eig) {
  return guard([&] {
    DEREK_REQUIRE(c && eig, "null argument");
    scf_get(c, 0, eig);
  });
}

int32_t derek_hip_get_occupations(derek_ctx* c, double* occ) {
  return guard([&] {
    DEREK_REQUIRE(c && occ, "null argument");
    scf_get(c, 1, occ);
  });
}

int32_t derek_hip_get_residuals(derek_ctx* c, double* res) {
  return guard([&] {
    DEREK_REQUIRE(c && res, "null argument");
    scf_get(c, 4, res);
  });
}

int32_t derek_hip_get_density(derek_ctx* c, double* rho) {
  return guard([&] {
    DEREK_REQUIRE(c && rho, "null argument");
    scf_get(c, 2, rho);
  });
}

int32_t derek_hip_get_potential(derek_ctx* c, double* v) {
  return guard([&] {
    DEREK_REQUIRE(c && v, "null argument");
    scf_get(c, 3, v);
  });
}

int32_t derek_hip_get_psi(derek_ctx* c, int32_t k, double* psi) {
  return guard([&] {
    DEREK_REQUIRE(c && psi, "null argument");
    c->require_basis(k);
    scf_psi(c, k, psi, false);
  });
}

int32_t derek_hip_put_psi(derek_ctx* c, int32_t k, const double* psi) {
  return guard([&] {
    DEREK_REQUIRE(c && psi, "null argument");
    c->require_basis(k);
    scf_psi(c, k, const_cast<double*>(psi), true);
  });
}

int32_t derek_hip_apply_h_resident(derek_ctx* c, int32_t k, int32_t nb) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    if (k == -1) c->require_all_bases();
    else c->require_basis(k);
    if (!c->vloc_set) throw StateError("local potential not set");
    DEREK_REQUIRE(nb >= 1 && nb <= c->nb, "nb must be in 1..nbands");
    scf_apply_h_resident(c, k, nb);
  });
}

int32_t derek_hip_get_tuning(derek_ctx* c, int32_t out[8]) {
  return guard([&] {
    DEREK_REQUIRE(c && out, "null argument");
    out[0] = c->bpw;
    out[1] = c->item_cap;
    scf_tuning(c, out + 2);
    out[4] = c->fused;
    out[5] = c->zsl ? c->zsl : (c->prec == DEREK_FP64 ? 16 : 32);
    out[6] = c->ov_rounds;
    out[7] = c->ov_tile;
  });
}

int32_t derek_hip_set_launch(derek_ctx* c, int32_t bands_per_wg, int32_t items_per_launch, int32_t fused) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    DEREK_REQUIRE(bands_per_wg == 1 || bands_per_wg == 2 || bands_per_wg == 4 || bands_per_wg == 8,
                  "bands per workgroup must be 1, 2, 4 or 8");
    DEREK_REQUIRE(items_per_launch >= 1, "items per launch must be positive");
    DEREK_REQUIRE(fused >= 0 && fused <= 3, "fused must be 0..3 (bit 0: H|psi>, bit 1: density)");
    c->bpw = bands_per_wg;
    c->item_cap = items_per_launch;
    c->fused = fused;
    ++c->graph_epoch;
  });
}

int32_t derek_hip_set_tiles(derek_ctx* c, int32_t z_sticks, int32_t gemm_rounds, int32_t gemm_tile) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    DEREK_REQUIRE(z_sticks == 0 || z_sticks == 16 || z_sticks == 32, "sticks per z-pass workgroup must be 0, 16 or 32");
    DEREK_REQUIRE(gemm_rounds == 1 || gemm_rounds == 2 || gemm_rounds == 4, "GEMM rounds must be 1, 2 or 4");
    DEREK_REQUIRE(gemm_tile == 64 || gemm_tile == 128, "GEMM tile must be 64 or 128");
    c->zsl = z_sticks;
    c->ov_rounds = gemm_rounds;
    c->ov_tile = gemm_tile;
    ++c->graph_epoch;
  });
}

int32_t derek_hip_set_gemm(derek_ctx* c, int32_t bf16_terms) {
  return guard([&] {
    DEREK_REQUIRE(c, "null context");
    DEREK_REQUIRE(bf16_terms >= 0 && bf16_terms <= 3, "bf16 terms must be 0 (FP32 matrix cores) or 1..3");
    c->gemm_bf16 = bf16_terms;
    ++c->graph_epoch;
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
