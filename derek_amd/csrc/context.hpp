// The library context (derek_ctx): grid, per-k-point spheres, device
// buffers of the pruned FFT pipeline, profiler, and the batched SCF engine.
#pragma once
#include <algorithm>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/derek_hip.h"
#include "dense_gpu.hpp"
#include "launch.hpp"
#include "pw_fft.hpp"
#include "scf_kernels.hpp"

namespace derek {

// Device buffer that grows on demand.  Every free or reallocation bumps
// *epoch (when set): captured SCF graphs hold these addresses, and the
// engine drops its graphs when the epoch it captured them at has moved.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  uint64_t* epoch = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      if (epoch) ++*epoch;
    }
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T* get(size_t n) {
    const size_t need = std::max<size_t>(n, 1) * sizeof(T);
    if (need > bytes) {
      release();
      DEREK_HIP(hipMalloc(&p, need));
      bytes = need;
      if (epoch) ++*epoch;
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

inline int64_t round_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// Allocation-free overlap (noalloc.hip): a file-scope device array of this
// many double2 (32 MiB per GPU) holds its split-K partial sums.
constexpr int64_t kNoAllocScratchElems = (int64_t)1 << 21;

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
  KC_DENSE,
  KC_OTHER,
  KC_SLAB_VMUL,
  KC_SLAB_DENSITY,
  KC_COUNT
};

struct Profiler {
  struct Pending {
    int cls;
    hipEvent_t a, b;
    double bytes, flops;
  };
  bool on = false;
  bool paused = false;  // no events while a graph is being captured
  std::vector<hipEvent_t> pool;
  std::vector<Pending> pending;
  int64_t count[KC_COUNT] = {};
  double ms[KC_COUNT] = {}, bytes[KC_COUNT] = {}, flops[KC_COUNT] = {};
  hipEvent_t take();
  void resolve();
  void reset();
  ~Profiler();
};

struct KState {
  int64_t npw = 0;
  int32_t* idx = nullptr;
  double* kin = nullptr;
  double weight = 0.0;
  bool set = false;
  PwMaps maps;                  // sphere maps (device pointers into mapbuf)
  int32_t* mapbuf = nullptr;
};

struct ScfEngine;
struct EngineDeleter {
  void operator()(ScfEngine* e) const;  // defined with the engine (scf.hip)
};

}  // namespace derek

struct derek_ctx {
  int device = 0;  // HIP device current at derek_hip_init
  int dims[3] = {0, 0, 0};
  int64_t N = 0;
  int nk = 0, nb = 0, prec = DEREK_FP64;
  double volume = 0.0;
  int64_t max_npw = 0;
  hipStream_t s = nullptr;
  derek::FftPlan plan[3];
  double2* tw64[3] = {nullptr, nullptr, nullptr};
  float2* tw32[3] = {nullptr, nullptr, nullptr};
  derek::Axis<double> ax64[3];
  derek::Axis<float> ax32[3];
  std::vector<derek::KState> ks;
  derek::PwMaps* d_maps = nullptr;  // device table of every k-point's maps
  int* d_npw = nullptr;             // [nk]
  double* d_wk = nullptr;           // [nk] k-point weights
  int nst_max = 0, nax_max = 0;     // largest stick count (rounded up to 16) / plane count over the k-points
  double* gsq = nullptr;
  double* vext = nullptr;
  double* vloc = nullptr;
  float* vloc32 = nullptr;
  // vloc / N with x and z swapped, [z][y][x], FP64 and FP32, for the fused y/x pass
  double* vt64 = nullptr;
  float* vt32 = nullptr;
  bool gsq_set = false, vloc_set = false, vext_set = false;
  // launch shapes (autotuned by the SCF engine when asked)
  int bpw = 2;         // bands per y/x-pass workgroup
  int item_cap = 256;  // (k-point, band) items per pipeline launch
  // fused y/x passes on z-major sticks where the grid allows them: bit 0 for
  // H|psi> (kFusedHpsi), bit 1 for the density (kFusedDensity)
  int fused = 3;
  // complex64 GEMMs (overlaps, rotations): 0 on the FP32 matrix cores,
  // 1..3 on the BF16 ones with each value split into that many bf16 terms
  // (set to 2 at init for DEREK_MIXED contexts)
  int gemm_bf16 = 0;
  // sticks per workgroup of the z passes (16 or 32; 0: 16 for complex128,
  // 32 for complex64) and the rounds of resident workgroups the overlaps'
  // row chunks fill (1, 2 or 4); both autotuned by the SCF engine when asked
  int zsl = 0;
  int ov_rounds = 2;
  // C tile edge of the BF16 overlaps (64 or 128), autotuned as well
  int ov_tile = 64;
  // Moves whenever a work buffer below is freed or reallocated, or a launch
  // setting changes (set_launch, set_gemm): a graph captured at another
  // epoch may hold freed addresses or stale shapes and is not replayed.
  uint64_t graph_epoch = 0;
  // pruned-FFT work buffers for item_cap items
  derek::DevBuf S64, P64, S32, P32;
  derek::DevBuf rho_part;  // partial densities of the fused density pass
  template <typename R> derek::cx_t<R>* sticks(int items);
  template <typename R> derek::cx_t<R>* planes(int items);
  double2* cgrid = nullptr;
  derek::DevBuf tA, tB, tC, tD, ovwork, partial;
  derek::Profiler prof;
  std::unique_ptr<derek::ScfEngine, derek::EngineDeleter> scf;
  template <typename F> void timed(int cls, double bytes, double flops, F&& launch) {
    if (!prof.on || prof.paused) {
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

  derek_ctx();
  ~derek_ctx();
  template <typename R> const derek::Axis<R>* axes() const;
  template <typename R> const R* vlocal() const;
  template <typename R> const R* vtrans() const;
  static constexpr int kFusedHpsi = 1, kFusedDensity = 2;
  template <typename R> bool use_slab(int part) const;
  // after vloc changed: its FP32 copy (when f32) and the transposed copies
  void vloc_changed(bool f32);
  void require_basis(int k) const;
  void require_all_bases() const;
  void sync() { DEREK_HIP(hipStreamSynchronize(s)); }

  // Pipelines over the items (k, b) for k = k0..k0+nkc-1, b < nbands; psi
  // and hpsi blocks as in PwBlock.  Launches are split so that no launch has
  // more than item_cap items.
  template <typename R>
  void apply_h(int k0, int nkc, derek::PwBlock<R> psi, derek::PwBlock<R> hpsi, int nbands);
  template <typename R>
  void density(int k0, int nkc, derek::PwBlock<R> psi, int nbands, const double* occw, int occ_ks, double* rho,
               bool keep);
  template <typename R> void fft3d(derek::cx_t<R>* g, int howmany, int dir);
  // Grid operations on device buffers (FP64); energies land in rep[slot]
  // when rep is given (no host synchronisation).
  void hartree(const double* rho, double* vh_out, double* rep, int slot);
  void lda(const double* rho, double* vxc_out, double* rep, int slot);
  void dot_into(const double* a, const double* b, int64_t n, double scale, bool root, double* rep, int slot);
  void kerker(const double* f, double* out, double q0);
};

namespace derek {
// resolves the allocation-free overlap's scratch address on device `dev`
// (current); once per device
void noalloc_setup(int dev);
// S = X^H Y (device pointers, complex128) on the context's stream, partial
// sums in the scratch array; no allocation, no synchronisation
void overlap_noalloc(derek_ctx* c, int64_t npw, int m, int n, const double2* X, const double2* Y, double2* S);
}  // namespace derek
