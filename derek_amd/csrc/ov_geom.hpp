// Launch geometry of the split-K overlap C = X^H Y and the size of its
// partial-sum buffer.  Plain C++ (no HIP): the GEMM kernels take OvGeom by
// value, the SCF engine sizes its buffer with ov_engine_work_elems, and a
// host test checks that buffer against every overlap the engine launches.
#pragma once
#include <algorithm>
#include <cstdint>

namespace derek {

constexpr int OV_T = 64;    // C tile edge (the BF16 kernels also take 128)
constexpr int OV_KS = 16;   // rows per LDS slab (FP64 / FP32 matrix cores)
constexpr int BF_KS = 32;   // rows per slab on the BF16 matrix cores (one MFMA)

struct OvGeom {
  int m, n, tm, tn, ntiles, nchunk;
  int64_t npw, kch;
  bool herm;
  int t;  // tile edge
};

inline int64_t ov_ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Row chunks per tile: at least `rounds` full rounds of workgroups over the
// chip's resident slots (`slots` = resident workgroups per CU x CUs), then,
// among the next few counts, the one that leaves the smallest idle tail in
// the last round; every chunk keeps at least 256 rows and a whole number of
// `ks`-row slabs.  The SCF autotuner picks `rounds` among kOvRounds.
constexpr int kOvRounds[3] = {1, 2, 4};
inline OvGeom ov_geom(int m, int n, int64_t npw, bool herm, int batch, int slots, int ks = OV_KS, int rounds = 2,
                      int tile = OV_T) {
  OvGeom g{};
  g.m = m;
  g.n = n;
  g.npw = npw;
  g.herm = herm;
  g.t = tile;
  g.tm = (int)ov_ceil_div(m, tile);
  g.tn = (int)ov_ceil_div(n, tile);
  g.ntiles = herm ? g.tm * (g.tm + 1) / 2 : g.tm * g.tn;
  const int64_t units = (int64_t)g.ntiles * std::max(1, batch);
  const int64_t cmax = std::max<int64_t>(1, npw / 256);
  const int64_t c0 = std::min(cmax, std::max<int64_t>(1, ov_ceil_div((int64_t)rounds * slots, units)));
  int64_t best = c0;
  double best_eff = -1.0;
  for (int64_t c = c0; c <= std::min(cmax, c0 + 16); ++c) {
    const int64_t w = units * c;
    const double eff = (double)w / (double)(ov_ceil_div(w, (int64_t)slots) * slots);
    if (eff > best_eff + 1e-9) {
      best_eff = eff;
      best = c;
    }
  }
  int64_t kch = ov_ceil_div(std::max<int64_t>(npw, 1), best);
  kch = std::max<int64_t>(ks, ov_ceil_div(kch, ks) * ks);
  g.kch = kch;
  g.nchunk = (int)std::max<int64_t>(1, ov_ceil_div(npw, kch));
  return g;
}

// At most max_chunks row chunks (a fixed partial-sum buffer): each chunk
// takes more rows, still a whole number of `ks`-row slabs.
inline void ov_limit_chunks(OvGeom& g, int64_t max_chunks, int ks) {
  if (max_chunks < 1 || g.nchunk <= max_chunks) return;
  int64_t kch = ov_ceil_div(std::max<int64_t>(g.npw, 1), max_chunks);
  kch = std::max<int64_t>(ks, ov_ceil_div(kch, ks) * ks);
  g.kch = kch;
  g.nchunk = (int)std::max<int64_t>(1, ov_ceil_div(g.npw, kch));
}

// double2 elements of the partial sums of an m x n overlap of `batch`
// k-points, for any of the kernels whose resident slots are listed (both
// slab depths, Hermitian or not, every choice of rounds, both tile edges).
inline int64_t ov_work_elems(int m, int n, int64_t npw, int batch, const int* slots, int nslots) {
  int64_t chunks = 1;
  for (int i = 0; i < nslots; ++i)
    for (int ks : {OV_KS, BF_KS})
      for (bool h : {false, m == n})
        for (int rounds : kOvRounds)
          for (int tile : {OV_T, 2 * OV_T})
            chunks = std::max<int64_t>(chunks, ov_geom(m, n, npw, h, batch, slots[i], ks, rounds, tile).nchunk);
  return (int64_t)batch * chunks * m * n;
}

// The SCF engine's buffer: every overlap it launches is nb x nb or
// 2nb x 2nb (Hermitian or not) over a full batch of kc k-points or the last,
// shorter batch of nk mod kc.
inline int64_t ov_engine_work_elems(int nb, int nk, int kc, int64_t npw, const int* slots, int nslots) {
  int64_t cap = 0;
  for (int b : {kc, nk % kc})
    if (b > 0)
      for (int m : {nb, 2 * nb}) cap = std::max(cap, ov_work_elems(m, m, npw, b, slots, nslots));
  return cap;
}

}  // namespace derek
