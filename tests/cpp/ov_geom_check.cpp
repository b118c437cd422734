// Host check of the SCF engine's overlap partial-sum buffer: for each
// system shape, every overlap the engine launches must fit the buffer that
// ov_engine_work_elems sizes.  The launches are listed here from the
// engine's call sites, independently of the sizing function:
//   scf_init           X^H X            nb x nb, Hermitian, per batch
//   Rayleigh-Ritz      X^H (HX)         nb x nb, Hermitian
//                      X^H X (FP32)     nb x nb, Hermitian
//   Davidson, full     Y^H Y, Y^H (HY)  2nb x 2nb, Hermitian
//   Davidson, blocks   X^H R, X^H (HR)  nb x nb, general
//                      R^H R, R^H (HR)  nb x nb, Hermitian
// each over a full batch of kc k-points and the last batch of nk mod kc,
// on any of the five overlap kernels (FP64, FP32, BF16 with 1..3 terms:
// slabs of 16 or 32 rows), with each choice of chunk rounds (1, 2, 4) and
// tile edge (64, 128), for resident-slot counts of 1..8 workgroups per CU
// on 256 CUs.  Prints one line per failure; exit status 1 if any.
#include <cstdio>
#include <vector>

#include "../../derek_amd/csrc/ov_geom.hpp"

using namespace derek;

struct Sys {
  const char* name;
  int nb, nk, kc;
  long long npw;
};

int main() {
  const Sys systems[] = {
      {"300 bands, 8 k, 144^3", 300, 8, 8, 161214}, {"300 bands, 8 k, one k per batch", 300, 8, 1, 161214},
      {"300 bands, 8 k, batches of 3", 300, 8, 3, 161214}, {"80 bands, Gamma, 108^3", 80, 1, 1, 62669},
      {"24 bands, 64 k, 72^3", 24, 64, 64, 17529}, {"24 bands, 64 k, batches of 10", 24, 64, 10, 17529},
      {"48 bands, 64 k, 72^3", 48, 64, 64, 17529}, {"128 bands, Gamma", 128, 1, 1, 5000},
      {"10 bands, Gamma, 32^3", 10, 1, 1, 1503}, {"8 bands, 4 k, tiny", 8, 4, 4, 300},
      {"320 bands, 1 k, tiny basis", 320, 1, 1, 700}};
  int bad = 0, checked = 0;
  for (const Sys& s : systems) {
    for (int per_cu = 1; per_cu <= 8; ++per_cu) {
      // the five kernels may differ in occupancy: all at one value, and
      // mixed sets
      std::vector<std::vector<int>> slot_sets = {{256 * per_cu},
                                                 {256 * per_cu, 256, 512, 768, 256 * (9 - per_cu)}};
      for (const auto& slots : slot_sets) {
        const long long cap = ov_engine_work_elems(s.nb, s.nk, s.kc, s.npw, slots.data(), (int)slots.size());
        struct L {
          int m, n;
          bool herm;
        };
        const L launches[] = {{s.nb, s.nb, true}, {2 * s.nb, 2 * s.nb, true}, {s.nb, s.nb, false}};
        for (int b : {s.kc, s.nk % s.kc}) {
          if (b <= 0) continue;
          for (const L& l : launches)
            for (int sl : slots)
              for (int ks : {OV_KS, BF_KS})
                for (int rounds : kOvRounds)
                  for (int tile : {64, 128}) {
                const OvGeom g = ov_geom(l.m, l.n, s.npw, l.herm, b, sl, ks, rounds, tile);
                const long long need = (long long)b * g.nchunk * l.m * l.n;
                ++checked;
                if (need > cap) {
                  ++bad;
                  std::printf("FAIL %s: %dx%d herm=%d batch=%d slots=%d ks=%d rounds=%d tile=%d needs %lld > %lld\n",
                              s.name, l.m, l.n, (int)l.herm, b, sl, ks, rounds, tile, need, cap);
                }
              }
        }
      }
    }
  }
  std::printf("%d overlap launches checked, %d do not fit\n", checked, bad);
  return bad ? 1 : 0;
}
