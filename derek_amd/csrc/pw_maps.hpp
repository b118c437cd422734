// Plane-wave sphere maps for the pruned FFT pipelines (host side).
#pragma once
#include <cstdint>
#include <vector>

namespace derek {

// Device pointers describing one k-point's sphere on the FFT grid.  The
// plane waves are visited in stick order p = 0..npw-1 (perm[p] is the
// position in the caller's order); sticks are sorted by (x bin, y bin).
struct PwMaps {
  int npw = 0;
  int nst = 0;                    // sticks (distinct (x, y) bins)
  int nax = 0;                    // active x planes (distinct x bins)
  const int32_t* perm = nullptr;  // [npw] caller index of stick-ordered plane wave p
  const int32_t* pw_z = nullptr;  // [npw] z bin of p
  const int32_t* pw_st = nullptr; // [npw] stick of p
  const int32_t* st_pw = nullptr; // [nst+1] first plane wave of each stick
  const int32_t* st_y = nullptr;  // [nst] y bin of each stick
  const int32_t* pl_st = nullptr; // [nax+1] first stick of each active plane
  const int32_t* ax_x = nullptr;  // [nax] x bin of each active plane
  const int32_t* pl_y = nullptr;  // [nax*ny] stick at (active plane a, y bin), -1 if none
  const double* kin = nullptr;    // [npw] kinetic energy 1/2 |G+k|^2, caller order
};

// One launch of a pruned pipeline covers the items (k, band) for
// k = k0 .. k0+nkc-1 and nbb bands of each.  `mt` is a device array of the
// maps of every k-point; the stick and plane buffers hold nst_cap sticks and
// nax_cap planes per item.
struct PwBatch {
  const PwMaps* mt = nullptr;
  int k0 = 0, nkc = 1, nbb = 1;
  int nst_cap = 0, nax_cap = 0;
  // sticks per workgroup of the compiled z passes (16 or 32; 0: 16 for
  // complex128, 32 for complex64)
  int zsl = 0;
};

// Host construction of the maps from flat FFT-bin indices (z fastest).
struct PwMapsHost {
  int npw = 0, nst = 0, nax = 0;
  std::vector<int32_t> perm, pw_z, pw_st, st_pw, st_y, pl_st, ax_x, pl_y;
};
PwMapsHost build_pw_maps(const int32_t* fft_index, int64_t npw, int nx, int ny, int nz);

}  // namespace derek
