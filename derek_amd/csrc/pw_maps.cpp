// Host construction of the plane-wave sphere maps used by pw_fft.hip.
#include <algorithm>
#include <numeric>
#include <string>

#include "common.hpp"
#include "pw_maps.hpp"

namespace derek {

PwMapsHost build_pw_maps(const int32_t* idx, int64_t npw, int nx, int ny, int nz) {
  (void)nx;
  PwMapsHost h;
  h.npw = (int)npw;
  std::vector<int32_t> ord(npw);
  std::iota(ord.begin(), ord.end(), 0);
  // stick order: by (x bin, y bin), then z bin; the flat index is already
  // x-major, y, z-minor, so sorting by it gives exactly that order
  std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return idx[a] < idx[b]; });
  for (int64_t p = 1; p < npw; ++p)
    DEREK_REQUIRE(idx[ord[p]] != idx[ord[p - 1]], "fft_index holds the same grid point twice");
  h.perm = ord;
  h.pw_z.resize(npw);
  h.pw_st.resize(npw);
  const int64_t nyz = (int64_t)ny * nz;
  int64_t prev_col = -1, prev_x = -1;
  for (int64_t p = 0; p < npw; ++p) {
    const int64_t f = idx[ord[p]];
    const int64_t col = f / nz;  // x*ny + y
    const int64_t x = f / nyz;
    h.pw_z[p] = (int32_t)(f % nz);
    if (col != prev_col) {
      if (x != prev_x) {
        h.pl_st.push_back(h.nst);
        h.ax_x.push_back((int32_t)x);
        h.nax += 1;
        prev_x = x;
      }
      h.st_pw.push_back((int32_t)p);
      h.st_y.push_back((int32_t)(col % ny));
      h.nst += 1;
      prev_col = col;
    }
    h.pw_st[p] = h.nst - 1;
  }
  h.st_pw.push_back((int32_t)npw);
  h.pl_st.push_back(h.nst);
  h.pl_y.assign((size_t)h.nax * ny, -1);
  for (int a = 0; a < h.nax; ++a)
    for (int st = h.pl_st[a]; st < h.pl_st[a + 1]; ++st) h.pl_y[(size_t)a * ny + h.st_y[st]] = st;
  return h;
}

}  // namespace derek
