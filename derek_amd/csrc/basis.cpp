#include "basis.hpp"

#include <cmath>
#include <string>

#include "common.hpp"

namespace derek {
namespace basis {

static const double kPi = 3.141592653589793238462643383279502884197;  // constants.f90:29
static const double kTwoPi = 2.0 * kPi;                                 // constants.f90:30

// basis.f90:257-309: smallest n' >= n with prime factors in {2,3,5,7,11,13},
// up to 100 increments; each sweep divides by every prime at most once.
int prime_fact(int n) {
  static const int primes[6] = {2, 3, 5, 7, 11, 13};
  for (int k = 0; k < 100; ++k) {
    int np = n;
    int factors[6] = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < 100; ++j) {
      const int old = np;
      for (int i = 0; i < 6; ++i) {
        if (np % primes[i] == 0) {
          np /= primes[i];
          factors[i] += 1;
        }
      }
      if (old == np) break;
    }
    long long tot = 1;
    for (int i = 0; i < 6; ++i)
      for (int f = 0; f < factors[i]; ++f) tot *= primes[i];
    if (tot == n) break;
    n += 1;
  }
  return n;
}

static double col_norm(const double c[9], int j) {
  // io.f90:1956-1966: sqrt(cell(1,j)**2 + cell(2,j)**2 + cell(3,j)**2)
  const double a = c[0 + 3 * j], b = c[1 + 3 * j], d = c[2 + 3 * j];
  return std::sqrt(a * a + b * b + d * d);
}

void grid_dims(const double cell[9], double ecut, double g, int s[3], int f[3]) {
  const double k_cut = std::sqrt(2.0 * ecut);  // basis.f90:85
  for (int j = 0; j < 3; ++j) {
    const double l = col_norm(cell, j);
    s[j] = prime_fact((int)std::ceil(k_cut * l / kPi));
    f[j] = prime_fact((int)std::ceil(g * k_cut * l / kPi));
  }
}

void inv_cell(const double cell[9], double inv[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      DEREK_REQUIRE(i == j || cell[i + 3 * j] == 0.0,
                    "only orthorhombic cells have a LAPACK-independent inverse");
  // transpose(twopi * io_inv(cell)), io.f90:1954; DGETRI of a diagonal
  // matrix gives ONE/a_jj.
  for (int k = 0; k < 9; ++k) inv[k] = 0.0;
  for (int j = 0; j < 3; ++j) inv[j + 3 * j] = kTwoPi * (1.0 / cell[j + 3 * j]);
}

double volume(const double cell[9]) {
  // io_volume, io.f90:1983-1989 (a, b, c are the columns)
  const double* a = cell;
  const double* b = cell + 3;
  const double* c = cell + 6;
  return c[0] * (a[1] * b[2] - a[2] * b[1]) + c[1] * (b[0] * a[2] - a[2] * b[0]) +
         c[2] * (a[0] * b[1] - a[1] * b[0]);
}

void g_map(const double inv[9], const int dims[3], double* g, int32_t* m) {
  const int64_t N = (int64_t)dims[0] * dims[1] * dims[2];
  int64_t ig = 0;
  for (int mx = -(dims[0] / 2); mx <= dims[0] - dims[0] / 2 - 1; ++mx)
    for (int my = -(dims[1] / 2); my <= dims[1] - dims[1] / 2 - 1; ++my)
      for (int mz = -(dims[2] / 2); mz <= dims[2] - dims[2] / 2 - 1; ++mz) {
        const double t[3] = {(double)mx, (double)my, (double)mz};
        for (int i = 0; i < 3; ++i) {
          double acc = inv[i + 0] * t[0];
          acc = acc + inv[i + 3] * t[1];
          acc = acc + inv[i + 6] * t[2];
          if (g) g[ig + N * i] = acc;
        }
        if (m) {
          m[ig] = mx;
          m[ig + N] = my;
          m[ig + 2 * N] = mz;
        }
        ++ig;
      }
}

void real_map(const double cell[9], const int dims[3], double* r, double* frac) {
  const int64_t N = (int64_t)dims[0] * dims[1] * dims[2];
  int64_t ig = 0;
  for (int i = 1; i <= dims[0]; ++i)
    for (int j = 1; j <= dims[1]; ++j)
      for (int k = 1; k <= dims[2]; ++k) {
        const double f[3] = {(double)i * 1.0 / (double)dims[0], (double)j * 1.0 / (double)dims[1],
                             (double)k * 1.0 / (double)dims[2]};
        for (int c = 0; c < 3; ++c) {
          double acc = cell[c + 0] * f[0];
          acc = acc + cell[c + 3] * f[1];
          acc = acc + cell[c + 6] * f[2];
          if (r) r[ig + N * c] = acc;
          if (frac) frac[ig + N * c] = f[c];
        }
        ++ig;
      }
}

int kpoints(const int mp[3], double* kf) {
  const int nk = mp[0] * mp[1] * mp[2];
  DEREK_REQUIRE(nk > 0, "invalid number of kpoints");
  const double dk[3] = {1.0 / (double)mp[0], 1.0 / (double)mp[1], 1.0 / (double)mp[2]};
  int ik = 0;
  for (int x = 1; x <= mp[0]; ++x)
    for (int y = 1; y <= mp[1]; ++y)
      for (int z = 1; z <= mp[2]; ++z) {
        if (kf) {
          kf[ik + nk * 0] = x * dk[0] - (1.0 + dk[0]) / 2;
          kf[ik + nk * 1] = y * dk[1] - (1.0 + dk[1]) / 2;
          kf[ik + nk * 2] = z * dk[2] - (1.0 + dk[2]) / 2;
        }
        ++ik;
      }
  return nk;
}

static inline int64_t bin(int m, int n) { return ((m % n) + n) % n; }

int64_t sphere(const double inv[9], const int dims[3], const double kf[3], double ecut, int64_t cap,
               int32_t* fft_index, double* kin, int32_t* mo) {
  double kc[3];
  for (int i = 0; i < 3; ++i) kc[i] = inv[i + 0] * kf[0] + inv[i + 3] * kf[1] + inv[i + 6] * kf[2];
  int64_t cnt = 0;
  for (int mx = -(dims[0] / 2); mx <= dims[0] - dims[0] / 2 - 1; ++mx)
    for (int my = -(dims[1] / 2); my <= dims[1] - dims[1] / 2 - 1; ++my)
      for (int mz = -(dims[2] / 2); mz <= dims[2] - dims[2] / 2 - 1; ++mz) {
        const double t[3] = {(double)mx, (double)my, (double)mz};
        double gk[3];
        for (int i = 0; i < 3; ++i) {
          double acc = inv[i + 0] * t[0];
          acc = acc + inv[i + 3] * t[1];
          acc = acc + inv[i + 6] * t[2];
          gk[i] = acc + kc[i];
        }
        double g2 = gk[0] * gk[0];
        g2 = g2 + gk[1] * gk[1];
        g2 = g2 + gk[2] * gk[2];
        const double e = 0.5 * g2;
        if (!(e <= ecut)) continue;
        if (cnt < cap) {
          if (fft_index)
            fft_index[cnt] = (int32_t)((bin(mx, dims[0]) * dims[1] + bin(my, dims[1])) * dims[2] + bin(mz, dims[2]));
          if (kin) kin[cnt] = e;
          if (mo) {
            mo[3 * cnt + 0] = mx;
            mo[3 * cnt + 1] = my;
            mo[3 * cnt + 2] = mz;
          }
        }
        ++cnt;
      }
  return cnt;
}

void gsq(const double inv[9], const int dims[3], double* out) {
  for (int mx = -(dims[0] / 2); mx <= dims[0] - dims[0] / 2 - 1; ++mx)
    for (int my = -(dims[1] / 2); my <= dims[1] - dims[1] / 2 - 1; ++my)
      for (int mz = -(dims[2] / 2); mz <= dims[2] - dims[2] / 2 - 1; ++mz) {
        const double t[3] = {(double)mx, (double)my, (double)mz};
        double gv[3];
        for (int i = 0; i < 3; ++i) {
          double acc = inv[i + 0] * t[0];
          acc = acc + inv[i + 3] * t[1];
          acc = acc + inv[i + 6] * t[2];
          gv[i] = acc;
        }
        double g2 = gv[0] * gv[0];
        g2 = g2 + gv[1] * gv[1];
        g2 = g2 + gv[2] * gv[2];
        out[(bin(mx, dims[0]) * dims[1] + bin(my, dims[1])) * dims[2] + bin(mz, dims[2])] = g2;
      }
}

void gaussian_pot(const double cell[9], const int dims[3], int nwells, const double* centres,
                  const double* heights, const double* widths, double cos_amp, double* v) {
  const int64_t N = (int64_t)dims[0] * dims[1] * dims[2];
  std::vector<double> r((size_t)N * 3);
  real_map(cell, dims, r.data(), nullptr);
  double L[3];
  for (int a = 0; a < 3; ++a) L[a] = cell[a + 3 * a];
  for (int64_t p = 0; p < N; ++p) v[p] = 0.0;
  std::vector<double> fac((size_t)N), s((size_t)N);
  for (int w = 0; w < nwells; ++w) {
    const double h = heights[w], wd = widths[w];
    for (int64_t p = 0; p < N; ++p) fac[p] = 1.0;
    for (int a = 0; a < 3; ++a) {
      const int nimg = (int)std::ceil(std::sqrt(92.0) * wd / L[a]) + 1;
      const double c = centres[3 * w + a] * L[a];
      for (int64_t p = 0; p < N; ++p) s[p] = 0.0;
      for (int j = -nimg; j <= nimg; ++j)
        for (int64_t p = 0; p < N; ++p) {
          const double x = (r[p + N * a] - c) - j * L[a];
          s[p] += std::exp(-(x * x) / (2.0 * wd * wd));
        }
      for (int64_t p = 0; p < N; ++p) fac[p] *= s[p];
    }
    for (int64_t p = 0; p < N; ++p) v[p] += h * fac[p];
  }
  if (cos_amp != 0.0)
    for (int64_t p = 0; p < N; ++p) v[p] += cos_amp * std::cos(2.0 * kPi * r[p] / L[0]);
}

void std_index(const int fine[3], const int std_dims[3], int64_t n, const int32_t* fft_index, int64_t* out) {
  const int64_t nyz = (int64_t)fine[1] * fine[2];
  for (int64_t p = 0; p < n; ++p) {
    const int64_t b = fft_index[p];
    const int64_t bin[3] = {b / nyz, (b / fine[2]) % fine[1], b % fine[2]};
    int64_t pos = 0;
    bool inside = true;
    for (int a = 0; a < 3; ++a) {
      // FFT bin -> Miller index in [-n/2, n - n/2 - 1] of the fine grid
      const int64_t m = bin[a] < fine[a] - fine[a] / 2 ? bin[a] : bin[a] - fine[a];
      const int64_t lo = -(std_dims[a] / 2), hi = std_dims[a] - std_dims[a] / 2 - 1;
      if (m < lo || m > hi) inside = false;
      pos = pos * std_dims[a] + (m - lo);
    }
    out[p] = inside ? pos : -1;
  }
}

}  // namespace basis
}  // namespace derek
