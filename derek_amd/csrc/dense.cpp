#include "dense.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <stdexcept>
#include <string>

#include "common.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace derek {
namespace dense {

void hermitize(int n, cd* A, int lda) {
  for (int j = 0; j < n; ++j) {
    A[j + (size_t)j * lda] = cd(A[j + (size_t)j * lda].real(), 0.0);
    for (int i = j + 1; i < n; ++i) {
      const cd a = A[i + (size_t)j * lda], b = std::conj(A[j + (size_t)i * lda]);
      const cd m = 0.5 * (a + b);
      A[i + (size_t)j * lda] = m;
      A[j + (size_t)i * lda] = std::conj(m);
    }
  }
}

bool cholesky(int n, cd* A, int lda) {
  for (int j = 0; j < n; ++j) {
    double d = A[j + (size_t)j * lda].real();
    for (int k = 0; k < j; ++k) d -= std::norm(A[j + (size_t)k * lda]);
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    A[j + (size_t)j * lda] = d;
    const double id = 1.0 / d;
    // column j below the diagonal: A(i,j) = (A(i,j) - sum_k L(i,k) conj(L(j,k))) / d
    for (int k = 0; k < j; ++k) {
      const cd ljk = std::conj(A[j + (size_t)k * lda]);
      const cd* Lk = A + (size_t)k * lda;
      cd* Aj = A + (size_t)j * lda;
      for (int i = j + 1; i < n; ++i) Aj[i] -= Lk[i] * ljk;
    }
    for (int i = j + 1; i < n; ++i) A[i + (size_t)j * lda] *= id;
    for (int i = 0; i < j; ++i) A[i + (size_t)j * lda] = 0.0;
  }
  return true;
}

void inv_lower(int n, cd* L, int lda) {
  // Column by column: X = L^-1 is lower triangular; solve L X(:,j) = e_j.
  std::vector<cd> x(n);
  std::vector<cd> out((size_t)n * n, cd(0.0));
  for (int j = 0; j < n; ++j) {
    std::fill(x.begin(), x.end(), cd(0.0));
    x[j] = 1.0;
    for (int k = j; k < n; ++k) {
      // x[k] = (e_j[k] - sum_{i<k} L(k,i) x[i]) / L(k,k), done column-oriented below
      x[k] /= L[k + (size_t)k * lda];
      const cd xk = x[k];
      const cd* Lk = L + (size_t)k * lda;
      for (int i = k + 1; i < n; ++i) x[i] -= Lk[i] * xk;
    }
    for (int i = j; i < n; ++i) out[i + (size_t)j * n] = x[i];
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) L[i + (size_t)j * lda] = out[i + (size_t)j * n];
}

// Implicit QL with Wilkinson shifts on a real symmetric tridiagonal matrix
// (diagonal d, sub-diagonal e[i] = T(i+1,i), e[n-1] = 0).  Eigenvalues are
// left in d; every Givens rotation (plane i,i+1, cosine c, sine s) is
// recorded so that the eigenvectors can be formed afterwards row by row.
struct Givens {
  int i;
  double c, s;
};

static void tridiag_ql(int n, double* d, double* e, std::vector<Givens>* rot) {
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m != l) {
        if (++iter > 100) throw Error("hermitian eigensolver: QL iteration did not converge");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        bool deflated = false;
        for (i = m - 1; i >= l; --i) {
          const double f = s * e[i], b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            deflated = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          if (rot) rot->push_back({i, c, s});
        }
        if (deflated) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
}

static int nthreads(int work_rows) {
#ifdef _OPENMP
  int t = std::min(16, omp_get_max_threads());
  return std::max(1, std::min(t, work_rows / 16));
#else
  (void)work_rows;
  return 1;
#endif
}

void heev(int n, cd* A, int lda, double* w, cd* Z, int ldz) { heev_lowest(n, A, lda, n, w, Z, ldz); }

void heev_lowest(int n, cd* A, int lda, int nev, double* w, cd* Z, int ldz) {
  if (n <= 0 || nev <= 0) return;
  nev = std::min(nev, n);
  // Householder reduction A <- H_k A H_k, H_k = I - 2 v v^H acting on rows/cols k+1..n-1.
  std::vector<cd> V((size_t)n * n, cd(0.0));  // v_k stored in column k (rows k+1..)
  std::vector<cd> sub(n, cd(0.0));            // T(k+1,k)
  std::vector<cd> u(n), wv(n);
  for (int k = 0; k + 2 < n; ++k) {
    cd* col = A + (size_t)k * lda;
    double xn2 = 0.0;
    for (int i = k + 1; i < n; ++i) xn2 += std::norm(col[i]);
    const double xn = std::sqrt(xn2);
    const cd x0 = col[k + 1];
    const double ax0 = std::abs(x0);
    if (xn == 0.0 || xn2 - ax0 * ax0 <= 1e-300) {
      sub[k] = x0;  // already tridiagonal in this column
      continue;
    }
    const cd ph = ax0 > 0.0 ? x0 / ax0 : cd(1.0);
    const cd alpha = -ph * xn;
    cd* v = V.data() + (size_t)k * n;
    for (int i = k + 1; i < n; ++i) v[i] = col[i];
    v[k + 1] -= alpha;
    double vn2 = 0.0;
    for (int i = k + 1; i < n; ++i) vn2 += std::norm(v[i]);
    const double ivn = 1.0 / std::sqrt(vn2);
    for (int i = k + 1; i < n; ++i) v[i] *= ivn;
    sub[k] = alpha;
    const int nt = nthreads(n - k - 1);
    // u = A22 v, row i = sum_j A(i,j) v(j) = sum_j conj(A(j,i)) v(j) (Hermitian: read column i)
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
    for (int i = k + 1; i < n; ++i) {
      const cd* Ai = A + (size_t)i * lda;
      cd acc = 0.0;
      for (int j = k + 1; j < n; ++j) acc += std::conj(Ai[j]) * v[j];
      u[i] = acc;
    }
    double gam = 0.0;
    for (int i = k + 1; i < n; ++i) gam += (std::conj(v[i]) * u[i]).real();
    for (int i = k + 1; i < n; ++i) wv[i] = 2.0 * u[i] - 2.0 * gam * v[i];
    // A22 -= v w^H + w v^H
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
    for (int j = k + 1; j < n; ++j) {
      const cd cw = std::conj(wv[j]), cv = std::conj(v[j]);
      cd* Aj = A + (size_t)j * lda;
      for (int i = k + 1; i < n; ++i) Aj[i] -= v[i] * cw + wv[i] * cv;
    }
    col[k + 1] = alpha;
    for (int i = k + 2; i < n; ++i) col[i] = 0.0;
  }
  if (n >= 2) sub[n - 2] = A[(n - 1) + (size_t)(n - 2) * lda];
  // phases making the sub-diagonal real and non-negative
  std::vector<cd> ph(n, cd(1.0));
  std::vector<double> d(n), e(n, 0.0);
  for (int i = 0; i < n; ++i) d[i] = A[i + (size_t)i * lda].real();
  for (int k = 0; k + 1 < n; ++k) {
    const double a = std::abs(sub[k]);
    e[k] = a;
    ph[k + 1] = a > 0.0 ? ph[k] * (sub[k] / a) : ph[k];
  }
  std::vector<Givens> rot;
  rot.reserve((size_t)3 * n * n / 2 + 16);
  tridiag_ql(n, d.data(), e.data(), &rot);
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  // Eigenvectors of the real tridiagonal: start from the identity and replay
  // the rotations; each thread owns a block of rows (the column-pair updates
  // of one rotation stay contiguous and vectorise).
  std::vector<double> Zf((size_t)n * n, 0.0);  // column-major n x n
  for (int i = 0; i < n; ++i) Zf[i + (size_t)i * n] = 1.0;
  const int ntr = nthreads(n / 2);
#pragma omp parallel num_threads(ntr) if (ntr > 1)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num(), T = omp_get_num_threads();
#else
    const int t = 0, T = 1;
#endif
    const int r0 = (int)((int64_t)n * t / T), r1 = (int)((int64_t)n * (t + 1) / T);
    for (const Givens& g : rot) {
      double* zi = Zf.data() + (size_t)g.i * n;
      double* zj = zi + n;
      const double c = g.c, sn = g.s;
      for (int r = r0; r < r1; ++r) {
        const double fz = zj[r];
        zj[r] = sn * zi[r] + c * fz;
        zi[r] = c * zi[r] - sn * fz;
      }
    }
  }
  std::vector<double> Zr((size_t)nev * n);  // column j = eigenvector ord[j]
  for (int j = 0; j < nev; ++j) std::copy(Zf.begin() + (size_t)ord[j] * n, Zf.begin() + (size_t)(ord[j] + 1) * n,
                                          Zr.begin() + (size_t)j * n);
  for (int j = 0; j < nev; ++j) {
    w[j] = d[ord[j]];
    cd* zd = Z + (size_t)j * ldz;
    const double* zs = Zr.data() + (size_t)j * n;
    for (int i = 0; i < n; ++i) zd[i] = ph[i] * zs[i];
  }
  // back-transform: Z <- H_0 H_1 ... H_{n-3} Z, column by column
  const int ntc = nthreads(nev * 4);
#pragma omp parallel for num_threads(ntc) schedule(static) if (ntc > 1)
  for (int j = 0; j < nev; ++j) {
    cd* zj = Z + (size_t)j * ldz;
    for (int k = n - 3; k >= 0; --k) {
      const cd* v = V.data() + (size_t)k * n;
      if (v[k + 1] == cd(0.0)) {
        bool any = false;
        for (int i = k + 2; i < n && !any; ++i) any = v[i] != cd(0.0);
        if (!any) continue;
      }
      cd s = 0.0;
      for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * zj[i];
      s *= 2.0;
      for (int i = k + 1; i < n; ++i) zj[i] -= v[i] * s;
    }
  }
}

std::vector<int> cholesky_active(int n, const cd* S, int lda, double tol) {
  std::vector<cd> L((size_t)n * n, cd(0.0));  // row-major rows of L over kept columns
  std::vector<int> act;
  for (int j = 0; j < n; ++j) {
    const int c = (int)act.size();
    const cd* lj = L.data() + (size_t)j * n;
    double d = S[j + (size_t)j * lda].real();
    for (int q = 0; q < c; ++q) d -= std::norm(lj[q]);
    if (!(d > tol * S[j + (size_t)j * lda].real())) continue;
    const double dj = std::sqrt(d);
    L[(size_t)j * n + c] = dj;
    for (int i = j + 1; i < n; ++i) {
      cd v = S[i + (size_t)j * lda];
      const cd* li = L.data() + (size_t)i * n;
      for (int q = 0; q < c; ++q) v -= li[q] * std::conj(lj[q]);
      L[(size_t)i * n + c] = v / dj;
    }
    act.push_back(j);
  }
  return act;
}

bool hegv_lowest(int n, cd* S, cd* H, int ld, int nev, double* w, cd* C, int ldc) {
  hermitize(n, S, ld);
  hermitize(n, H, ld);
  if (!cholesky(n, S, ld)) return false;
  inv_lower(n, S, ld);  // S <- L^-1
  // T = L^-1 H ; A = T L^-H
  std::vector<cd> T((size_t)n * n, cd(0.0)), A((size_t)n * n, cd(0.0));
  const int ntp = nthreads(n * 4);
#pragma omp parallel for num_threads(ntp) schedule(static) if (ntp > 1)
  for (int j = 0; j < n; ++j) {
    cd* Tj = T.data() + (size_t)j * n;
    for (int k = 0; k < n; ++k) {
      const cd hkj = H[k + (size_t)j * ld];
      const cd* Lk = S + (size_t)k * ld;
      for (int i = k; i < n; ++i) Tj[i] += Lk[i] * hkj;
    }
  }
  // A(:,j) = sum_k T(:,k) conj(Linv(j,k)), Linv(j,k) nonzero for k <= j
#pragma omp parallel for num_threads(ntp) schedule(static) if (ntp > 1)
  for (int j = 0; j < n; ++j) {
    cd* Aj = A.data() + (size_t)j * n;
    for (int k = 0; k <= j; ++k) {
      const cd c = std::conj(S[j + (size_t)k * ld]);
      const cd* Tk = T.data() + (size_t)k * n;
      for (int i = 0; i < n; ++i) Aj[i] += Tk[i] * c;
    }
  }
  std::vector<double> ww(nev);
  std::vector<cd> Z((size_t)n * nev);
  heev_lowest(n, A.data(), n, nev, ww.data(), Z.data(), n);
  // C = L^-H Z; (L^-H)(i,k) = conj(Linv(k,i)), nonzero for k >= i
  const int nt = nthreads(nev * 4);
#pragma omp parallel for num_threads(nt) schedule(static) if (nt > 1)
  for (int j = 0; j < nev; ++j) {
    const cd* zj = Z.data() + (size_t)j * n;
    cd* cj = C + (size_t)j * ldc;
    for (int i = 0; i < n; ++i) {
      cd s = 0.0;
      for (int k = i; k < n; ++k) s += std::conj(S[k + (size_t)i * ld]) * zj[k];
      cj[i] = s;
    }
  }
  for (int j = 0; j < nev; ++j) w[j] = ww[j];
  return true;
}

}  // namespace dense
}  // namespace derek
