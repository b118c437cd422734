"""Spin-unpolarised LDA: Slater exchange + Perdew-Zunger 1981 correlation.

The reference's ``module xc`` (Source/xc.f90:18-26) only imports libxc; the
functional it names is libxc's XC_LDA_X (id 1) + XC_LDA_C_PZ (id 9).  The
formulas are the published ones restated from the libxc sources:
``ExLib/libxc/maple/lda_exc/lda_x.mpl`` with ``maple/util.mpl:97``
(eps_x = -(3/4)(3/pi)^(1/3) n^(1/3)), and
``ExLib/libxc/maple/lda_exc/lda_c_pz.mpl`` with parameters
``ExLib/libxc/src/lda_c_pz.c:51-59`` (paramagnetic set, rs >= 1 branch
``gamma/(1+beta1*sqrt(rs)+beta2*rs)``, else ``a ln rs + b + c rs ln rs + d rs``).
Screening as libxc: points with n < 1e-15 are skipped (output 0;
src/work_lda_inc.c:47-52, threshold src/lda_c_pz.c:102) and exchange is zero
when n/2 <= 1e-15 (src/maple2c/lda_exc/lda_x.c, ``t2``).

Pinned by libxc's own regression data (testsuite/regression/lda_x,
lda_c; 16 digits) and the DFT-repository data (testsuite/df_repo), committed
as fixtures under tests/golden/.
"""
import numpy as np

DENS_THRESHOLD = 1e-15
PZ = dict(gamma=-0.1423, beta1=1.0529, beta2=0.3334,
          a=0.0311, b=-0.048, c=0.0020, d=-0.0116)


def lda_x(n):
    """(eps_x, v_x) per point."""
    n = np.asarray(n, dtype=np.float64)
    eps = np.zeros_like(n)
    v = np.zeros_like(n)
    ok = (n >= DENS_THRESHOLD) & (n / 2.0 > DENS_THRESHOLD)
    c = -0.75 * (3.0 / np.pi) ** (1.0 / 3.0)
    n13 = np.cbrt(n[ok])
    eps[ok] = c * n13
    v[ok] = (4.0 / 3.0) * c * n13
    return eps, v


def lda_c_pz(n):
    """(eps_c, v_c) per point, v_c = eps_c - (rs/3) d eps_c / d rs."""
    n = np.asarray(n, dtype=np.float64)
    eps = np.zeros_like(n)
    v = np.zeros_like(n)
    ok = n >= DENS_THRESHOLD
    rs = np.cbrt(3.0 / (4.0 * np.pi * np.maximum(n[ok], DENS_THRESHOLD)))
    p = PZ
    hi = rs >= 1.0
    e = np.empty_like(rs)
    vv = np.empty_like(rs)
    s = np.sqrt(rs[hi])
    den = 1.0 + p["beta1"] * s + p["beta2"] * rs[hi]
    e[hi] = p["gamma"] / den
    vv[hi] = e[hi] * (1.0 + (7.0 / 6.0) * p["beta1"] * s + (4.0 / 3.0) * p["beta2"] * rs[hi]) / den
    lo = ~hi
    r = rs[lo]
    lr = np.log(r)
    e[lo] = p["a"] * lr + p["b"] + p["c"] * r * lr + p["d"] * r
    vv[lo] = (p["a"] * lr + (p["b"] - p["a"] / 3.0) + (2.0 / 3.0) * p["c"] * r * lr
              + ((2.0 * p["d"] - p["c"]) / 3.0) * r)
    eps[ok] = e
    v[ok] = vv
    return eps, v


def lda_xc(n):
    ex, vx = lda_x(n)
    ec, vc = lda_c_pz(n)
    return ex + ec, vx + vc
