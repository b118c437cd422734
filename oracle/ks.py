"""Plane-wave Kohn-Sham pieces the reference leaves as stubs.

* ``apply_h``   - H|psi> = 1/2|G+k|^2 psi + FFT[V(r) * FFT^-1[psi]] / N
                  (``pot_apply_wfn`` stub, Source/pot.f90:344-382; FFT
                  conventions fft.f90:222-303).  Pinned by a dense
                  H_GG' = T_G delta_GG' + V((G-G') mod n) built on a tiny
                  basis and by V=const / V=0 closed forms.
* ``density``   - rho(r) = sum_k w_k sum_n f_nk |psi_nk(r)|^2 / Omega
                  (``density_calculate`` stub, Source/density.f90:101-131).
                  Pinned by int rho = N_e, rho >= 0 and the single-plane-wave
                  uniform density.
* ``hartree``   - V_H(G) = 4 pi rho(G)/|G|^2, G=0 -> 0 (absent in the
                  reference).  Pinned by a single cosine mode and a periodic
                  Gaussian charge (closed forms).
* ``overlap`` / ``rotate`` - Psi^H Phi and Psi U (``wave_dot_*`` stubs,
                  Source/wave.f90:929-1036); pinned against explicit loops.
* ``overlap_bf16`` / ``rotate_bf16`` - the same products with complex64
                  operands on BF16 matrix cores: each FP32 value split into
                  nt bf16 terms (round to nearest even), the term products
                  with a + b < nt kept, three real products per complex one.
                  No reference counterpart (a GPU precision option); pinned
                  by bf16 rounding known answers, torch's bfloat16 cast, and
                  the nt = 3 product sitting within FP32 error of the exact
                  product while nt = 1 does not.
* ``gaussian_wells`` - the benchmark's external potential at the reference's
                  ``real_fine_grid_points`` (basis.f90:175-188), wells as in
                  the ``$DATA gaussian_pot`` block (io.f90:1888-1922).
"""
import numpy as np

from . import fft as _fft
from . import xc as _xc


def gaussian_wells(R, lengths, centres, heights, widths, cos_amp=0.0):
    """V(r) = sum_i h_i sum_R exp(-|r - c_i L - R|^2 / (2 w_i^2))
    + cos_amp * cos(2 pi x / L_x), periodic images summed per axis until the
    image term is below exp(-46)."""
    lengths = np.asarray(lengths, dtype=np.float64)
    V = np.zeros(R.shape[0])
    for c, h, w in zip(np.asarray(centres, dtype=np.float64), heights, widths):
        fac = np.ones(R.shape[0])
        for a in range(3):
            L = lengths[a]
            nimg = int(np.ceil(np.sqrt(92.0) * w / L)) + 1
            d = R[:, a] - c[a] * L
            s = np.zeros(R.shape[0])
            for j in range(-nimg, nimg + 1):
                x = d - j * L
                s += np.exp(-(x * x) / (2.0 * w * w))
            fac *= s
        V += h * fac
    if cos_amp != 0.0:
        V += cos_amp * np.cos(2.0 * np.pi * R[:, 0] / lengths[0])
    return V


def to_grid(psi, fft_index, dims):
    """Scatter plane-wave coefficients (nb, npw) into (nb, N) grids."""
    N = int(np.prod(dims))
    g = np.zeros((psi.shape[0], N), dtype=np.complex128)
    g[:, fft_index] = psi
    return g


def apply_h(psi, fft_index, kin, v_r, dims):
    psi = np.atleast_2d(psi)
    N = int(np.prod(dims))
    grid = _fft.fft3d(to_grid(psi, fft_index, dims), dims, -1)
    grid *= v_r[None, :]
    grid = _fft.fft3d(grid, dims, 1) / float(N)
    return grid[:, fft_index] + kin[None, :] * psi


def apply_h_dense(npw_g_int, kin, v_r, dims):
    """Dense H_GG' for brute-force checks: V((G-G') mod n) from FFT(V)/N."""
    N = int(np.prod(dims))
    vg = _fft.fft3d(v_r.astype(np.complex128), dims, 1).reshape(dims) / float(N)
    M = np.asarray(npw_g_int)
    d = M[:, None, :] - M[None, :, :]
    H = vg[d[..., 0] % dims[0], d[..., 1] % dims[1], d[..., 2] % dims[2]]
    return H + np.diag(kin)


def density(psis, occs, weights, fft_indices, dims, volume):
    N = int(np.prod(dims))
    rho = np.zeros(N)
    for psi, occ, w, idx in zip(psis, occs, weights, fft_indices):
        g = _fft.fft3d(to_grid(psi, idx, dims), dims, -1)
        rho += w * np.einsum("b,bn->n", np.asarray(occ, dtype=np.float64), np.abs(g) ** 2)
    return rho / volume


def hartree(rho, gsq, dims, volume):
    """Returns (V_H(r), E_H) with E_H = 1/2 (Omega/N) sum V_H rho."""
    N = int(np.prod(dims))
    rg = _fft.fft3d(rho.astype(np.complex128), dims, 1) / float(N)
    fac = np.zeros(N)
    nz = gsq > 0.0
    fac[nz] = 4.0 * np.pi / gsq[nz]
    vh = _fft.fft3d(rg * fac, dims, -1).real
    eh = 0.5 * volume / N * float(np.dot(vh, rho))
    return vh, eh


def xc_on_grid(rho, dims, volume):
    """Returns (V_xc(r), E_xc, eps_xc(r))."""
    N = int(np.prod(dims))
    eps, v = _xc.lda_xc(rho)
    return v, volume / N * float(np.dot(rho, eps)), eps


def overlap(X, Y):
    """S[i,j] = <X_i|Y_j> for row-stacked bands (nb, npw)."""
    return np.conj(X) @ Y.T


def rotate(X, U):
    """Y_j = sum_i U[i,j] X_i."""
    return U.T @ X


def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def bf16_split(x, nt):
    """x (float32) = t_0 + t_1 + ... + t_{nt-1} + remainder, t_j bfloat16."""
    r = np.asarray(x, dtype=np.float32).copy()
    terms = []
    for _ in range(nt):
        t = bf16_round(r)
        terms.append(t)
        r = (r - t).astype(np.float32)
    return terms


def _split_product(A, B, nt):
    """sum over a + b < nt of A_a^T-style products, in float64:
    A (p, K) and B (K, q) float32 -> (p, q)."""
    As = bf16_split(A, nt)
    Bs = bf16_split(B, nt)
    out = 0.0
    for s_ in range(nt):
        for a in range(s_ + 1):
            out = out + As[a].astype(np.float64) @ Bs[s_ - a].astype(np.float64)
    return out


def overlap_bf16(X, Y, nt):
    """conj(X) Y^T with complex64 X (m, npw), Y (n, npw): planes
    xr, xi, xr - xi and yr, yi, yr + yi (float32), each product split into
    nt bf16 terms; Re = P1 + P2, Im = P3 - P1 + P2."""
    X = np.asarray(X, dtype=np.complex64)
    Y = np.asarray(Y, dtype=np.complex64)
    xr, xi = X.real.astype(np.float32), X.imag.astype(np.float32)
    yr, yi = Y.real.astype(np.float32), Y.imag.astype(np.float32)
    p1 = _split_product(xr, yr.T, nt)
    p2 = _split_product(xi, yi.T, nt)
    p3 = _split_product((xr - xi).astype(np.float32), (yr + yi).astype(np.float32).T, nt)
    return (p1 + p2) + 1j * (p3 - p1 + p2)


def rotate_bf16(X, U, nt):
    """U^T X with complex64 X (m, npw) and U (m, n) rounded to complex64:
    planes xr, xi, xr + xi and ur, ui, ur + ui, nt bf16 terms;
    Re = P1 - P2, Im = P3 - P1 - P2."""
    X = np.asarray(X, dtype=np.complex64)
    U = np.asarray(U, dtype=np.complex64)
    xr, xi = X.real.astype(np.float32), X.imag.astype(np.float32)
    ur, ui = U.real.astype(np.float32), U.imag.astype(np.float32)
    p1 = _split_product(ur.T, xr, nt)
    p2 = _split_product(ui.T, xi, nt)
    p3 = _split_product((ur + ui).astype(np.float32).T, (xr + xi).astype(np.float32), nt)
    return (p1 - p2) + 1j * (p3 - p1 - p2)


EKIN_FLOOR = 0.1


def tpa_precondition(R, kin, ekin):
    """Teter-Payne-Allan preconditioner, x = T_G / max(<T>_n, 0.1 Ha).

    The floor keeps a (near) G = 0 band's correction from collapsing onto
    its own plane wave."""
    x = kin[None, :] / np.maximum(ekin, EKIN_FLOOR)[:, None]
    p = 27.0 + x * (18.0 + x * (12.0 + 8.0 * x))
    return R * (p / (p + 16.0 * x ** 4))


def fermi_occupations(eigs, weights, nelec, kT, spin_deg=2.0):
    """Occupations f_nk in [0, spin_deg] and the Fermi level.

    kT == 0: the lowest nelec/spin_deg bands at every k are filled (insulator
    convention).  kT > 0: Fermi-Dirac, mu by bisection on sum_k w_k f_nk = nelec.
    Returns (occ, mu, -TS).
    """
    eigs = np.asarray(eigs)
    nk, nb = eigs.shape
    if kT <= 0.0:
        nocc = int(round(nelec / spin_deg))
        occ = np.zeros((nk, nb))
        occ[:, :nocc] = spin_deg
        return occ, float(np.max(eigs[:, nocc - 1])), 0.0
    w = np.asarray(weights)[:, None]
    lo, hi = float(eigs.min()) - 50 * kT, float(eigs.max()) + 50 * kT

    def count(mu):
        x = np.clip((eigs - mu) / kT, -200.0, 200.0)
        return float(np.sum(w * spin_deg / (1.0 + np.exp(x))))

    for _ in range(200):
        mu = 0.5 * (lo + hi)
        if count(mu) < nelec:
            lo = mu
        else:
            hi = mu
    mu = 0.5 * (lo + hi)
    x = np.clip((eigs - mu) / kT, -200.0, 200.0)
    f = 1.0 / (1.0 + np.exp(x))
    occ = spin_deg * f
    fs = np.clip(f, 1e-300, 1.0)
    gs = np.clip(1.0 - f, 1e-300, 1.0)
    ent = -np.sum(w * spin_deg * (f * np.log(fs) + (1.0 - f) * np.log(gs)))
    return occ, mu, -kT * float(ent)
