"""FP64 Kohn-Sham SCF restated in numpy: the loop of ``electronic_scf``
(Source/electronic.f90:86-175; stop when |E_i - E_{i-1}| < energy_tol,
electronic.f90:156-165) with the physics the reference leaves as stubs.

Per SCF step, for every k: H X, Rayleigh-Ritz, then ``n_david`` block-Davidson
steps on [X, K R] (K = Teter-Payne-Allan preconditioner) with Cholesky of the
2nb x 2nb overlap and a Hermitian eigensolve; occupations (fixed or
Fermi-Dirac); rho_out; E_tot = E_band - int V_Hxc,in rho_out + E_H[rho_out]
+ E_xc[rho_out] - TS; Pulay mixing of rho with a Kerker-preconditioned step.

Pinned by: non-interacting runs against dense diagonalisation of H_GG'
(scipy/numpy eigh) and closed-shell jellium (V_ext = 0), whose exact solution
is uniform rho, V_H = 0, eps_n = 1/2|G|^2 + v_xc(N_e/Omega),
E = sum_n f_n 1/2|G_n|^2 + Omega n eps_xc(n) (tests/test_oracle_scf.py).
"""
import numpy as np

from . import basis as _basis
from . import ks as _ks
from . import rng as _rng


class System:
    """All inputs of one SCF: an orthorhombic cell, cut-off, k-mesh, electrons,
    bands, smearing, Gaussian wells (fractional centres, heights (Ha, signed),
    widths (bohr)), an optional cosine term along x and the wavefunction seed."""

    def __init__(self, lengths, ecut, g_fine_scale, mp, nelec, nbands, kT=0.0,
                 centres=(), heights=(), widths=(), cos_amp=0.0, seed=1):
        self.cell = np.diag(np.asarray(lengths, dtype=np.float64))
        self.lengths = [float(x) for x in lengths]
        self.ecut = float(ecut)
        self.g_fine_scale = float(g_fine_scale)
        self.mp = tuple(int(x) for x in mp)
        self.nelec = float(nelec)
        self.nbands = int(nbands)
        self.kT = float(kT)
        self.centres = np.asarray(centres, dtype=np.float64).reshape(-1, 3)
        self.heights = np.asarray(heights, dtype=np.float64)
        self.widths = np.asarray(widths, dtype=np.float64)
        self.cos_amp = float(cos_amp)
        self.seed = int(seed)
        self.setup()

    def setup(self):
        self.inv_cell = _basis.inv_cell_orthorhombic(self.cell)
        self.std_dims, self.dims = _basis.grid_dims(self.cell, self.ecut, self.g_fine_scale)
        self.N = int(np.prod(self.dims))
        self.volume = _basis.volume(self.cell)
        self.G, self.M = _basis.g_map(self.inv_cell, self.dims)
        self.R, self.F = _basis.real_map(self.cell, self.dims)
        self.kfrac = _basis.kpoint_grid(self.mp)
        self.nk = self.kfrac.shape[0]
        self.weights = np.full(self.nk, 1.0 / self.nk)
        self.fft_index, self.kin, self.gk, self.Mk = [], [], [], []
        for ik in range(self.nk):
            kc = _basis.kcart(self.inv_cell, self.kfrac[ik])
            idx, kin, gk = _basis.sphere(self.G, self.M, self.dims, kc, self.ecut)
            self.fft_index.append(idx)
            self.kin.append(kin)
            self.gk.append(gk)
            gk2 = self.G + kc[None, :]
            sel = 0.5 * ((gk2[:, 0] * gk2[:, 0] + gk2[:, 1] * gk2[:, 1]) + gk2[:, 2] * gk2[:, 2]) <= self.ecut
            self.Mk.append(self.M[sel])
        self.gsq = _basis.gsq_on_fft_grid(self.inv_cell, self.dims)
        self.vext = _ks.gaussian_wells(self.R, self.lengths, self.centres, self.heights,
                                       self.widths, self.cos_amp)

    def initial_psi(self, ik):
        """(N(0,1) + i N(0,1)) / (1 + 1/2|G+k|^2) per band, then Cholesky-orthonormalised."""
        npw = len(self.kin[ik])
        X = np.empty((self.nbands, npw), dtype=np.complex128)
        for b in range(self.nbands):
            X[b] = _rng.normal_pairs(self.seed, _rng.psi_stream(ik, b), npw) / (1.0 + self.kin[ik])
        return orthonormalise(X)


def orthonormalise(X):
    S = _ks.overlap(X, X)
    S = 0.5 * (S + S.conj().T)
    L = np.linalg.cholesky(S)
    U = np.linalg.inv(L).conj().T
    return _ks.rotate(X, U)


def _herm(A):
    return 0.5 * (A + A.conj().T)


def rayleigh_ritz(X, HX):
    lam, Z = np.linalg.eigh(_herm(_ks.overlap(X, HX)))
    return _ks.rotate(X, Z), _ks.rotate(HX, Z), lam


def cholesky_active(S, tol=1e-8):
    """Columns of the (unit-diagonal) overlap S kept by a Cholesky that skips
    every column whose remaining pivot is below tol (numerically dependent on
    the columns already kept), in order."""
    n = S.shape[0]
    L = np.zeros((n, n), dtype=np.complex128)
    act = []
    for j in range(n):
        c = len(act)
        lj = L[j, :c].copy()
        d = S[j, j].real - float(np.vdot(lj, lj).real)
        if not d > tol * S[j, j].real:
            continue
        dj = np.sqrt(d)
        L[j, c] = dj
        L[j + 1:, c] = (S[j + 1:, j] - L[j + 1:, :c] @ np.conj(lj)) / dj
        act.append(j)
    return np.array(act, dtype=np.int64)


def davidson(X, HX, kin, apply, n_iter, res_tol=0.0, drop_tol=1e-10):
    """Returns (X, HX, eps, residual norms).  X must be orthonormal, HX = H X.

    Bands whose residual norm is below ``drop_tol`` get no correction vector;
    corrections numerically dependent on the vectors before them are dropped
    from the subspace problem (``cholesky_active``)."""
    nb = X.shape[0]
    X, HX, eps = rayleigh_ritz(X, HX)
    rn = np.linalg.norm(HX - eps[:, None] * X, axis=1)
    for _ in range(n_iter):
        if rn.max() < res_tol:
            break
        R = HX - eps[:, None] * X
        act = rn > drop_tol
        if not np.any(act):
            break
        ekin = np.einsum("g,bg->b", kin, np.abs(X) ** 2)
        W = _ks.tpa_precondition(R[act], kin, ekin[act])
        HW = apply(W)
        Y = np.vstack([X, W])
        HY = np.vstack([HX, HW])
        S = _herm(_ks.overlap(Y, Y))
        Hm = _herm(_ks.overlap(Y, HY))
        d = 1.0 / np.sqrt(np.real(np.diag(S)))
        S = _herm(d[:, None] * S * d[None, :])
        Hm = _herm(d[:, None] * Hm * d[None, :])
        a = cholesky_active(S)
        Sa, Ha = S[np.ix_(a, a)], Hm[np.ix_(a, a)]
        L = np.linalg.cholesky(Sa)
        Li = np.linalg.inv(L)
        lam, Z = np.linalg.eigh(_herm(Li @ Ha @ Li.conj().T))
        C = np.zeros((Y.shape[0], nb), dtype=np.complex128)
        C[a] = d[a, None] * (Li.conj().T @ Z[:, :nb])
        X = _ks.rotate(Y, C)
        HX = _ks.rotate(HY, C)
        eps = lam[:nb]
        rn = np.linalg.norm(HX - eps[:, None] * X, axis=1)
    return X, HX, eps, rn


class Pulay:
    """Pulay (DIIS) density mixing with a Kerker-preconditioned step:
    rho_next = sum_i a_i rho_in,i + beta * K[sum_i a_i F_i],
    F = rho_out - rho_in, a = A^-1 1 / (1^T A^-1 1), A_ij = <F_i|F_j>,
    K(G) = G^2/(G^2 + q0^2) (G = 0 removed)."""

    def __init__(self, system, beta=0.5, q0=1.0, history=8):
        self.s = system
        self.beta, self.q0sq, self.m = beta, q0 * q0, history
        self.rin, self.res = [], []

    def kerker(self, F):
        from . import fft as _fft
        s = self.s
        g = _fft.fft3d(F.astype(np.complex128), s.dims, 1)
        g *= s.gsq / (s.gsq + self.q0sq)
        return _fft.fft3d(g, s.dims, -1).real / s.N

    def __call__(self, rho_in, rho_out):
        F = rho_out - rho_in
        self.rin.append(rho_in.copy())
        self.res.append(F)
        if len(self.rin) > self.m:
            self.rin.pop(0)
            self.res.pop(0)
        n = len(self.res)
        A = np.array([[np.dot(self.res[i], self.res[j]) for j in range(n)] for i in range(n)])
        c = solve_pulay(A)
        r_opt = sum(ci * ri for ci, ri in zip(c, self.rin))
        f_opt = sum(ci * fi for ci, fi in zip(c, self.res))
        return r_opt + self.beta * self.kerker(f_opt)


def solve_pulay(A):
    """a = A^-1 1 normalised to sum 1; A scaled by its largest diagonal and
    regularised by 1e-12 on the diagonal."""
    n = A.shape[0]
    s = np.max(np.diag(A))
    if s <= 0.0:
        return np.full(n, 1.0 / n)
    c = np.linalg.solve(A / s + 1e-12 * np.eye(n), np.ones(n))
    return c / np.sum(c)


def scf(s, max_scf=40, energy_tol=1e-7, rho_tol=None, n_david_first=4, n_david=2,
        beta=0.5, q0=1.0, history=8, final_res_tol=None, log=None):
    """Run the SCF; returns a dict with energies, eigenvalues, occupations,
    density, per-step history and the converged flag."""
    X = [s.initial_psi(ik) for ik in range(s.nk)]
    rho_in = np.full(s.N, s.nelec / s.volume)
    mixer = Pulay(s, beta=beta, q0=q0, history=history)
    e_hist = None
    hist = []
    converged = False
    for iscf in range(1, max_scf + 1):
        vh, _ = _ks.hartree(rho_in, s.gsq, s.dims, s.volume)
        vxc, _, _ = _ks.xc_on_grid(rho_in, s.dims, s.volume)
        v_in = s.vext + vh + vxc
        eps = np.zeros((s.nk, s.nbands))
        rn_all = np.zeros((s.nk, s.nbands))
        for ik in range(s.nk):
            def apply(P, ik=ik):
                return _ks.apply_h(P, s.fft_index[ik], s.kin[ik], v_in, s.dims)
            HX = apply(X[ik])
            X[ik], _, eps[ik], rn_all[ik] = davidson(
                X[ik], HX, s.kin[ik], apply, n_david_first if iscf == 1 else n_david)
        occ, mu, mts = _ks.fermi_occupations(eps, s.weights, s.nelec, s.kT)
        rho_out = _ks.density(X, occ, s.weights, s.fft_index, s.dims, s.volume)
        _, eh = _ks.hartree(rho_out, s.gsq, s.dims, s.volume)
        _, exc, _ = _ks.xc_on_grid(rho_out, s.dims, s.volume)
        dv = s.volume / s.N
        e_band = float(np.sum(s.weights[:, None] * occ * eps))
        e_dc = dv * float(np.dot(vh + vxc, rho_out))
        etot = e_band - e_dc + eh + exc + mts
        drho = float(np.sqrt(dv * np.sum((rho_out - rho_in) ** 2)))
        de = None if e_hist is None else etot - e_hist
        hist.append(dict(iscf=iscf, etot=etot, de=de, drho=drho, e_band=e_band,
                         e_h=eh, e_xc=exc, mts=mts, mu=mu))
        if log:
            log(f"scf {iscf:3d}  E={etot:.12f}  dE={de if de is not None else 0.0:.3e}  drho={drho:.3e}")
        ok = de is not None and abs(de) < energy_tol
        if rho_tol is not None:
            ok = ok and drho < rho_tol
        e_hist = etot
        if ok:
            converged = True
            break
        rho_in = mixer(rho_in, rho_out)
    out = dict(etot=etot, eigs=eps, occ=occ, mu=mu, rho_out=rho_out, rho_in=rho_in,
               v_in=v_in, history=hist, converged=converged, iterations=iscf, X=X, res=rn_all)
    if final_res_tol is not None:
        for ik in range(s.nk):
            def apply(P, ik=ik):
                return _ks.apply_h(P, s.fft_index[ik], s.kin[ik], v_in, s.dims)
            HX = apply(X[ik])
            X[ik], _, eps[ik], rn_all[ik] = davidson(X[ik], HX, s.kin[ik], apply, 1000, res_tol=final_res_tol)
        out["eigs_final"] = eps.copy()
    return out
