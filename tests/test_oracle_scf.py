"""The oracle's SCF against exact solutions: closed-shell jellium, and
fixed-potential diagonalisation against a dense eigensolve."""
import numpy as np

from derek_amd import config
from oracle import ks, scf, xc
from tests.conftest import oracle_system


def test_davidson_matches_dense_eigh():
    s = oracle_system(config.small_system(seed=9, length=6.0, ecut=2.5, g=2.0, nwells=2, nbands=5))
    idx, kin, M = s.fft_index[0], s.kin[0], s.Mk[0]
    H = ks.apply_h_dense(M, kin, s.vext, s.dims)
    ref = np.linalg.eigvalsh(H)[:5]
    X = s.initial_psi(0)

    def apply(P):
        return ks.apply_h(P, idx, kin, s.vext, s.dims)

    X, HX, eps, rn = scf.davidson(X, apply(X), kin, apply, 300, res_tol=1e-10)
    assert rn.max() < 1e-9
    assert np.max(np.abs(eps - ref)) < 1e-12
    assert np.allclose(scf._ks.overlap(X, X), np.eye(5), atol=1e-12)


def test_closed_shell_jellium_is_exact():
    # V_ext = 0, 14 electrons fill G = 0 and the six |m| = 1 waves: rho is
    # uniform, V_H = 0, eps = 1/2|G|^2 + v_xc(n0), E = Omega n0 eps_xc(n0) + 2*6*(1/2)b^2.
    sp = config.small_system(seed=3, length=7.0, ecut=3.0, g=2.0, nwells=0, nelec=14, nbands=7)
    s = oracle_system(sp)
    r = scf.scf(s, max_scf=40, energy_tol=1e-12, rho_tol=1e-9)
    assert r["converged"]
    n0 = 14.0 / s.volume
    exc, vxc = xc.lda_xc(np.array([n0]))
    b2 = (2 * np.pi / 7.0) ** 2
    want = np.array([0.0] + [0.5 * b2] * 6) + vxc[0]
    assert np.max(np.abs(r["eigs"][0] - want)) < 1e-9
    assert abs(r["etot"] - (s.volume * n0 * exc[0] + 12 * 0.5 * b2)) < 1e-9
    assert np.max(np.abs(r["rho_out"] - n0)) < 1e-9


def test_scf_energy_equals_ks_functional():
    sp = config.small_system(seed=21)
    s = oracle_system(sp)
    r = scf.scf(s, max_scf=60, energy_tol=1e-12, rho_tol=1e-9)
    assert r["converged"]
    rho = r["rho_out"]
    dv = s.volume / s.N
    X = r["X"][0]
    ts = float(np.sum(r["occ"][0] * np.einsum("g,bg->b", s.kin[0], np.abs(X) ** 2)))
    _, eh = ks.hartree(rho, s.gsq, s.dims, s.volume)
    _, exc, _ = ks.xc_on_grid(rho, s.dims, s.volume)
    e_ks = ts + dv * float(np.dot(s.vext, rho)) + eh + exc
    assert abs(e_ks - r["etot"]) < 1e-8
    assert abs(rho.sum() * dv - sp.nelec) < 1e-10
