"""Spectral tasks on the GPU library: ``task : dos`` and ``task : bandstructure``
(io.f90:522-527; the reference's ``spec_dos`` / ``spec_bands`` are empty,
spectral.f90:33-44).  After an SCF run, the converged total potential is
held fixed and the bands are found on another k-point set by the library's
non-self-consistent Davidson (``derek_hip_nscf``): the Monkhorst-Pack mesh
``dos_kpt_mp_grid`` (io.f90:111, default 11 11 11) for the density of
states, or a path through the Brillouin zone for the band structure.
"""
import numpy as np

from . import lib
from .scf import Setup

# simple-cubic path Gamma - X - M - Gamma - R (fractional coordinates)
CUBIC_PATH = (("G", (0.0, 0.0, 0.0)), ("X", (0.5, 0.0, 0.0)), ("M", (0.5, 0.5, 0.0)),
              ("G", (0.0, 0.0, 0.0)), ("R", (0.5, 0.5, 0.5)))


def kpath(vertices, points_per_segment=10):
    """Fractional k-points along straight segments between the vertices (the
    last vertex included) and the cumulative path length in fractional units."""
    v = np.asarray(vertices, dtype=np.float64)
    ks, dist = [v[0]], [0.0]
    for a, b in zip(v[:-1], v[1:]):
        for i in range(1, points_per_segment + 1):
            k = a + (b - a) * (i / points_per_segment)
            dist.append(dist[-1] + float(np.linalg.norm(k - ks[-1])))
            ks.append(k)
    return np.array(ks), np.array(dist)


def nscf_bands(spec, v_total, kfrac, nbands=None, precision=lib.DEREK_FP64, max_iter=300, res_tol=1e-9,
               seed=None):
    """Eigenvalues (nk x nbands, Ha) at the fixed total potential v_total on
    the fine grid of `spec`, for the fractional k-points kfrac; also returns
    the largest residual."""
    su = Setup(spec, kfrac=kfrac)
    ctx = su.context(precision, nbands=nbands)
    try:
        ctx.scf_init(spec.nelec, kT=spec.kT, seed=spec.seed if seed is None else seed)
        ctx.set_vloc(v_total)
        res = ctx.nscf(max_iter=max_iter, res_tol=res_tol)
        return ctx.eigenvalues(), res
    finally:
        ctx.close()


def dos(eigs, sigma, energies, weights=None):
    """Gaussian-broadened density of states per Ha and per cell, counting
    both spin channels (each band holds two electrons)."""
    e = np.asarray(eigs, dtype=np.float64)
    nk = e.shape[0]
    w = np.full(nk, 1.0 / nk) if weights is None else np.asarray(weights, dtype=np.float64)
    x = (np.asarray(energies)[:, None] - e.ravel()[None, :]) / sigma
    g = np.exp(-0.5 * x * x) / (sigma * np.sqrt(2.0 * np.pi))
    return 2.0 * g @ np.repeat(w, e.shape[1])
