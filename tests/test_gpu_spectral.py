"""Non-self-consistent eigenvalues (task : dos / bandstructure) on the GPU:
at the SCF's own k-points they reproduce the converged SCF bands, and at
new k-points they match the oracle's dense H_GG' diagonalised with numpy."""
import numpy as np
import pytest

from derek_amd import config, lib, spectral
from derek_amd import scf as gscf
from oracle import basis as ob
from oracle import ks as oks

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def converged():
    spec = config.small_system(seed=7, mp=(2, 1, 1), nelec=4, nbands=6, kT=0.01)
    su = gscf.Setup(spec)
    ctx = su.context()
    try:
        res = gscf.run(spec, max_scf=60, energy_tol=1e-10, setup=su, ctx=ctx, log=None)
        assert res["converged"]
        ctx.scf_polish(1000, 1e-10)
        return spec, su, ctx.get_potential(), ctx.eigenvalues()
    finally:
        ctx.close()


def test_nscf_reproduces_scf_bands(converged):
    spec, su, v, eig = converged
    print("nscf on the SCF mesh", flush=True)
    e, res = spectral.nscf_bands(spec, v, su.kfrac, res_tol=1e-10, max_iter=1000, seed=spec.seed + 3)
    assert res < 1e-9
    assert np.max(np.abs(e - eig)) < 1e-8


def test_nscf_matches_dense_oracle(converged):
    spec, su, v, _ = converged
    print("nscf vs dense oracle", flush=True)
    kf = np.array([[0.3, 0.1, -0.2], [0.5, 0.5, 0.5]])
    e, res = spectral.nscf_bands(spec, v, kf, res_tol=1e-10, max_iter=1000)
    assert res < 1e-9
    G, M = ob.g_map(su.inv, su.dims)
    for i, k in enumerate(kf):
        kc = ob.kcart(su.inv, k)
        idx, kin, _ = ob.sphere(G, M, su.dims, kc, spec.ecut)
        m = M[np.isin(ob.fft_bin(M, su.dims), idx)]
        # same order as idx (map order) for the dense matrix
        order = {int(b): j for j, b in enumerate(ob.fft_bin(m, su.dims))}
        m = m[[order[int(b)] for b in idx]]
        H = oks.apply_h_dense(m, kin, v, su.dims)
        ref = np.linalg.eigvalsh(H)[:spec.nbands]
        assert np.max(np.abs(e[i] - ref)) < 1e-8, (i, e[i] - ref)
