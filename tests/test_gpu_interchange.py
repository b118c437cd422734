"""Interchange files from and into the SCF engine's device state: the
density, total potential and wavefunctions written after an SCF read back
bit for bit, a .wvfn restarts a new context with the same wavefunctions, and
an external potential supplied as .potex gives the same SCF as the in-memory
one (external_pot: <file>.potex, io.f90:565-581, pot.f90:167-184)."""
import numpy as np
import pytest

from derek_amd import config, interchange, lib
from derek_amd import scf as gscf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    spec = config.small_system(seed=5, mp=(2, 1, 1), nelec=4, nbands=6, kT=0.01)
    return spec, gscf.Setup(spec)


def _scf(spec, su, ctx):
    return gscf.run(spec, max_scf=40, energy_tol=1e-9, setup=su, ctx=ctx, log=None)


def test_state_round_trip(tmp_path, system):
    spec, su = system
    print("interchange round trip", flush=True)
    ctx = su.context()
    try:
        res = _scf(spec, su, ctx)
        assert res["converged"]
        rho, v = ctx.get_density(), ctx.get_potential()
        interchange.write_density(str(tmp_path / "s.den"), su, rho)
        interchange.write_potential(str(tmp_path / "s.pot"), su, v)
        interchange.write_wvfn(str(tmp_path / "s.wvfn"), su, ctx)
        assert np.array_equal(interchange.read_density(str(tmp_path / "s.den"), su), rho)
        assert np.array_equal(interchange.read_potential(str(tmp_path / "s.pot"), su), v)
        psis = [ctx.get_psi(k) for k in range(su.nk)]
    finally:
        ctx.close()
    # a fresh context restarted from the .wvfn holds the same wavefunctions
    ctx2 = su.context()
    try:
        ctx2.scf_init(spec.nelec, kT=spec.kT, seed=spec.seed + 1)
        interchange.load_wvfn(str(tmp_path / "s.wvfn"), su, ctx2)
        for k in range(su.nk):
            assert np.array_equal(ctx2.get_psi(k), psis[k])
    finally:
        ctx2.close()


def test_external_potential_from_potex(tmp_path, system):
    spec, su = system
    print("interchange potex", flush=True)
    p = str(tmp_path / "wells.potex")
    interchange.write_potential(p, su, su.vext)
    v = interchange.read_potential(p, su)
    assert np.array_equal(v, su.vext)
    e = []
    for vext in (su.vext, v):
        ctx = su.context()
        try:
            ctx.set_vext(vext)
            e.append(_scf(spec, su, ctx)["etot"])
        finally:
            ctx.close()
    assert abs(e[0] - e[1]) < 1e-12
