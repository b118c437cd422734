"""The fused y/x pass of H|psi> (one workgroup per z plane on z-major
sticks) against the oracle, against the separate y and x passes, and inside
the SCF.

It runs at the grid lengths with compiled kernels (32, 72, 108, 144), FP64
to a relative 1e-12 and complex64 to 5e-5 of the oracle, as the separate
passes do.  The kernel profile shows which path ran: the FP64 working set of
a 144^3 plane does not fit the 160 KiB of LDS, so that case must fall back
to the separate passes.
"""
import numpy as np
import pytest

from derek_amd import config, lib
from derek_amd import scf as gscf
from oracle import ks as oks
from tests.conftest import assert_close

pytestmark = pytest.mark.gpu


def rnd_c(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def grid_spec(n):
    """Systems on an n^3 grid: the benchmark cells, with one or two k-points."""
    if n == 32:
        return config.workload("gamma32")
    if n == 72:
        return config.SystemSpec("g72", 16.0, 20.0, 2.2, (2, 1, 1), 32, 24, kT=0.01, nwells=8, seed=4)
    if n == 108:
        return config.workload("gamma108")
    return config.SystemSpec("g144", 30.0, 25.0, 2.125, (1, 1, 1), 512, 300, kT=0.002, nwells=64, seed=5)


def fused_expected(n, precision):
    return not (n == 144 and precision == lib.DEREK_FP64)


@pytest.mark.parametrize("n", [32, 72, 108, 144])
@pytest.mark.parametrize("precision", [lib.DEREK_FP64, lib.DEREK_MIXED])
@pytest.mark.parametrize("z_sticks", [16, 32])
def test_fused_apply_h_matches_oracle(n, precision, z_sticks):
    """z_sticks: sticks per workgroup of the z passes."""
    print("fused apply_h", n, precision, "z sticks", z_sticks, flush=True)
    su = gscf.Setup(grid_spec(n))
    assert tuple(su.dims) == (n, n, n)
    nb = 5
    ctx = su.context(precision, nbands=nb)
    ctx.set_tiles(z_sticks, 2, 64)
    rng = np.random.default_rng(n)
    v = su.vext + 0.3 * rng.normal(size=su.N)
    ctx.set_vloc(v)
    # 5 bands in launches of 3 and 2 items, workgroups of 2 bands: every tail
    ctx.set_launch(bands_per_wg=2, items_per_launch=3, fused=True)
    ctx.profile(True)
    tol = 1e-12 if precision == lib.DEREK_FP64 else 5e-5
    for k in range(su.nk):
        psi = rnd_c(rng, nb, su.npw[k])
        got = ctx.apply_h(k, psi)
        ref = oks.apply_h(psi, su.fft_index[k], su.kin[k], v, su.dims)
        assert_close(got, ref, tol)
    prof = ctx.profile_read()
    fz = fused_expected(n, precision)
    assert (prof["slab_yx_vmul"][0] > 0) == fz
    assert (prof["planes_x_vmul"][0] > 0) == (not fz)
    assert ctx.tuning()["fused"]
    ctx.close()


@pytest.mark.parametrize("precision", [lib.DEREK_FP64, lib.DEREK_MIXED])
def test_fused_matches_separate_passes(precision):
    """108^3, 62669 plane waves, 19 bands: the same H|psi> for every launch shape."""
    print("fused vs separate", precision, flush=True)
    su = gscf.Setup(config.workload("gamma108"))
    nb = 19
    ctx = su.context(precision, nbands=nb)
    rng = np.random.default_rng(5)
    ctx.set_vloc(su.vext + 0.2 * rng.normal(size=su.N))
    psi = rnd_c(rng, nb, su.npw[0])
    ctx.set_launch(4, 16, fused=False)
    ref = ctx.apply_h(0, psi)
    tol = 1e-13 if precision == lib.DEREK_FP64 else 1e-5
    for bpw, items in [(1, 19), (2, 7), (4, 16), (8, 19)]:
        ctx.set_launch(bpw, items, fused=True)
        assert_close(ctx.apply_h(0, psi), ref, tol)
    ctx.close()


@pytest.mark.parametrize("n", [32, 72, 108, 144])
@pytest.mark.parametrize("precision", [lib.DEREK_FP64, lib.DEREK_MIXED])
def test_fused_density_matches_oracle(n, precision):
    """rho = sum occ |psi(r)|^2 / volume through the fused pass, with zero
    occupations skipped and the accumulate (keep) form."""
    print("fused density", n, precision, flush=True)
    su = gscf.Setup(grid_spec(n))
    nb = 7
    ctx = su.context(precision, nbands=nb)
    ctx.set_launch(2, 5, fused=True)  # 7 bands in launches of 5 and 2
    ctx.profile(True)
    rng = np.random.default_rng(n + 1)
    occ = np.array([2.0, 0.0, 1.5, 2.0, 0.25, 0.0, 1.0])
    tol = 1e-12 if precision == lib.DEREK_FP64 else 5e-5
    for k in range(su.nk):
        psi = rnd_c(rng, nb, su.npw[k]) / np.sqrt(su.npw[k])
        got = ctx.density(k, psi, occ)
        ref = oks.density([psi], [occ], [su.weights[k]], [su.fft_index[k]], su.dims, su.volume)
        assert_close(got, ref, tol)
        assert_close(ctx.density(k, psi, occ, rho=got), 2 * ref, tol)
    prof = ctx.profile_read()
    fz = fused_expected(n, precision)
    assert (prof["slab_yx_density"][0] > 0) == fz
    assert (prof["planes_x_density"][0] > 0) == (not fz)
    ctx.close()


def test_fused_hermitian_and_kinetic_only():
    """<x|Hy> = conj(<y|Hx>) through the fused pass; V = 0 gives the kinetic term exactly."""
    su = gscf.Setup(config.workload("gamma108"))
    ctx = su.context(nbands=2)
    ctx.set_launch(4, 16, fused=True)
    ctx.set_vloc(su.vext)
    rng = np.random.default_rng(13)
    xy = rnd_c(rng, 2, su.npw[0])
    h = ctx.apply_h(0, xy)
    a = np.vdot(xy[0], h[1])
    b = np.conj(np.vdot(xy[1], h[0]))
    assert abs(a - b) < 1e-11 * abs(a)
    ctx.set_vloc(np.zeros(su.N))
    assert_close(ctx.apply_h(0, xy), su.kin[0][None, :] * xy, 1e-13)
    ctx.close()


def test_set_launch_rejects_bad_shapes():
    su = gscf.Setup(config.workload("gamma32"))
    ctx = su.context(nbands=2)
    for bpw, items, fused in [(3, 16, 1), (4, 0, 1), (4, 16, 4), (4, 16, -1)]:
        with pytest.raises(lib.DerekError) as e:
            lib.check(ctx.lib.derek_hip_set_launch(ctx.h, bpw, items, fused))
        assert e.value.code == lib.DEREK_ERR_ARG
    ctx.close()


@pytest.mark.parametrize("n,precision", [(32, lib.DEREK_FP64), (72, lib.DEREK_FP64), (72, lib.DEREK_MIXED)])
def test_fused_scf_matches_separate(n, precision):
    """Per-step energies of the SCF with the fused pass equal those with the
    separate passes (FP64: to rounding; mixed: to 5e-6 Ha, about four float32
    ulps of the ~9 Ha total energy, since the two paths round V / N and their
    sums to complex64 differently and the SCF carries that rounding from step
    to step; observed 4e-7 to 1.7e-6 Ha).  The mixed runs use the FP32
    matrix cores: the BF16 split GEMMs carry ~1e-5 relative overlap error,
    which moves FP32-phase energies at first order and would hide the
    comparison of the two FFT paths."""
    print("fused scf", n, precision, flush=True)
    spec = grid_spec(n)
    out = []
    for fused in (False, True):
        su = gscf.Setup(spec)
        ctx = su.context(precision)
        ctx.set_gemm(0)
        ctx.set_launch(4, 64, fused=fused)
        res = gscf.run(spec, max_scf=6, energy_tol=0.0, setup=su, ctx=ctx, log=None)
        out.append(np.array([h["etot"] for h in res["history"]]))
        ctx.close()
    tol = 1e-10 if precision == lib.DEREK_FP64 else 5e-6
    assert np.max(np.abs(out[1] - out[0])) < tol
