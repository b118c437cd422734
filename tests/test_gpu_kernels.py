"""GPU kernels through the C-ABI against the oracle on the same seeded inputs.

FP64: relative error <= 1e-12 (FFT/H|psi>/density/Hartree/GEMM: sums of
O(N) terms in double).  Mixed mode (complex64 FFT and H|psi>): <= 5e-5.
"""
import json
import os

import numpy as np
import pytest

from derek_amd import config, lib
from derek_amd.scf import Setup
from oracle import fft as offt
from oracle import ks as oks
from oracle import xc as oxc
from tests.conftest import assert_close, rel_err

pytestmark = pytest.mark.gpu


def rnd_c(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


@pytest.mark.parametrize("dims", [(32, 32, 32), (8, 12, 10), (14, 22, 26), (15, 27, 20), (30, 16, 24), (72, 72, 72),
                                  (108, 108, 108), (144, 144, 144)])
@pytest.mark.parametrize("precision", [lib.DEREK_FP64, lib.DEREK_MIXED])
def test_fft3d(dims, precision):
    print("fft3d", dims, precision, flush=True)
    N = int(np.prod(dims))
    howmany = 2 if N < 200000 else 1
    ctx = lib.Context(dims, 1, 1, 1.0, 1, precision)
    rng = np.random.default_rng(N)
    a = rnd_c(rng, howmany, N)
    tol = 1e-12 if precision == lib.DEREK_FP64 else 5e-5
    for d in (1, -1):
        got = ctx.fft3d(a, d)
        ref = offt.fft3d(a, dims, d)
        assert_close(got, ref, tol)
    ctx.close()


def test_fft3d_roundtrip_and_bad_direction():
    dims = (108, 108, 108)
    N = int(np.prod(dims))
    ctx = lib.Context(dims, 1, 1, 1.0, 1)
    a = rnd_c(np.random.default_rng(0), N)
    back = ctx.fft3d(ctx.fft3d(a, 1), -1)
    assert_close(back, N * a, 1e-12)
    with pytest.raises(lib.DerekError) as e:
        ctx.fft3d(a, 0)
    assert e.value.code == lib.DEREK_ERR_ARG
    ctx.close()


@pytest.mark.parametrize("dims,n", [((8, 12, 10), 8), ((8, 12, 10), 12), ((8, 12, 10), 10), ((14, 22, 26), 26),
                                    ((14, 22, 26), 22), ((14, 22, 26), 14), ((32, 32, 32), 32), ((72, 72, 72), 72),
                                    ((108, 108, 108), 108), ((144, 144, 144), 144)])
@pytest.mark.parametrize("precision", [lib.DEREK_FP64, lib.DEREK_MIXED])
def test_fft1d(dims, n, precision):
    """fft_1d (fft.f90:305-376) against the oracle, which the brute-force DFT
    pins: 1D lines of one of the grid's lengths, both directions."""
    print("fft1d", dims, n, precision, flush=True)
    ctx = lib.Context(dims, 1, 1, 1.0, 1, precision)
    rng = np.random.default_rng(n)
    for howmany in (1, 7, 1000):
        a = rnd_c(rng, howmany, n)
        tol = 1e-12 if precision == lib.DEREK_FP64 else 5e-5
        for d in (1, -1):
            assert_close(ctx.fft1d(a, d), offt.fft1d(a, d), tol, f"fft1d n={n} howmany={howmany} dir={d}")
    with pytest.raises(lib.DerekError) as e:
        ctx.fft1d(rnd_c(rng, 2, 9), 1)
    assert e.value.code == lib.DEREK_ERR_ARG
    with pytest.raises(lib.DerekError) as e:
        ctx.fft1d(rnd_c(rng, 2, n), 0)
    assert e.value.code == lib.DEREK_ERR_ARG
    ctx.close()


def test_wave_axpby():
    """The wavefunction operators of wave.f90 (+, -, real and complex *,
    copies) against the oracle's restatement, on a wavefunction-shaped
    array (k-points x bands x plane waves) and a slice."""
    from oracle import wave as ow
    print("wave algebra", flush=True)
    ctx = lib.Context((8, 8, 8), 1, 1, 1.0, 1)
    rng = np.random.default_rng(9)
    for shape in [(3, 5, 1001), (7, 333)]:
        x, y = rnd_c(rng, *shape), rnd_c(rng, *shape)
        cases = [
            (ctx.wave_axpby(1.0, x, 1.0, y), ow.add(x, y)),
            (ctx.wave_axpby(1.0, x, -1.0, y), ow.sub(x, y)),
            (ctx.wave_axpby(-2.5, x), ow.scale(-2.5, x)),
            (ctx.wave_axpby(0.3 - 1.7j, x), ow.scale(0.3 - 1.7j, x)),
            (ctx.wave_axpby(1.0, x), ow.copy(x)),
        ]
        for got, ref in cases:
            assert got.shape == ref.shape
            assert_close(got, ref, 1e-15)
        # exact cases: a copy and a sum of dyadic values are bit-exact
        assert np.array_equal(ctx.wave_axpby(1.0, x), x)
        xd = np.round(x * 64) / 64
        yd = np.round(y * 64) / 64
        assert np.array_equal(ctx.wave_axpby(1.0, xd, 1.0, yd), ow.add(xd, yd))
    with pytest.raises(lib.DerekError) as e:
        lib.check(ctx.lib.derek_hip_wave_axpby(ctx.h, 0, (lib.C.c_double * 2)(1, 0), None,
                                               (lib.C.c_double * 2)(0, 0), None, None))
    assert e.value.code == lib.DEREK_ERR_ARG
    ctx.close()


def make_setup(spec):
    su = Setup(spec)
    return su


@pytest.mark.parametrize("name", ["small", "small_k", "gamma32"])
@pytest.mark.parametrize("precision", [lib.DEREK_FP64, lib.DEREK_MIXED])
def test_apply_h(name, precision):
    print("apply_h", name, precision, flush=True)
    if name == "small":
        spec = config.small_system(seed=3)
    elif name == "small_k":
        spec = config.small_system(seed=4, mp=(2, 1, 2), ecut=5.0)
    else:
        spec = config.workload("gamma32")
    su = make_setup(spec)
    ctx = su.context(precision, nbands=11)
    rng = np.random.default_rng(7)
    v = su.vext + 0.3 * rng.normal(size=su.N)
    ctx.set_vloc(v)
    tol = 1e-12 if precision == lib.DEREK_FP64 else 5e-5
    for k in range(su.nk):
        psi = rnd_c(rng, 11, su.npw[k])
        got = ctx.apply_h(k, psi)
        ref = oks.apply_h(psi, su.fft_index[k], su.kin[k], v, su.dims)
        assert_close(got, ref, tol)
    ctx.close()


def test_apply_h_state_errors():
    su = make_setup(config.small_system(seed=3))
    ctx = su.context()
    psi = np.ones((1, su.npw[0]), dtype=complex)
    with pytest.raises(lib.DerekError) as e:
        ctx.apply_h(0, psi)
    assert e.value.code == lib.DEREK_ERR_STATE
    with pytest.raises(lib.DerekError) as e:
        ctx.set_basis(0, np.array([0, su.N], dtype=np.int32), np.zeros(2), 1.0)
    assert e.value.code == lib.DEREK_ERR_ARG
    ctx.close()


def test_apply_h_hermitian_full_size():
    """108^3 grid, 62669 plane waves: <x|Hy> = conj(<y|Hx>)."""
    su = make_setup(config.workload("gamma108"))
    ctx = su.context(nbands=2)
    ctx.set_vloc(su.vext)
    rng = np.random.default_rng(11)
    xy = rnd_c(rng, 2, su.npw[0])
    h = ctx.apply_h(0, xy)
    a = np.vdot(xy[0], h[1])
    b = np.conj(np.vdot(xy[1], h[0]))
    assert abs(a - b) < 1e-11 * abs(a)
    # kinetic-only check: V = 0 gives H = diag(1/2|G|^2) exactly
    ctx.set_vloc(np.zeros(su.N))
    h0 = ctx.apply_h(0, xy)
    assert_close(h0, su.kin[0][None, :] * xy, 1e-13)
    ctx.close()


@pytest.mark.parametrize("m,n,npw", [(10, 10, 1503), (37, 53, 777), (160, 160, 5003), (80, 80, 62669)])
@pytest.mark.parametrize("precision", [lib.DEREK_FP64, lib.DEREK_MIXED])
def test_overlap(m, n, npw, precision):
    """FP64: f64 MFMA against numpy to 1e-13.  MIXED with FP32 matrix cores:
    operands rounded to complex64, f32 MFMA accumulation; checked against the
    FP64 product of the rounded operands (tolerance 1e-5 of the largest
    entry)."""
    print("overlap", m, n, npw, precision, flush=True)
    rng = np.random.default_rng(m * 1000 + n)
    ctx = lib.Context((8, 8, 8), 1, 1, 1.0, 1, precision)
    ctx.set_gemm(0)
    X = rnd_c(rng, m, npw)
    Y = rnd_c(rng, n, npw)
    tol = 1e-13 if precision == lib.DEREK_FP64 else 1e-5
    Xr = X if precision == lib.DEREK_FP64 else X.astype(np.complex64).astype(np.complex128)
    Yr = Y if precision == lib.DEREK_FP64 else Y.astype(np.complex64).astype(np.complex128)
    assert_close(ctx.overlap(X, Y), np.conj(Xr) @ Yr.T, tol)
    if m == n:
        S = ctx.overlap(X, X)
        assert_close(S, np.conj(Xr) @ Xr.T, tol)
        assert np.array_equal(S, S.conj().T)
    ctx.close()


@pytest.mark.parametrize("m,n,npw", [(10, 10, 1503), (37, 53, 777), (160, 80, 4003)])
@pytest.mark.parametrize("precision", [lib.DEREK_FP64, lib.DEREK_MIXED])
def test_rotate(m, n, npw, precision):
    """MIXED with FP32 matrix cores: X and U rounded to complex64, f32 MFMA
    accumulation."""
    print("rotate", m, n, npw, precision, flush=True)
    rng = np.random.default_rng(m + n)
    ctx = lib.Context((8, 8, 8), 1, 1, 1.0, 1, precision)
    ctx.set_gemm(0)
    X = rnd_c(rng, m, npw)
    U = rnd_c(rng, m, n)
    if precision == lib.DEREK_FP64:
        assert_close(ctx.rotate(X, U), U.T @ X, 1e-13)
    else:
        Xr = X.astype(np.complex64).astype(np.complex128)
        Ur = U.astype(np.complex64).astype(np.complex128)
        assert_close(ctx.rotate(X, U), Ur.T @ Xr, 1e-5)
    ctx.close()


BF16_ACC = {1: 2e-2, 2: 1e-4, 3: 2e-6}  # against the exact product of the complex64 operands


@pytest.mark.parametrize("m,n,npw", [(10, 10, 1503), (37, 53, 777), (160, 160, 5003), (80, 80, 62669),
                                     (300, 260, 3001)])
@pytest.mark.parametrize("nt", [1, 2, 3])
@pytest.mark.parametrize("tile", [64, 128])
def test_overlap_bf16(m, n, npw, nt, tile):
    """BF16 MFMA overlaps of complex64 operands split into nt bf16 terms, on
    64- or 128-wide C tiles (three terms always take 64): against the
    oracle's split product (FP32 accumulation: 1e-5 of the largest entry),
    against the exact product to the split's accuracy, and exactly Hermitian
    for X^H X."""
    print("overlap_bf16", m, n, npw, nt, tile, flush=True)
    rng = np.random.default_rng(m * 1000 + n + nt)
    ctx = lib.Context((8, 8, 8), 1, 1, 1.0, 1, lib.DEREK_MIXED)
    ctx.set_gemm(nt)
    ctx.set_tiles(0, 2, tile)
    X = rnd_c(rng, m, npw).astype(np.complex64)
    Y = rnd_c(rng, n, npw).astype(np.complex64)
    got = ctx.overlap(X, Y)
    assert_close(got, oks.overlap_bf16(X, Y, nt), 1e-5)
    assert_close(got, np.conj(X.astype(np.complex128)) @ Y.astype(np.complex128).T, BF16_ACC[nt])
    if m == n:
        # the Hermitian kernel multiplies the upper triangle and mirrors it
        # (real diagonal); the split product of conj(x_j) x_i is not the
        # conjugate of that of conj(x_i) x_j to the last bf16 bit
        S = ctx.overlap(X, X)
        R = oks.overlap_bf16(X, X, nt)
        R = np.triu(R, 1) + np.triu(R, 1).conj().T + np.diag(np.diag(R).real)
        assert_close(S, R, 1e-5)
        assert np.array_equal(S, S.conj().T)
    ctx.set_gemm(0)
    Xr = X.astype(np.complex128)
    assert_close(ctx.overlap(X, Y), np.conj(Xr) @ Y.astype(np.complex128).T, 1e-5)
    ctx.close()


@pytest.mark.parametrize("m,n,npw", [(10, 10, 1503), (37, 53, 777), (160, 80, 4003), (600, 300, 2048)])
@pytest.mark.parametrize("nt", [1, 2, 3])
def test_rotate_bf16(m, n, npw, nt):
    """BF16 MFMA rotation Z = X U (U rounded to complex64, both split into nt
    bf16 terms) against the oracle's split product and the exact product."""
    print("rotate_bf16", m, n, npw, nt, flush=True)
    rng = np.random.default_rng(m + n + nt)
    ctx = lib.Context((8, 8, 8), 1, 1, 1.0, 1, lib.DEREK_MIXED)
    ctx.set_gemm(nt)
    X = rnd_c(rng, m, npw).astype(np.complex64)
    U = rnd_c(rng, m, n).astype(np.complex64)
    got = ctx.rotate(X, U)
    assert_close(got, oks.rotate_bf16(X, U, nt), 1e-5)
    assert_close(got, U.astype(np.complex128).T @ X.astype(np.complex128), BF16_ACC[nt])
    ctx.close()


def test_set_gemm_errors():
    """Out-of-range term counts are refused; a DEREK_MIXED context starts on
    the two-term BF16 GEMMs (its overlap equals the oracle's nt = 2 split)."""
    ctx = lib.Context((8, 8, 8), 1, 1, 1.0, 1, lib.DEREK_MIXED)
    rng = np.random.default_rng(2)
    X = rnd_c(rng, 20, 900).astype(np.complex64)
    Y = rnd_c(rng, 12, 900).astype(np.complex64)
    assert_close(ctx.overlap(X, Y), oks.overlap_bf16(X, Y, 2), 1e-5)
    for bad in (-1, 4):
        with pytest.raises(lib.DerekError):
            ctx.set_gemm(bad)
    ctx.close()


@pytest.mark.parametrize("precision", [lib.DEREK_FP64, lib.DEREK_MIXED])
def test_density(precision):
    spec = config.workload("gamma32")
    su = make_setup(spec)
    ctx = su.context(precision)
    rng = np.random.default_rng(5)
    psi = rnd_c(rng, 10, su.npw[0]) / np.sqrt(su.npw[0])
    occ = np.array([2.0, 2.0, 2.0, 2.0, 1.0, 0.0, 0.5, 0.0, 0.0, 0.25])
    got = ctx.density(0, psi, occ)
    ref = oks.density([psi], [occ], [1.0], [su.fft_index[0]], su.dims, su.volume)
    tol = 1e-12 if precision == lib.DEREK_FP64 else 5e-5
    assert_close(got, ref, tol)
    got2 = ctx.density(0, psi, occ, rho=got)
    assert_close(got2, 2 * ref, tol)
    ctx.close()


def test_lda_matches_libxc_fixture_and_oracle():
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "libxc_lda.json")))
    ctx = lib.Context((8, 8, 8), 1, 1, 1.0, 1)
    rho = np.array(fx["regression"]["lda_x.BrOH"]["rho"])
    e, v = ctx.lda(rho)
    ex = np.array(fx["regression"]["lda_x.BrOH"]["exc"]) + np.array(fx["regression"]["lda_c_pz.BrOH"]["exc"])
    vx = np.array(fx["regression"]["lda_x.BrOH"]["vxc"]) + np.array(fx["regression"]["lda_c_pz.BrOH"]["vxc"])
    assert np.max(np.abs(e - ex) / (1 + np.abs(ex))) < 1e-13
    assert np.max(np.abs(v - vx) / (1 + np.abs(vx))) < 1e-13
    r = np.concatenate([np.geomspace(1e-16, 1e3, 5000), [0.0, 1e-15, 1.5e-15, 3 / (4 * np.pi)]])
    e, v = ctx.lda(r)
    eo, vo = oxc.lda_xc(r)
    assert np.max(np.abs(e - eo) / (1e-300 + np.abs(eo) + (eo == 0))) < 1e-13
    assert np.max(np.abs(v - vo) / (1e-300 + np.abs(vo) + (vo == 0))) < 1e-13
    ctx.close()


def test_hartree():
    su = make_setup(config.workload("gamma32"))
    ctx = su.context()
    rng = np.random.default_rng(9)
    rho = 0.01 + 0.001 * rng.random(su.N)
    vh, eh = ctx.hartree(rho)
    vr, er = oks.hartree(rho, su.gsq, su.dims, su.volume)
    assert_close(vh, vr, 1e-12)
    assert abs(eh - er) < 1e-12 * abs(er)
    ctx.close()
