"""Reference-layout FFT (fft.f90 conventions) against an explicit DFT and
closed forms."""
import numpy as np
import pytest

from oracle import fft


@pytest.mark.parametrize("dims", [(4, 6, 5), (3, 5, 4), (2, 2, 7)])
@pytest.mark.parametrize("direction", [1, -1])
def test_bruteforce(dims, direction):
    rng = np.random.default_rng(3)
    N = int(np.prod(dims))
    a = rng.normal(size=N) + 1j * rng.normal(size=N)
    assert np.allclose(fft.fft3d(a, dims, direction), fft.dft_bruteforce(a, dims, direction), atol=1e-11)


def test_sign_and_layout():
    dims = (4, 6, 5)
    nx, ny, nz = dims
    # e^{+i G r} with m = (1, 2, -1) -> forward (e^{-iGr}) gives N at bin (1, 2, 4)
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    pw = np.exp(2j * np.pi * (1 * ix / nx + 2 * iy / ny - 1 * iz / nz)).reshape(-1)
    out = fft.fft3d(pw, dims, 1)
    k = (1 * ny + 2) * nz + 4
    assert abs(out[k] - 120) < 1e-10
    out[k] = 0
    assert np.max(np.abs(out)) < 1e-10
    # round trip multiplies by N
    assert np.allclose(fft.fft3d(fft.fft3d(pw, dims, 1), dims, -1), 120 * pw)


def test_bad_direction():
    with pytest.raises(ValueError):
        fft.fft3d(np.zeros(8), (2, 2, 2), 0)
