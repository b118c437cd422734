"""The oracle's wavefunction algebra and 1D FFT against closed forms."""
import numpy as np
import pytest

from oracle import fft
from oracle import wave


def test_wave_algebra_hand_case():
    x = np.array([1.0 + 2.0j, -3.0 + 0.5j])
    y = np.array([0.25 - 1.0j, 2.0 + 2.0j])
    assert np.array_equal(wave.add(x, y), [1.25 + 1.0j, -1.0 + 2.5j])
    assert np.array_equal(wave.sub(x, y), [0.75 + 3.0j, -5.0 - 1.5j])
    assert np.array_equal(wave.scale(2.0, x), [2.0 + 4.0j, -6.0 + 1.0j])
    # (0 + 1i)(a + bi) = -b + ai
    assert np.array_equal(wave.scale(1j, x), [-2.0 + 1.0j, -0.5 - 3.0j])
    c = wave.copy(x)
    assert np.array_equal(c, x) and c is not x


def test_wave_algebra_identities():
    rng = np.random.default_rng(3)
    # dyadic values: sums and differences are exact
    x = (rng.integers(-512, 512, size=(7, 33)) + 1j * rng.integers(-512, 512, size=(7, 33))) / 64.0
    y = (rng.integers(-512, 512, size=(7, 33)) + 1j * rng.integers(-512, 512, size=(7, 33))) / 64.0
    assert np.array_equal(wave.sub(wave.add(x, y), y), x)
    assert np.array_equal(wave.sub(x, x), np.zeros_like(x))
    assert np.array_equal(wave.scale(1j, wave.scale(1j, x)), -x)
    assert np.array_equal(wave.scale(-1.0, x), wave.sub(np.zeros_like(x), x))


@pytest.mark.parametrize("n", [1, 2, 5, 8, 12, 13, 20])
@pytest.mark.parametrize("direction", [1, -1])
def test_fft1d_bruteforce(n, direction):
    rng = np.random.default_rng(n)
    a = rng.normal(size=(3, n)) + 1j * rng.normal(size=(3, n))
    assert np.allclose(fft.fft1d(a, direction), fft.dft1d_bruteforce(a, direction), atol=1e-12 * n)


def test_fft1d_closed_forms():
    n = 12
    # delta at 0 -> constant 1; exp(+2 pi i m j/n) -> n delta at m under the forward sign
    d = np.zeros(n, dtype=complex)
    d[0] = 1.0
    assert np.allclose(fft.fft1d(d, 1), np.ones(n))
    m = 5
    e = np.exp(2j * np.pi * m * np.arange(n) / n)
    out = fft.fft1d(e, 1)
    ref = np.zeros(n, dtype=complex)
    ref[m] = n
    assert np.allclose(out, ref, atol=1e-12)
    # round trip multiplies by n (unnormalised both ways)
    rng = np.random.default_rng(1)
    a = rng.normal(size=n) + 1j * rng.normal(size=n)
    assert np.allclose(fft.fft1d(fft.fft1d(a, 1), -1), n * a)
    with pytest.raises(ValueError):
        fft.fft1d(a, 0)
