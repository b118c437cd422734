"""Spectral-task helpers on the host: k-point paths and the broadened DOS."""
import numpy as np

from derek_amd import spectral


def test_kpath_vertices_and_length():
    verts = [v for _, v in spectral.CUBIC_PATH]
    ks, d = spectral.kpath(verts, points_per_segment=5)
    assert ks.shape == (1 + 4 * 5, 3)
    for i, v in enumerate(verts):
        assert np.allclose(ks[5 * i], v)
    assert np.all(np.diff(d) > 0)
    # G-X 0.5, X-M 0.5, M-G sqrt(0.5), G-R sqrt(0.75)
    assert np.isclose(d[-1], 0.5 + 0.5 + np.sqrt(0.5) + np.sqrt(0.75))


def test_dos_counts_two_electrons_per_band():
    rng = np.random.default_rng(0)
    eigs = np.sort(rng.uniform(-1.0, 1.0, size=(7, 5)), axis=1)
    e = np.linspace(-3.0, 3.0, 6001)
    g = spectral.dos(eigs, 0.02, e)
    assert abs(np.trapezoid(g, e) - 2 * 5) < 1e-9
    # a single level: a Gaussian of height 2 / (sigma sqrt(2 pi)) at its energy
    g1 = spectral.dos(np.array([[0.3]]), 0.05, np.array([0.3]))
    assert np.isclose(g1[0], 2.0 / (0.05 * np.sqrt(2 * np.pi)))
