"""The seeded generator behind the initial wavefunctions and the wells."""
import numpy as np

from derek_amd import config
from oracle import rng


def test_splitmix64_published_vectors():
    assert int(rng.splitmix64_sequence(0, 1)[0]) == 0xE220A8397B1DCDAF
    assert [int(x) for x in rng.splitmix64_sequence(1234567, 5)] == [
        6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
        16408922859458223821]


def test_product_and_oracle_streams_agree():
    # derek_amd.config (product, pure Python) and oracle.rng (numpy) restate
    # the same counter-based stream
    u = rng.uniforms(7, 2 << 48, 50)
    assert np.array_equal(u, np.array(config.seeded_uniforms(7, 2 << 48, 50)))


def test_normals_moments():
    z = rng.normal_pairs(3, rng.psi_stream(0, 1), 200000)
    for part in (z.real, z.imag):
        assert abs(part.mean()) < 0.01 and abs(part.std() - 1.0) < 0.01
    assert abs(np.mean(z.real * z.imag)) < 0.01
