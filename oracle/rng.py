"""Counter-based generator shared by the library's initial wavefunctions.

The reference seeds gfortran's ``random_number`` (utils.f90:238-285,
wave.f90:859-894), which cannot be reproduced without gfortran; the build
defines its own: SplitMix64 finaliser as a hash of (seed, stream, index),
53-bit uniforms, Box-Muller normals.  Pinned by the published SplitMix64
test vector (seed 1234567 -> 6457827717110365317, ...) and by moment checks.
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
S30, S27, S31 = np.uint64(30), np.uint64(27), np.uint64(31)


def mix64(z):
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> S30)) * M1
        z = (z ^ (z >> S27)) * M2
    return z ^ (z >> S31)


def splitmix64_sequence(state, n):
    """Plain sequential SplitMix64 (state += golden; output = mix(state))."""
    with np.errstate(over="ignore"):
        st = np.uint64(state) + GOLDEN * np.arange(1, n + 1, dtype=np.uint64)
    return mix64(st)


def stream_key(seed, stream):
    with np.errstate(over="ignore"):
        return mix64(np.uint64(seed) * GOLDEN ^ mix64(np.uint64(stream) + GOLDEN))


def uniforms(seed, stream, n, start=0):
    """u_i = (mix(key + (i+1)*golden) >> 11) * 2^-53, i = start..start+n-1."""
    key = stream_key(seed, stream)
    idx = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = mix64(key + idx * GOLDEN)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def psi_stream(k, band):
    """Stream id of band ``band`` at k-point ``k`` (purpose tag 1)."""
    return (1 << 48) | (int(k) << 24) | int(band)


def normal_pairs(seed, stream, n):
    """n complex normals (re, im each N(0,1)) by Box-Muller on (u_2g, u_2g+1)."""
    u = uniforms(seed, stream, 2 * n)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(1.0 - u1))
    t = 2.0 * np.pi * u2
    return r * np.cos(t) + 1j * (r * np.sin(t))
