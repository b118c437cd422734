"""Reference-convention 3D FFT (fft.f90:222-303, grids.f90:54-306).

Layout: a flat array with z fastest and x slowest (the reshape with
order=(3,2,1) around ``fftw_execute_dft``).  ``direction=+1`` is
FFTW_FORWARD (exp(-iGr), grids ``real2recip``), ``direction=-1`` is
FFTW_BACKWARD (exp(+iGr), ``recip2real``); both unnormalised, so a round trip
multiplies by N.  Any other direction is an error (fft.f90:268).

Pinned by an O(N^2) brute-force DFT on small grids and by closed forms
(delta -> constant, single plane wave -> N*delta at bin m mod n).
"""
import numpy as np


def fft3d(grid, dims, direction):
    nx, ny, nz = dims
    a = np.asarray(grid, dtype=np.complex128)
    lead = a.shape[:-1]
    a = a.reshape(lead + (nx, ny, nz))
    if direction == 1:
        out = np.fft.fftn(a, axes=(-3, -2, -1))
    elif direction == -1:
        out = np.fft.ifftn(a, axes=(-3, -2, -1)) * float(nx * ny * nz)
    else:
        raise ValueError("direction must be +1 or -1")
    return out.reshape(lead + (nx * ny * nz,))


def dft_bruteforce(grid, dims, direction):
    """Explicit triple-sum DFT, sign = -direction (tiny grids only)."""
    if direction not in (1, -1):
        raise ValueError("direction must be +1 or -1")
    nx, ny, nz = dims
    a = np.asarray(grid, dtype=np.complex128).reshape(nx, ny, nz)
    sign = -1.0 if direction == 1 else 1.0
    out = np.zeros_like(a)
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    for kx in range(nx):
        for ky in range(ny):
            for kz in range(nz):
                ph = np.exp(sign * 2j * np.pi * (kx * ix / nx + ky * iy / ny + kz * iz / nz))
                out[kx, ky, kz] = np.sum(a * ph)
    return out.reshape(-1)
