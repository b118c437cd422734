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
import scipy.fft as _sfft

# Threads for the transforms (scipy's pocketfft, the same algorithm as
# numpy.fft); 1 by default, raised by the fixture scripts for large grids.
WORKERS = 1


def set_workers(n):
    global WORKERS
    WORKERS = max(1, int(n))


def fft3d(grid, dims, direction):
    nx, ny, nz = dims
    a = np.asarray(grid, dtype=np.complex128)
    lead = a.shape[:-1]
    a = a.reshape(lead + (nx, ny, nz))
    if direction == 1:
        out = _sfft.fftn(a, axes=(-3, -2, -1), workers=WORKERS)
    elif direction == -1:
        out = _sfft.ifftn(a, axes=(-3, -2, -1), workers=WORKERS) * float(nx * ny * nz)
    else:
        raise ValueError("direction must be +1 or -1")
    return out.reshape(lead + (nx * ny * nz,))


def fft1d(lines, direction):
    """``fft_1d`` (fft.f90:305-376): each row of ``lines`` transformed with
    the plan of fft.f90:209-212 (length nz), direction +1 FFTW_FORWARD
    (exp(-2 pi i jk/n)), -1 FFTW_BACKWARD (exp(+2 pi i jk/n)), unnormalised."""
    a = np.asarray(lines, dtype=np.complex128)
    n = a.shape[-1]
    if direction == 1:
        return _sfft.fft(a, axis=-1, workers=WORKERS)
    if direction == -1:
        return _sfft.ifft(a, axis=-1, workers=WORKERS) * float(n)
    raise ValueError("direction must be +1 or -1")


def dft1d_bruteforce(lines, direction):
    """Explicit sum over j of x_j exp(-direction 2 pi i jk/n)."""
    if direction not in (1, -1):
        raise ValueError("direction must be +1 or -1")
    a = np.atleast_2d(np.asarray(lines, dtype=np.complex128))
    n = a.shape[-1]
    j = np.arange(n)
    out = np.zeros_like(a)
    for k in range(n):
        out[:, k] = a @ np.exp(-direction * 2j * np.pi * j * k / n)
    return out


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
