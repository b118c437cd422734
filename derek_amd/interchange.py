"""The reference's binary interchange files for a system set up by
``derek_amd.scf.Setup``: the total / external potential (.pot, .potex,
pot.f90:194-224,303-340), the density (.den, density.f90:133-163) and the
wavefunctions (.wvfn, wave.f90:896-926), through the library's writers and
readers (include/derek_hip.h, ``derek_io_*``).

The reference stores the potential as a 2x2 block nc_pot(:, 2, 2); a scalar
potential V sits on the diagonal (nc_pot(:,1,1) = nc_pot(:,2,2) = V) and is
read back as their mean (xsf.f90:43-45).  The density is the charge with
zero spin density; the wavefunctions are the sphere coefficients placed on
the standard grid, spinor 1, zeros elsewhere.
"""
import numpy as np

from . import lib


def nc_pot_from_v(v):
    v = np.asarray(v, dtype=np.float64)
    nc = np.zeros((len(v), 2, 2), dtype=np.complex128, order="F")
    nc[:, 0, 0] = v
    nc[:, 1, 1] = v
    return nc


def v_from_nc_pot(nc):
    return (0.5 * (nc[:, 0, 0] + nc[:, 1, 1])).real.copy()


def write_potential(path, setup, v):
    lib.pot_write(path, setup.std_dims, nc_pot_from_v(v))


def read_potential(path, setup):
    """Scalar potential of a .pot / .potex record on this system's grid."""
    nc, _ = lib.pot_read(path, setup.std_dims, setup.N)
    return v_from_nc_pot(nc)


def write_density(path, setup, rho):
    lib.den_write(path, setup.std_dims, np.asarray(rho, dtype=np.float64).astype(np.complex128))


def read_density(path, setup):
    ch, _, _ = lib.den_read(path, setup.std_dims, setup.N)
    return ch.real.copy()


def std_indices(setup):
    """Standard-grid positions of each k-point's plane waves."""
    out = []
    for k in range(setup.nk):
        idx = lib.std_index(setup.dims, setup.std_dims, setup.fft_index[k])
        if (idx < 0).any():
            raise ValueError("a plane wave of k-point %d lies outside the standard grid" % k)
        out.append(idx)
    return out


def write_wvfn(path, setup, ctx):
    """Every k-point's bands of a context's resident wavefunctions."""
    psis = [ctx.get_psi(k).T for k in range(setup.nk)]  # npw x nbands
    lib.wvfn_write(path, setup.std_dims, psis, std_indices(setup))


def read_wvfn(path, setup, nbands):
    """Per k-point (npw x nbands) sphere coefficients of a .wvfn record."""
    psis, _ = lib.wvfn_read(path, setup.std_dims, std_indices(setup), nbands)
    return psis


def load_wvfn(path, setup, ctx):
    """Reads a .wvfn record into a context's resident wavefunctions."""
    for k, p in enumerate(read_wvfn(path, setup, ctx.nbands)):
        ctx.put_psi(k, p.T)
