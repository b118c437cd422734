"""Reference discretisation: grid sizes, G/real/fractional maps, k-points.

Restates ``Source/basis.f90:65-309`` (``basis_init``, ``basis_prime_fact``),
``Source/io.f90:1938-1980`` (``io_lattice_to_abc``), ``Source/io.f90:2472-2498``
(``io_inv``) and ``Source/io.f90:1992-2035`` (``io_kpoint_grid``) with the same
floating-point operation order, so the maps are bit-identical to the
reference's for orthorhombic cells (the only cells whose 3x3 inverse does not
depend on LAPACK rounding).

Pinned by: known grid sizes for the benchmark cells, brute-force smoothness
search for ``prime_fact``, exact enumeration order/range checks, and the
requirement that the FFT bin of every map entry round-trips (tests/test_oracle_basis.py).
"""
import numpy as np

PI = 3.141592653589793238462643383279502884197   # constants.f90:29
TWOPI = 2.0 * PI                                  # constants.f90:30
PRIMES = (2, 3, 5, 7, 11, 13)                     # basis.f90:272


def prime_fact(n):
    """Smallest n' >= n whose prime factors are in {2,3,5,7,11,13}.

    basis.f90:257-309: up to 100 increments; each inner sweep divides by every
    prime at most once and sweeps repeat (up to 100) until nothing divides.
    """
    n = int(n)
    for _ in range(100):
        n_p = n
        factors = [0] * len(PRIMES)
        for _ in range(100):
            old = n_p
            for i, p in enumerate(PRIMES):
                if n_p % p == 0:
                    n_p //= p
                    factors[i] += 1
            if old == n_p:
                break
        tot = 1
        for p, f in zip(PRIMES, factors):
            tot *= p ** f
        if tot == n:
            break
        n += 1
    return n


def inv_cell_orthorhombic(cell):
    """``transpose(twopi*io_inv(cell))`` (io.f90:1954) for a diagonal cell.

    DGETRF/DGETRI of a diagonal matrix produce ``1/a_jj`` on the diagonal
    (DTRTI2: ``A(j,j) = ONE/A(j,j)``) and signed zeros elsewhere; the product
    with twopi is then ``fl(twopi * fl(1/a_jj))``.
    """
    cell = np.asarray(cell, dtype=np.float64)
    if np.count_nonzero(cell - np.diag(np.diag(cell))):
        raise ValueError("only orthorhombic (diagonal) cells have a LAPACK-independent inverse")
    inv = np.zeros((3, 3))
    for j in range(3):
        inv[j, j] = TWOPI * (1.0 / cell[j, j])
    return inv.T.copy()


def lattice_lengths(cell):
    """Column norms, io.f90:1956-1966 (sqrt(c1j**2 + c2j**2 + c3j**2))."""
    cell = np.asarray(cell, dtype=np.float64)
    return [float(np.sqrt(cell[0, j] ** 2 + cell[1, j] ** 2 + cell[2, j] ** 2)) for j in range(3)]


def volume(cell):
    """io_volume, io.f90:1983-1989 (exact for diagonal cells; the reference's
    middle term ``b(1)*a(3)-a(3)*b(1)`` is kept as written)."""
    a, b, c = (np.asarray(cell, dtype=np.float64)[:, j] for j in range(3))
    return float(c[0] * (a[1] * b[2] - a[2] * b[1]) + c[1] * (b[0] * a[2] - a[2] * b[0])
                 + c[2] * (a[0] * b[1] - a[1] * b[0]))


def grid_dims(cell, ecut, g_fine_scale):
    """Standard and fine FFT grid sizes, basis.f90:87-105."""
    k_cut = np.sqrt(2.0 * ecut)
    la = lattice_lengths(cell)
    std = [prime_fact(int(np.ceil(k_cut * l / PI))) for l in la]
    fine = [prime_fact(int(np.ceil(g_fine_scale * k_cut * l / PI))) for l in la]
    return std, fine


def m_range(n):
    """Integer G indices in enumeration order: -n/2 .. n-n/2-1 (Fortran
    truncating division of a positive n), basis.f90:124-126."""
    return np.arange(-(n // 2), n - n // 2, dtype=np.int64)


def g_map(inv_cell, dims):
    """``grid_points``/``fine_grid_points`` (N,3): x outer, z inner,
    ``matmul(inv_cell, (mx,my,mz))`` (basis.f90:124-149).

    Also returns the integer indices (N,3).  For diagonal inv_cell each
    component is ``fl(inv_cell(i,i)*m_i)`` plus exact (signed) zeros.
    """
    mx, my, mz = (m_range(n) for n in dims)
    M = np.stack(np.meshgrid(mx, my, mz, indexing="ij"), axis=-1).reshape(-1, 3)
    Mf = M.astype(np.float64)
    G = np.empty_like(Mf)
    for i in range(3):
        acc = inv_cell[i, 0] * Mf[:, 0]
        acc = acc + inv_cell[i, 1] * Mf[:, 1]
        acc = acc + inv_cell[i, 2] * Mf[:, 2]
        G[:, i] = acc
    return G, M


def real_map(cell, dims):
    """``real_*grid_points`` and ``*frac_points`` (basis.f90:155-188):
    frac = fl(fl(i*1.0)/n) for i = 1..n (so the last point is 1.0), real =
    matmul(cell, frac)."""
    fx, fy, fz = ((np.arange(1, n + 1, dtype=np.float64) * 1.0) / float(n) for n in dims)
    F = np.stack(np.meshgrid(fx, fy, fz, indexing="ij"), axis=-1).reshape(-1, 3)
    cell = np.asarray(cell, dtype=np.float64)
    R = np.empty_like(F)
    for i in range(3):
        acc = cell[i, 0] * F[:, 0]
        acc = acc + cell[i, 1] * F[:, 1]
        acc = acc + cell[i, 2] * F[:, 2]
        R[:, i] = acc
    return R, F


def kpoint_grid(mp):
    """Monkhorst-Pack fractions, io.f90:2016-2031: k = n*dk - (1+dk)/2,
    n = 1..N, x outer / z inner."""
    dk = [1.0 / float(m) for m in mp]
    out = []
    for nx in range(1, mp[0] + 1):
        for ny in range(1, mp[1] + 1):
            for nz in range(1, mp[2] + 1):
                out.append([nx * dk[0] - (1.0 + dk[0]) / 2,
                            ny * dk[1] - (1.0 + dk[1]) / 2,
                            nz * dk[2] - (1.0 + dk[2]) / 2])
    return np.array(out, dtype=np.float64).reshape(-1, 3)


def fft_bin(M, dims):
    """Flat index of integer frequency m in the reference FFT layout (flat,
    z fastest, x slowest; fft.f90:258-273): bin = m mod n per axis."""
    b = np.mod(M, np.asarray(dims, dtype=np.int64))
    return ((b[:, 0] * dims[1] + b[:, 1]) * dims[2] + b[:, 2]).astype(np.int64)


def kcart(inv_cell, kfrac):
    """Cartesian k = matmul(inv_cell, k_frac) in the map's operation order."""
    k = np.empty(3)
    for i in range(3):
        k[i] = inv_cell[i, 0] * kfrac[0] + inv_cell[i, 1] * kfrac[1] + inv_cell[i, 2] * kfrac[2]
    return k


def sphere(G, M, dims, kc, ecut):
    """Plane-wave set of one k-point: fine-map entries with 1/2|G+k|^2 <= ecut,
    in map order.  Returns (fft_index int32, kinetic 1/2|G+k|^2, G+k).

    Operation order (no fused multiply-add): gk = G + k; g2 = gx*gx + gy*gy + gz*gz
    (left to right); kin = 0.5*g2.
    """
    gk = G + kc[None, :]
    g2 = gk[:, 0] * gk[:, 0]
    g2 = g2 + gk[:, 1] * gk[:, 1]
    g2 = g2 + gk[:, 2] * gk[:, 2]
    kin = 0.5 * g2
    sel = kin <= ecut
    return fft_bin(M[sel], dims).astype(np.int32), kin[sel].copy(), gk[sel].copy()


def gsq_on_fft_grid(inv_cell, dims):
    """|G|^2 for every FFT bin of the fine grid (flat reference layout), using
    the map's G components for the representative m in -n/2..n-n/2-1."""
    G, M = g_map(inv_cell, dims)
    g2 = G[:, 0] * G[:, 0]
    g2 = g2 + G[:, 1] * G[:, 1]
    g2 = g2 + G[:, 2] * G[:, 2]
    out = np.empty(int(np.prod(dims)))
    out[fft_bin(M, dims)] = g2
    return out
