"""Reference interchange records (test infrastructure: only tests/ use it).

Restates the byte layout of the files the reference writes with gfortran
and ``-fconvert=big-endian`` (``Makefile:86``) through its defined
unformatted I/O procedures:

* ``pot_write`` (``Source/pot.f90:194-224``): standard-grid dims, then
  ``nc_pot(nfine, 2, 2)`` and ``allocated``;
* ``density_write`` (``Source/density.f90:133-163``): dims, ``charge(nfine)``,
  ``spin(nfine, 3)``, ``allocated``;
* ``wave_write`` (``Source/wave.f90:896-926``): dims,
  ``coeff(nstd, nk, nb, 2)``, ``nbands``, ``kpts``, ``allocated``.

The statement ``write(unit) x`` with a derived-type ``x`` calls the type's
procedure, whose child data transfer statements continue the parent's
record (Fortran 2008, 9.6.4.8.3): one record per file.  gfortran's record
framing ("Unformatted sequential file format" in its manual): a 4-byte
marker holding the payload length before and after the payload; a record
longer than 2147483639 bytes is split into subrecords, whose leading marker
is negative when more subrecords follow and whose trailing marker is
negative when the subrecord continues an earlier one.  LOGICAL(4) .true. is
stored as 1.

Pinned by: a record spelled out byte by byte from that description, a
split record likewise, the mapping of standard-grid positions by brute-force
enumeration of ``basis.f90:124-135``'s loop (tests/test_interchange.py).
"""
import struct

import numpy as np

MAX_SUBRECORD = 2147483639


def frame(payload, max_sub=MAX_SUBRECORD):
    """One record around ``payload`` (bytes), split into subrecords."""
    n = len(payload)
    if n == 0:
        return struct.pack(">i", 0) * 2
    out = []
    pos = 0
    first = True
    while pos < n or first:
        take = min(max_sub, n - pos)
        more = pos + take < n
        out.append(struct.pack(">i", -take if more else take))
        out.append(payload[pos:pos + take])
        out.append(struct.pack(">i", take if first else -take))
        pos += take
        first = False
    return b"".join(out)


def unframe(data):
    """Payload of the first record of ``data`` (bytes)."""
    pos = 0
    parts = []
    first = True
    while True:
        (head,) = struct.unpack(">i", data[pos:pos + 4])
        n = abs(head)
        parts.append(data[pos + 4:pos + 4 + n])
        (tail,) = struct.unpack(">i", data[pos + 4 + n:pos + 8 + n])
        if abs(tail) != n or (tail < 0) == first:
            raise ValueError("record markers do not match")
        pos += 8 + n
        first = False
        if head >= 0:
            return b"".join(parts)


def _c128(a):
    return np.asarray(a, dtype=np.complex128).ravel(order="F").astype(">c16").tobytes()


def _i32(*v):
    return struct.pack(">%di" % len(v), *v)


def pot_payload(std_dims, nc_pot, allocated=True):
    return _i32(*std_dims) + _c128(nc_pot) + _i32(1 if allocated else 0)


def den_payload(std_dims, charge, spin, allocated=True):
    return _i32(*std_dims) + _c128(charge) + _c128(spin) + _i32(1 if allocated else 0)


def wvfn_payload(std_dims, coeff, allocated=True):
    """coeff: complex (nstd, nk, nb, 2)."""
    nb, nk = coeff.shape[2], coeff.shape[1]
    return _i32(*std_dims) + _c128(coeff) + _i32(nb, nk, 1 if allocated else 0)


def std_positions(fine_dims, std_dims):
    """{(mx, my, mz): standard-grid position} by the loop of basis.f90:124-135,
    and the same Miller index for every fine-grid FFT bin (bin = m mod n)."""
    pos = {}
    ig = 0
    for mx in range(-(std_dims[0] // 2), std_dims[0] - std_dims[0] // 2):
        for my in range(-(std_dims[1] // 2), std_dims[1] - std_dims[1] // 2):
            for mz in range(-(std_dims[2] // 2), std_dims[2] - std_dims[2] // 2):
                pos[(mx, my, mz)] = ig
                ig += 1
    miller = {}
    for mx in range(-(fine_dims[0] // 2), fine_dims[0] - fine_dims[0] // 2):
        for my in range(-(fine_dims[1] // 2), fine_dims[1] - fine_dims[1] // 2):
            for mz in range(-(fine_dims[2] // 2), fine_dims[2] - fine_dims[2] // 2):
                b = ((mx % fine_dims[0]) * fine_dims[1] + (my % fine_dims[1])) * fine_dims[2] + mz % fine_dims[2]
                miller[b] = (mx, my, mz)
    return pos, miller


def std_index(fine_dims, std_dims, fft_index):
    pos, miller = std_positions(fine_dims, std_dims)
    return np.array([pos.get(miller[int(b)], -1) for b in fft_index], dtype=np.int64)


def sphere_to_coeff(std_dims, psis, std_idx):
    """coeff(nstd, nk, nb, 2) with psis[k] (npw_k x nb) at std_idx[k], spinor 1."""
    nstd = int(np.prod(std_dims))
    nk, nb = len(psis), psis[0].shape[1]
    coeff = np.zeros((nstd, nk, nb, 2), dtype=np.complex128, order="F")
    for k in range(nk):
        coeff[std_idx[k], k, :, 0] = psis[k]
    return coeff
