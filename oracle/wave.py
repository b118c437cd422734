"""Wavefunction algebra of Source/wave.f90, restated on plain complex arrays.

The reference overloads +, - and * for its ``wavefunction`` (coeff(G, k,
band, spinor)) and ``wavefunction_slice`` (coeff(G, band, spinor)) types and
copies between them; every operator allocates its result and combines the
coefficients elementwise.  Shapes are checked by the caller (band counts and
k-points, wave.f90:246-259).

* ``add``   - wave_add_slice_slice / _slice_wfn / _wfn_slice / _wfn_wfn
              (wave.f90:226-406): x + y.
* ``sub``   - wave_sub_* (wave.f90:409-589): x - y.
* ``scale`` - wave_scale_slice_real / _complex, wave_scale_wfn_real /
              _complex (wave.f90:591-736): s * x, s real or complex.
* ``copy``  - wave_copy_* (wave.f90:740-855): a copy of x.

Pinned by identities a wrong sign or a swapped real/imaginary part fails
(tests/test_oracle_wave.py): (x + y) - y = x and x - x = 0 on values whose
sums are exact in binary, i * (i * x) = -x, scaling by (0, 1) swapping the
parts with one sign change, and a hand-written 2-element case.
"""
import numpy as np


def add(x, y):
    return np.asarray(x, dtype=np.complex128) + np.asarray(y, dtype=np.complex128)


def sub(x, y):
    return np.asarray(x, dtype=np.complex128) - np.asarray(y, dtype=np.complex128)


def scale(s, x):
    return complex(s) * np.asarray(x, dtype=np.complex128)


def copy(x):
    return np.array(x, dtype=np.complex128, copy=True)
