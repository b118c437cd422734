"""FP64 CPU oracle for DEReK's Kohn-Sham SCF hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library or its host layer
imports, calls or links anything in this package: it is the checker used by
``tests/``, by ``__graft_entry__.smoke()`` and by the CPU-baseline leg of
``bench.py``.  It restates, in plain numpy FP64,

* the reference's discretisation (``Source/basis.f90``, ``Source/io.f90``
  lattice/k-point routines, ``Source/fft.f90`` layout and sign conventions),
* the parts of the SCF the reference leaves as stubs (``Source/density.f90``
  ``density_calculate``, ``Source/pot.f90`` ``pot_apply_wfn``,
  ``Source/wave.f90`` ``wave_dot_*``, ``Source/xc.f90``,
  ``Source/electronic.f90`` ``electronic_scf``) as textbook plane-wave KS-DFT,
  with the exchange-correlation functional following the vendored libxc
  (``ExLib/libxc/src/lda_x.c``, ``ExLib/libxc/src/lda_c_pz.c``).

How each function is pinned (golden vectors, closed forms, brute force) is
listed in the module docstrings and exercised by ``tests/test_oracle_*.py``.
"""
