"""MI355X-native Kohn-Sham SCF hot path for DEReK (host layer).

The compute lives in ``libderek_hip.so`` (HIP, gfx950) behind the C-ABI in
``include/derek_hip.h``; this package loads it (``derek_amd.lib``), reads
DEReK ``.info`` inputs (``lib.Info``, the library's reader), builds the benchmark systems
(``derek_amd.config``) and drives the SCF loop the way
``Source/electronic.f90:86-175`` does (``derek_amd.scf``).
"""
