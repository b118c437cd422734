"""Benchmark systems and their DEReK ``.info`` form.

The five workloads are cubic cells with Gaussian wells.  Their keys map onto
the reference's input language (Source/io.f90:118-161): ``cut_off_energy``,
``g_fine_scale``, ``kpt_mp_grid``, ``n_electrons``, ``electronic_temp`` (always
Kelvin, io.f90:340), ``$DATA lattice`` and ``$DATA gaussian_pot`` rows
``cx cy cz height width`` (io.f90:1888-1922).  Build-specific keys:
``nbands`` (scalar spin-degenerate bands; the reference's
``n_electrons + conduction_bands`` counts spinor bands, io.f90:316),
``cos_pot_amp`` (V += amp*cos(2 pi x/L)), ``wfn_seed`` and ``precision``.

Well parameters come from a seeded SplitMix64 stream: centre ~ U[0,1)^3,
depth ~ U[1,3] Ha (height = -depth), width ~ U[0.8,1.5] bohr.
"""
from dataclasses import dataclass, field
from typing import List

import numpy as np

_GOLDEN = 0x9E3779B97F4A7C15
_MASK = (1 << 64) - 1


def _mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
    return z ^ (z >> 31)


def seeded_uniforms(seed, stream, n):
    key = _mix(((seed * _GOLDEN) & _MASK) ^ _mix((stream + _GOLDEN) & _MASK))
    return [(_mix((key + (i + 1) * _GOLDEN) & _MASK) >> 11) * (1.0 / 9007199254740992.0)
            for i in range(n)]


def seeded_wells(seed, nwells):
    """Fractional centres, heights (Ha) and widths (bohr) of ``nwells`` wells."""
    u = seeded_uniforms(seed, (2 << 48), 5 * nwells)
    centres, heights, widths = [], [], []
    for i in range(nwells):
        a = u[5 * i:5 * i + 5]
        centres.append([a[0], a[1], a[2]])
        heights.append(-(1.0 + 2.0 * a[3]))
        widths.append(0.8 + 0.7 * a[4])
    return centres, heights, widths


# Boltzmann constant in Hartree per Kelvin as the reference builds it
# (units.f90:81 kelvin = boltzmann_si*joule; constants.f90 CODATA 2018).
_c, _h, _e, _me = 299792458.0, 6.62607015e-34, 1.602176634e-19, 9.1093837015e-31
_NA, _R = 6.02214076e23, 8.314462618
_hbar = _h / (2.0 * 3.141592653589793238462643383279502884197)
_mu0 = 4.0 * 3.141592653589793238462643383279502884197 * 1.0e-7
_eps0 = 1.0 / (_mu0 * _c ** 2)
_alpha = _e ** 2 / (4.0 * 3.141592653589793238462643383279502884197 * _eps0 * _hbar * _c)
_joule = 1.0 / (_alpha ** 2 * _me * _c ** 2)
KELVIN_HA = (_R / _NA) * _joule


@dataclass
class SystemSpec:
    name: str
    length: float
    ecut: float
    g_fine_scale: float
    mp: tuple
    nelec: int
    nbands: int
    kT: float = 0.0
    nwells: int = 0
    cos_amp: float = 0.0
    seed: int = 1
    centres: List = field(default_factory=list)
    heights: List = field(default_factory=list)
    widths: List = field(default_factory=list)

    def __post_init__(self):
        if self.nwells and not self.centres:
            self.centres, self.heights, self.widths = seeded_wells(self.seed, self.nwells)

    @property
    def lengths(self):
        return [self.length] * 3

    def to_info(self, energy_tol=1e-7, max_scf=40, precision="fp64", extra=()):
        """The system as a reference-style .info file; ``extra`` adds
        ``key : value`` lines (e.g. ("write_density", "true"))."""
        lines = [
            "! %s: generated benchmark system" % self.name,
            "unit_energy : Ha",
            "unit_length : bohr",
            "xc_functional : LDA",
            "cut_off_energy : %r" % self.ecut,
            "g_fine_scale : %r" % self.g_fine_scale,
            "kpt_mp_grid : %d %d %d" % tuple(self.mp),
            "n_electrons : %d" % self.nelec,
            "nbands : %d" % self.nbands,
            "electronic_temp : %r" % (self.kT / KELVIN_HA if self.kT > 0 else 0.0),
            "energy_tol : %r" % energy_tol,
            "max_scf : %d" % max_scf,
            "cos_pot_amp : %r" % self.cos_amp,
            "wfn_seed : %d" % self.seed,
            "precision : %s" % precision,
        ] + ["%s : %s" % kv for kv in extra] + [
            # block rows start in column 1: the reference treats a row that
            # starts with a blank as a comment (io.f90:1818-1824)
            "$DATA lattice",
            "%r 0.0 0.0" % self.length,
            "0.0 %r 0.0" % self.length,
            "0.0 0.0 %r" % self.length,
            "$ENDDATA lattice",
        ]
        if self.nwells:
            lines.append("$DATA gaussian_pot")
            for c, h, w in zip(self.centres, self.heights, self.widths):
                lines.append("%r %r %r %r %r" % (c[0], c[1], c[2], h, w))
            lines.append("$ENDDATA gaussian_pot")
        return "\n".join(lines) + "\n"


WORKLOADS = ("gamma32", "gamma108", "kmesh72", "kmesh144")


def workload(name):
    """The benchmark systems, all cubic cells with seeded Gaussian wells:

    gamma32   10 bohr, Ecut 10 Ha, Gamma, 32^3 fine grid, 8 electrons, 10 bands, 4 wells
    gamma108  20 bohr, Ecut 30 Ha, Gamma, 108^3, 128 electrons, 80 bands, 16 wells
              + 0.2 Ha cos(2 pi x / L), kT 0.002 Ha
    kmesh72   16 bohr, Ecut 20 Ha, 4x4x4 mesh, 72^3, 32 electrons, 24 bands, 8 wells, kT 0.01 Ha
    kmesh144  30 bohr, Ecut 25 Ha, 2x2x2 mesh, 144^3, 512 electrons, 300 bands, 64 wells, kT 0.002 Ha
    """
    if name == "gamma32":
        return SystemSpec("gamma32", 10.0, 10.0, 2.15, (1, 1, 1), 8, 10, nwells=4, seed=1)
    if name == "gamma108":
        # metallic: the gap between bands 64 and 65 closes, fixed occupations
        # slosh; a 0.002 Ha Fermi-Dirac smearing (16 kT below the top band) converges
        return SystemSpec("gamma108", 20.0, 30.0, 2.15, (1, 1, 1), 128, 80, kT=0.002, nwells=16, cos_amp=0.2,
                          seed=2)
    if name == "kmesh72":
        return SystemSpec("kmesh72", 16.0, 20.0, 2.2, (4, 4, 4), 32, 24, kT=0.01, nwells=8, seed=4)
    if name == "kmesh144":
        return SystemSpec("kmesh144", 30.0, 25.0, 2.125, (2, 2, 2), 512, 300, kT=0.002, nwells=64, seed=5)
    raise ValueError("benchmark systems: " + ", ".join(WORKLOADS))


def small_system(seed=11, length=8.0, ecut=6.0, g=2.0, mp=(1, 1, 1), nelec=4, nbands=6,
                 kT=0.0, nwells=3, cos_amp=0.0):
    """A fast test system (grid ~20^3, a few hundred plane waves)."""
    return SystemSpec("small", length, ecut, g, mp, nelec, nbands, kT=kT, nwells=nwells,
                      cos_amp=cos_amp, seed=seed)
