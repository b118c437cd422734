"""SCF driver on the GPU library, shaped like ``electronic_scf``
(Source/electronic.f90:86-175): loop up to ``max_scf`` steps, print the
step/energy/change/time table, stop when |E_i - E_{i-1}| < energy_tol
(electronic.f90:156-165).

The discretisation comes from the library's restatement of
Source/basis.f90 (grid sizes, fine G map, plane-wave sphere, |G|^2) and the
external potential from its Gaussian-well evaluator, so a run needs nothing
but a system description (``derek_amd.config.SystemSpec``) or a ``.info``
file (``lib.Info``).
"""
import time

import numpy as np

from . import lib


class Setup:
    """Discretisation of one system, computed by the library (host side)."""

    def __init__(self, spec, kfrac=None):
        """kfrac: fractional k-points (nk x 3) instead of the system's
        Monkhorst-Pack mesh (e.g. a band-structure path), equal weights."""
        self.spec = spec
        self.cell = np.diag([float(x) for x in spec.lengths])
        self.std_dims, self.dims = lib.grid_dims(self.cell, spec.ecut, spec.g_fine_scale)
        self.N = int(np.prod(self.dims))
        self.inv = lib.inv_cell(self.cell)
        self.volume = float(np.prod(spec.lengths))
        self.kfrac = lib.kpoints(spec.mp) if kfrac is None else np.atleast_2d(np.asarray(kfrac, dtype=np.float64))
        self.nk = self.kfrac.shape[0]
        self.weights = np.full(self.nk, 1.0 / self.nk)
        self.fft_index, self.kin = [], []
        for k in range(self.nk):
            idx, kin = lib.sphere(self.inv, self.dims, self.kfrac[k], spec.ecut)
            self.fft_index.append(idx)
            self.kin.append(kin)
        self.npw = [len(i) for i in self.fft_index]
        self.gsq = lib.gsq(self.inv, self.dims)
        self.vext = lib.gaussian_pot(self.cell, self.dims, spec.centres, spec.heights, spec.widths,
                                     spec.cos_amp)

    def context(self, precision=lib.DEREK_FP64, nbands=None):
        nb = self.spec.nbands if nbands is None else nbands
        ctx = lib.Context(self.dims, self.nk, nb, self.volume, max(self.npw), precision)
        for k in range(self.nk):
            ctx.set_basis(k, self.fft_index[k], self.kin[k], self.weights[k])
        ctx.set_gsq(self.gsq)
        ctx.set_vext(self.vext)
        return ctx


def run(spec, max_scf=40, energy_tol=1e-7, rho_tol=None, precision=lib.DEREK_FP64, log=print,
        n_david_first=4, n_david=2, beta=0.5, q0=1.0, history=8, setup=None, ctx=None, kbatch=0, graph=False,
        autotune=False, switch_tol=0.0, refine_steps=0, gemm_bf16=None):
    """Run the SCF; returns a dict with the final energy, eigenvalues,
    per-step reports, wall time (init included in ``wall_init``) and the
    converged flag.  In DEREK_MIXED mode a step may only end the run once
    its report's ``refine_left`` is 0 (the FP64 refinement is done).
    ``gemm_bf16`` (1..3) runs the complex64 overlaps and rotations on the
    BF16 matrix cores with that many bf16 terms per value, 0 on the FP32 ones
    (None: the context's setting, 2 for DEREK_MIXED)."""
    su = setup if setup is not None else Setup(spec)
    own = ctx is None
    if ctx is None:
        ctx = su.context(precision)
    if gemm_bf16 is not None:
        ctx.set_gemm(gemm_bf16)
    ti = time.perf_counter()
    ctx.scf_init(spec.nelec, kT=spec.kT, beta=beta, q0=q0, n_david_first=n_david_first, n_david=n_david,
                 history=history, seed=spec.seed, kbatch=kbatch, graph=graph, autotune=autotune,
                 switch_tol=switch_tol, refine_steps=refine_steps)
    wall_init = time.perf_counter() - ti
    hist = []
    e_old = None
    converged = False
    t0 = time.perf_counter()
    if log:
        log(" +--------+-------------------+-------------------+---------+ <-- SCF")
        log(" |  SCF   |      Energy       |   Energy Change   |  Time   | <-- SCF")
        log(" |  Step  |       (Ha)        |        (Ha)       |   (s)   | <-- SCF")
        log(" +--------+-------------------+-------------------+---------+ <-- SCF")
    for _ in range(max_scf):
        r = ctx.scf_step()
        r["time"] = time.perf_counter() - t0
        hist.append(r)
        de = None if e_old is None else r["etot"] - e_old
        if log:
            des = "                 " if de is None else "%17.8E" % de
            log(" | %5d  | %17.8E | %s | %7.2f | <-- SCF" % (r["iscf"], r["etot"], des, r["time"]))
        ok = de is not None and abs(de) < energy_tol and r["refine_left"] == 0
        if rho_tol is not None:
            ok = ok and r["drho"] < rho_tol
        e_old = r["etot"]
        if ok:
            converged = True
            break
    wall = time.perf_counter() - t0
    if log:
        log(" +--------+-------------------+-------------------+---------+ <-- SCF")
    out = dict(etot=hist[-1]["etot"], eigs=ctx.eigenvalues(), occ=ctx.occupations(), history=hist,
               converged=converged, iterations=len(hist), wall=wall, wall_init=wall_init,
               fp32_steps=sum(1 for h in hist if h["fp32"]), tuning=ctx.tuning())
    if own:
        out["rho"] = ctx.get_density()
        ctx.close()
    return out
