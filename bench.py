"""Benchmark of the Kohn-Sham SCF hot path on one MI355X.

Headline workload (``value``): the plain FP64 system of a 20 bohr cubic
cell, Ecut 30 Ha (108^3 FFT grid, 62669 plane waves at Gamma), 128
electrons, 80 bands, 16 seeded Gaussian wells plus a 0.2 Ha cosine term.
A step is one H|psi> over all 80 resident bands (sphere-pruned inverse 3D
FFT, V(r) multiply, forward 3D FFT, gather + kinetic); value = bands/s over
the K timed steps with the inputs resident in HBM.  ``roofline`` is the
dominant kernel of that step against HBM, with the path's compulsory bytes
per band (2 npw s_psi + 12 npw + N_g s_r / B) times the bands of one launch
over its duration; ``roofline_valu`` is the same kernel against the FP64
vector ALU (its FFT flops).  The same run reports the time to SCF
convergence (|dE| < 1e-6 Ha) of

  * the 32^3 system (10 bohr, Ecut 10 Ha, 10 bands) in FP64, beside the
    numpy oracle's SCF of the same system on one CPU core;
  * the headline system in plain FP64 and in the mixed form (complex64 bulk,
    FP64 subspace solves and energies, FP64 refinement), with dE_tot and the
    largest eigenvalue difference between the two, and the FP64 step against
    its own roofline (compulsory bytes at 8 TB/s plus GEMM flops at the FP64
    matrix-core peak);
  * the 64-k-point system (16 bohr, Ecut 20 Ha, 4x4x4 mesh, 72^3 grid, 24
    bands, kT = 0.01 Ha) in plain FP64 (one k-point per launch, no graph),
    FP64 batched (all k-points per launch, graph replay) and mixed + batched
    + autotuned, with the speed-up of the last over the first;
  * the 8-k-point system (30 bohr, Ecut 25 Ha, 2x2x2 mesh, 144^3 grid,
    161214 plane waves per k, 300 bands) in FP64 batched and mixed batched;
  * the mixed runs put their complex64 overlaps and rotations on the BF16
    matrix cores with ``--bf16 N`` bf16 terms per value (default 2, the
    library's default; FP32 accumulation), and each is repeated with them on
    the FP32 matrix cores (``*_fp32_mfma``) for comparison (``--bf16 0``: the
    FP32 ones only, no repeat);

and CPU baselines, timed on this host before the GPU is touched: the numpy
oracle's H|psi> on a bounded sample of the headline system (one core and
every core of the host's share), one oracle FP64 SCF step of the headline
system (one core and every core) and the oracle's whole SCF of the 32^3
system (one core).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-scf] [--scf gamma32,gamma108,kmesh72,kmesh144]
                [--cpu-seconds S] [--no-cpu] [--fixed-shapes | --shapes BPW,ITEMS,FUSED] [--tiles ZS,ROUNDS[,TILE]]
                [--bf16 N]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("time-to-SCF (|ΔE|<1e-6 Ha) mixed/batched vs FP64 GPU; H|ψ⟩ bands/s; HBM GB/s")
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6  # FP64 vector (and dense MFMA) peak


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_calibration(mib):
    """One hipMemcpy device -> device of `mib` MiB (read mib, write mib)."""
    import ctypes as C
    hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"# copy calibration: {mib} MiB device to device")


def hpsi_rate(ctx, nb, steps, warmup):
    """bands/s of H|psi> on the resident bands (timed without events), then
    the same steps again with a HIP event pair around every launch for the
    per-kernel-class durations."""
    for _ in range(warmup):
        ctx.apply_h_resident(0, nb)
    ctx.sync()
    t1 = time.perf_counter()
    for _ in range(steps):
        ctx.apply_h_resident(0, nb)
    ctx.sync()
    dt = time.perf_counter() - t1
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(steps):
        ctx.apply_h_resident(0, nb)
    ctx.sync()
    prof = ctx.profile_read()
    ctx.profile(False)
    return steps * nb / dt, dt / steps * 1e3, prof


def _cpu_hpsi_worker(args):
    """Oracle H|psi> on 2-band batches for about `seconds` (one process)."""
    idx, kin, vext, dims, seconds, seed = args
    import numpy as np
    from oracle import ks as oks
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=(2, len(idx))) + 1j * rng.normal(size=(2, len(idx)))
    done, t0 = 0, time.perf_counter()
    while True:
        oks.apply_h(psi, idx, kin, vext, dims)
        done += 2
        if time.perf_counter() - t0 > seconds:
            return done, time.perf_counter() - t0


def cpu_baseline(su, seconds):
    """The numpy oracle's H|psi> on the headline system: one process, then
    one process per host core (bands split across processes).  Runs before
    the GPU is touched: the pool forks this process."""
    import multiprocessing as mp
    args = (su.fft_index[0], su.kin[0], su.vext, su.dims, seconds, 0)
    done, tc = _cpu_hpsi_worker(args)
    serial = {"value": round(done / tc, 3), "unit": "bands/s", "cores": 1, "kind": "port",
              "sample": f"numpy oracle H|psi> on the same system, {done} band applications in batches of 2 "
                        f"({tc:.1f} s, pocketfft single-threaded)"}
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count() or 1
    ncores = max(1, min(ncores, 16))  # the box's CPU share for one GPU
    with mp.get_context("fork").Pool(ncores) as pool:
        res = pool.map(_cpu_hpsi_worker, [args[:5] + (i + 1,) for i in range(ncores)])
    done_all = sum(r[0] for r in res)
    t_all = max(r[1] for r in res)
    allc = {"value": round(done_all / t_all, 3), "unit": "bands/s", "cores": ncores, "kind": "port",
            "sample": f"the same on {ncores} processes (one per core), {done_all} band applications in "
                      f"{t_all:.1f} s"}
    return serial, allc


def cpu_scf_baseline(ncores):
    """The numpy oracle's FP64 SCF on this host, before the GPU is touched:
    one SCF step of the headline system (initial wavefunctions, V_in, H X,
    Rayleigh-Ritz and two block-Davidson steps, occupations, density,
    energies, Pulay/Kerker mixing) on one core and on ``ncores`` threads
    (pocketfft workers and OpenBLAS threads), and the whole SCF of the
    32^3 system (10 bands) to |dE| < 1e-6 Ha on one core."""
    from threadpoolctl import threadpool_limits

    from derek_amd import config
    from oracle import fft as offt
    from oracle import scf as oscf

    def system(spec):
        return oscf.System(spec.lengths, spec.ecut, spec.g_fine_scale, spec.mp, spec.nelec, spec.nbands, spec.kT,
                           spec.centres, spec.heights, spec.widths, spec.cos_amp, spec.seed)

    out = {}
    s32 = system(config.workload("gamma32"))
    offt.set_workers(1)
    with threadpool_limits(1):
        t = time.perf_counter()
        r = oscf.scf(s32, max_scf=80, energy_tol=1e-6)
        dt = time.perf_counter() - t
    out["gamma32_time_to_scf"] = {"value": round(dt, 3), "unit": "s", "cores": 1, "kind": "port",
                                  "steps": r["iterations"], "converged": r["converged"],
                                  "etot_ha": float(r["etot"]),
                                  "sample": "numpy oracle FP64 SCF of the 32^3 system to |dE| < 1e-6 Ha"}
    s108 = system(config.workload("gamma108"))
    for n in sorted({1, ncores}):
        offt.set_workers(n)
        with threadpool_limits(n):
            t = time.perf_counter()
            oscf.scf(s108, max_scf=1, energy_tol=0.0, n_david_first=2)
            dt = time.perf_counter() - t
        out[f"gamma108_scf_step_{n}core"] = {
            "value": round(dt, 3), "unit": "s per SCF step", "cores": n, "kind": "port",
            "sample": "one numpy oracle FP64 SCF step (two Davidson steps) of the 108^3 system, 80 bands"}
        log(f"# CPU FP64 SCF step (108^3, 80 bands) on {n} core(s): {dt:.2f} s")
    offt.set_workers(1)
    return out


def scf_step_roofline(su, nb, n_david, s_psi, t_step_s):
    """Lower bound on one SCF step of a Gamma system from its compulsory work:
    HBM bytes of H|psi> (nb (1 + n_david) band applications, each
    2 npw s_psi + 12 npw + N_g 8 / nb), of the density (nb npw s_psi in,
    2 N_g 8 for rho) and of the grid stages (Hartree and LDA of rho_in and
    rho_out, Kerker, Pulay with a full 8-deep history: 16+16+24+24+16+(4+32)*8
    N_g bytes), at 8 TB/s, plus the overlap and rotation flops
    (8 npw m n per complex GEMM: Y^H HY of X, nb^2; per Davidson step the
    2nb x 2nb S and H and two 2nb -> nb rotations; two nb -> nb rotations
    after Rayleigh-Ritz) at the FP64 matrix-core peak.  The dense 2nb x 2nb
    solves are latency-bound and not counted."""
    npw, N = float(su.npw[0]), float(su.N)
    hp = nb * (1 + n_david) * (2 * npw * s_psi + 12 * npw + N * 8.0 / nb)
    dens = nb * npw * s_psi + 2 * N * 8.0
    grid = (16 + 16 + 24 + 24 + 16 + (4 + 32) * 8) * N
    flops = 8.0 * npw * nb * nb * (3 + 12 * n_david)
    t_bytes = (hp + dens + grid) / (HBM_PEAK_GBS * 1e9)
    t_flops = flops / (FP64_PEAK_TFS * 1e12)
    t_roof = t_bytes + t_flops
    return {"bytes": hp + dens + grid, "gemm_flops": flops, "t_bytes_ms": round(t_bytes * 1e3, 4),
            "t_flops_ms": round(t_flops * 1e3, 4), "t_roofline_ms": round(t_roof * 1e3, 4),
            "t_step_ms": round(t_step_s * 1e3, 3), "frac": round(t_roof / t_step_s, 4)}


def items_sweep(gscf, lib, config, name, caps, precision, steps=5):
    """H|psi> over every (k-point, band) item of a system's resident bands (all
    k-points in one batch) with at most `cap` items per pipeline launch, for
    each cap (0: every item in one launch per pass): ms per application and
    the launches per pass."""
    import numpy as np
    spec = config.workload(name)
    su = gscf.Setup(spec)
    ctx = su.context(precision)
    ctx.scf_init(spec.nelec, kT=spec.kT, seed=spec.seed)
    ctx.set_vloc(su.vext)
    nk, nb = su.nk, spec.nbands
    out = []
    t0 = ctx.tuning()
    for cap in caps:
        c = cap if cap > 0 else nk * nb
        ctx.set_launch(t0["bands_per_wg"], c, t0["fused"], t0["fused_density"])
        for _ in range(2):
            ctx.apply_h_batch()
        ctx.sync()
        t = time.perf_counter()
        for _ in range(steps):
            ctx.apply_h_batch()
        ctx.sync()
        ms = (time.perf_counter() - t) / steps * 1e3
        out.append({"items_per_launch": c, "launches_per_pass": int(np.ceil(nk * nb / c)), "ms": round(ms, 3),
                    "items": nk * nb, "us_per_item": round(ms * 1e3 / (nk * nb), 3)})
        log(f"# {name} H|psi> over {nk * nb} items, {c} per launch: {ms:.3f} ms")
    ctx.close()
    return out


def scf_run(gscf, lib, spec, su, precision, verbose=False, **kw):
    """Time to |dE| < 1e-6 Ha from a fresh engine (init included)."""
    ctx = su.context(precision)
    t0 = time.perf_counter()
    res = gscf.run(spec, max_scf=80, energy_tol=1e-6, setup=su, ctx=ctx, log=log if verbose else None, **kw)
    wall = time.perf_counter() - t0
    out = {"time_to_scf_s": round(wall, 4), "init_s": round(res["wall_init"], 4), "steps": res["iterations"],
           "fp32_steps": res["fp32_steps"], "ms_per_step": round(res["wall"] / res["iterations"] * 1e3, 2),
           "etot_ha": res["etot"], "converged": res["converged"], "tuning": res["tuning"]}
    eigs = res["eigs"]
    ctx.close()
    return out, eigs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-scf", action="store_true")
    ap.add_argument("--scf", default="gamma32,gamma108,kmesh72,kmesh144")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--refine", type=int, default=0,
                    help="FP64 refinement steps of the mixed runs (0: the library default)")
    ap.add_argument("--bf16", type=int, default=2,
                    help="bf16 terms per value of the mixed runs' GEMMs (0: FP32 matrix cores, no repeat)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--items-sweep", default="",
                    help="SYSTEM: also time H|psi> over every resident (k-point, band) item of that system at 64, "
                         "128, 256, 512 items per launch and all items in one launch (e.g. kmesh72)")
    ap.add_argument("--no-graph", action="store_true",
                    help="run the batched SCF modes eagerly (no hipGraph replay), e.g. under a kernel tracer")
    ap.add_argument("--fixed-shapes", action="store_true",
                    help="time H|psi> at the default launch shapes (fused y/x pass where the grid allows it, 2 bands "
                         "per workgroup, 256 items per launch) instead of autotuned ones")
    ap.add_argument("--shapes", default=None,
                    help="BPW,ITEMS,FUSED: time H|psi> at these launch shapes (bands per workgroup, items per "
                         "launch, fused passes: bit 0 H|psi>, bit 1 density) instead of autotuned ones")
    ap.add_argument("--tiles", default=None,
                    help="ZS,ROUNDS[,TILE]: time H|psi> with these FFT/GEMM tiles (sticks per z-pass workgroup "
                         "16/32, overlap chunk rounds 1/2/4, BF16 overlap tile edge 64/128) instead of autotuned ones")
    ap.add_argument("--heev", default="",
                    help="N1,N2,...: also run the dense eigensolver (lowest n/2 eigenpairs, three calls each) on "
                         "seeded random Hermitian matrices of these orders and report the eigenvalue error against "
                         "numpy (for kernel traces of the solver)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (the known-size kernel that "
                         "scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(k-point sharding over RCCL is a design note, not code); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from derek_amd import config, lib
    from derek_amd import scf as gscf

    spec = config.workload("gamma108")
    t0 = time.perf_counter()
    su = gscf.Setup(spec)  # host-side discretisation: no GPU call yet
    cpu = cpu_all = cpu_scf = None
    if not args.no_cpu:
        cpu, cpu_all = cpu_baseline(su, args.cpu_seconds)
        log(f"# CPU baseline: {cpu['value']} bands/s on 1 core, {cpu_all['value']} bands/s on "
            f"{cpu_all['cores']} cores")
        cpu_scf = cpu_scf_baseline(cpu_all["cores"])
    if args.copy_calib > 0:
        copy_calibration(args.copy_calib)
    shapes = [int(x) for x in args.shapes.split(",")] if args.shapes else None
    tiles = [int(x) for x in args.tiles.split(",")] if args.tiles else None
    tuned = not (args.fixed_shapes or shapes or tiles)
    ctx = su.context()
    ctx.scf_init(spec.nelec, kT=spec.kT, seed=spec.seed, autotune=tuned)
    ctx.set_vloc(su.vext)
    if shapes:
        ctx.set_launch(shapes[0], shapes[1], bool(shapes[2] & 1), bool(shapes[2] & 2))
    if tiles:
        ctx.set_tiles(tiles[0], tiles[1], tiles[2] if len(tiles) > 2 else 64)
    tune64 = ctx.tuning()
    nb = spec.nbands
    log(f"# setup {time.perf_counter() - t0:.2f}s grid {su.dims} npw {su.npw[0]} bands {nb}")
    bands_per_s, ms_per_step, prof = hpsi_rate(ctx, nb, args.steps, args.warmup)
    log(f"# H|psi> FP64: {ms_per_step:.3f} ms per {nb} bands, {bands_per_s:.1f} bands/s, shapes {tune64}")
    ctx.close()
    # the same H|psi> in complex64 (the mixed form's bulk)
    ctx = su.context(lib.DEREK_MIXED)
    ctx.scf_init(spec.nelec, kT=spec.kT, seed=spec.seed, autotune=tuned)
    ctx.set_vloc(su.vext)
    if shapes:
        ctx.set_launch(shapes[0], shapes[1], bool(shapes[2] & 1), bool(shapes[2] & 2))
    if tiles:
        ctx.set_tiles(tiles[0], tiles[1], tiles[2] if len(tiles) > 2 else 64)
    tune32 = ctx.tuning()
    bps32, ms32, prof32 = hpsi_rate(ctx, nb, args.steps, args.warmup)
    log(f"# H|psi> FP32: {ms32:.3f} ms per {nb} bands, {bps32:.1f} bands/s, shapes {tune32}")
    ctx.close()

    # dominant kernel of the timed FP64 region
    busy = {k: v for k, v in prof.items() if v[0] > 0}
    dom = max(busy, key=lambda k: busy[k][1])
    n_l, tot_ms, by, fl = busy[dom]
    avg_s = tot_ms / n_l * 1e-3
    achieved = by / n_l / avg_s / 1e9
    total = sum(x[1] for x in busy.values())
    kernels = {k: {"launches": v[0], "avg_us": v[1] / v[0] * 1e3, "GB_s": v[2] / (v[1] * 1e-3) / 1e9,
                   "share": v[1] / total} for k, v in busy.items()}
    for k, v in busy.items():
        if v[3] > 0:
            kernels[k]["TFLOP_s"] = v[3] / (v[1] * 1e-3) / 1e12
    busy32 = {k: v for k, v in prof32.items() if v[0] > 0}
    kernels32 = {k: {"avg_us": v[1] / v[0] * 1e3, "GB_s": v[2] / (v[1] * 1e-3) / 1e9} for k, v in busy32.items()}
    # HBM bytes per launch from the committed counter summary: FETCH_SIZE and
    # WRITE_SIZE per (band) item at each bands-per-workgroup setting, times
    # the items of one launch
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_108cubed.json")
    if os.path.exists(tfile):
        per = json.load(open(tfile)).get("per_item", {}).get(str(tune64["bands_per_wg"]))
        if per:
            items = min(tune64["items_per_launch"], nb)
            traffic = round(items * (per["fetch"] + per["write"]))
    # The path's roofline (HBM): the compulsory bytes of one band's H|psi> at
    # one k, 2 npw s_psi (psi in, H psi out) + 12 npw (FFT index and kinetic
    # energy) + N_g s_r / B (V read once per launch of B bands), times the
    # bands of one launch of the dominant kernel, over its launch duration.
    npw0, s_psi, s_r = float(su.npw[0]), 16.0, 8.0
    units = min(tune64["items_per_launch"], nb)
    band_bytes = 2 * npw0 * s_psi + 12 * npw0 + su.N * s_r / units
    hbm = band_bytes * units / avg_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(hbm, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(hbm / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "unit_of_work": "one band's H|psi> at one k: 2 npw s_psi + 12 npw + N_g s_r / B",
                "bytes_per_unit": round(band_bytes), "units_per_launch": units, "avg_launch_us": avg_s * 1e6}
    # the same kernel against the pipe that bounds it: the FFT flops
    # (5 n log2 n per line) and the V multiply on the FP64 vector ALU, and
    # the bytes the kernel itself moves (stick rows in and out, V once)
    tf = fl / n_l / avg_s / 1e12 if fl > 0 else 0.0
    roofline_valu = {"bound": "valu", "achieved": round(tf, 2), "peak": FP64_PEAK_TFS, "unit": "TFLOP/s",
                     "frac": round(tf / FP64_PEAK_TFS, 4), "kernel": dom, "flops_per_launch": fl / n_l,
                     "kernel_bytes_per_launch": by / n_l, "kernel_GB_s": round(achieved, 1),
                     "kernel_hbm_frac": round(achieved / HBM_PEAK_GBS, 4)}

    scf_info = {}
    todo = [] if args.no_scf else [x.strip() for x in args.scf.split(",") if x.strip()]
    graph = not args.no_graph
    if "gamma32" in todo:
        spec32 = config.workload("gamma32")
        su32 = gscf.Setup(spec32)
        r32, _ = scf_run(gscf, lib, spec32, su32, lib.DEREK_FP64, kbatch=1)
        log(f"# gamma32 FP64: {r32['steps']} steps, {r32['time_to_scf_s']:.3f} s, E = {r32['etot_ha']:.10f}")
        if cpu_scf:
            c32 = cpu_scf["gamma32_time_to_scf"]
            r32 = dict(r32, cpu_time_to_scf_s=c32["value"], dE_vs_cpu_oracle_ha=abs(r32["etot_ha"] - c32["etot_ha"]),
                       speedup_vs_cpu_1core=round(c32["value"] / r32["time_to_scf_s"], 2))
        scf_info["gamma32_fp64"] = r32
    if "gamma108" in todo:
        r64, e64 = scf_run(gscf, lib, spec, su, lib.DEREK_FP64, kbatch=1)
        r64["step_roofline"] = scf_step_roofline(su, nb, 2, 16.0, r64["ms_per_step"] * 1e-3)
        log(f"# gamma108 FP64: {r64['steps']} steps, {r64['time_to_scf_s']:.3f} s, E = {r64['etot_ha']:.10f}")
        rmx, emx = scf_run(gscf, lib, spec, su, lib.DEREK_MIXED, graph=graph, refine_steps=args.refine,
                           gemm_bf16=args.bf16)
        log(f"# gamma108 mixed: {rmx['steps']} steps ({rmx['fp32_steps']} FP32), {rmx['time_to_scf_s']:.3f} s, "
            f"E = {rmx['etot_ha']:.10f}")
        scf_info["gamma108_fp64"] = r64
        scf_info["gamma108_mixed"] = dict(rmx, bf16_terms=args.bf16, dE_vs_fp64_ha=abs(rmx["etot_ha"] - r64["etot_ha"]),
                                      max_deig_vs_fp64_ha=float(np.max(np.abs(emx - e64))),
                                      speedup_vs_fp64=round(r64["time_to_scf_s"] / rmx["time_to_scf_s"], 3))
        if args.bf16:
            rbf, ebf = scf_run(gscf, lib, spec, su, lib.DEREK_MIXED, graph=graph, refine_steps=args.refine,
                               gemm_bf16=0)
            log(f"# gamma108 mixed, FP32 MFMA GEMMs: {rbf['steps']} steps ({rbf['fp32_steps']} FP32), "
                f"{rbf['time_to_scf_s']:.3f} s")
            scf_info["gamma108_mixed_fp32_mfma"] = dict(rbf, bf16_terms=0,
                                                    dE_vs_fp64_ha=abs(rbf["etot_ha"] - r64["etot_ha"]),
                                               max_deig_vs_fp64_ha=float(np.max(np.abs(ebf - e64))),
                                               speedup_vs_fp64=round(r64["time_to_scf_s"] / rbf["time_to_scf_s"], 3))
    if "kmesh72" in todo:
        spec72 = config.workload("kmesh72")
        su72 = gscf.Setup(spec72)
        log(f"# kmesh72 grid {su72.dims} nk {su72.nk} npw {min(su72.npw)}..{max(su72.npw)}")
        p64, e_p = scf_run(gscf, lib, spec72, su72, lib.DEREK_FP64, kbatch=1)
        log(f"# kmesh72 FP64 plain: {p64['steps']} steps, {p64['time_to_scf_s']:.3f} s, E = {p64['etot_ha']:.10f}")
        b64, e_b = scf_run(gscf, lib, spec72, su72, lib.DEREK_FP64, kbatch=0, graph=graph)
        log(f"# kmesh72 FP64 batched: {b64['steps']} steps, {b64['time_to_scf_s']:.3f} s")
        bmx, e_m = scf_run(gscf, lib, spec72, su72, lib.DEREK_MIXED, kbatch=0, graph=graph, autotune=True,
                           refine_steps=args.refine, gemm_bf16=args.bf16)
        log(f"# kmesh72 mixed batched: {bmx['steps']} steps ({bmx['fp32_steps']} FP32), {bmx['time_to_scf_s']:.3f} s")
        scf_info["kmesh72_fp64_plain"] = p64
        scf_info["kmesh72_fp64_batched"] = dict(b64, dE_vs_plain_ha=abs(b64["etot_ha"] - p64["etot_ha"]),
                                             speedup_vs_plain=round(p64["time_to_scf_s"] / b64["time_to_scf_s"], 3))
        scf_info["kmesh72_mixed_batched"] = dict(bmx, bf16_terms=args.bf16,
                                              dE_vs_plain_ha=abs(bmx["etot_ha"] - p64["etot_ha"]),
                                              max_deig_vs_plain_ha=float(np.max(np.abs(e_m - e_p))),
                                              speedup_vs_plain=round(p64["time_to_scf_s"] / bmx["time_to_scf_s"], 3))
        if args.bf16:
            bbf, e_bf = scf_run(gscf, lib, spec72, su72, lib.DEREK_MIXED, kbatch=0, graph=graph, autotune=True,
                                refine_steps=args.refine, gemm_bf16=0)
            log(f"# kmesh72 mixed batched, FP32 MFMA GEMMs: {bbf['steps']} steps ({bbf['fp32_steps']} FP32), "
                f"{bbf['time_to_scf_s']:.3f} s")
            scf_info["kmesh72_mixed_batched_fp32_mfma"] = dict(
                bbf, bf16_terms=0, dE_vs_plain_ha=abs(bbf["etot_ha"] - p64["etot_ha"]),
                max_deig_vs_plain_ha=float(np.max(np.abs(e_bf - e_p))),
                speedup_vs_plain=round(p64["time_to_scf_s"] / bbf["time_to_scf_s"], 3))

    if "kmesh144" in todo:
        spec144 = config.workload("kmesh144")
        su144 = gscf.Setup(spec144)
        log(f"# kmesh144 grid {su144.dims} nk {su144.nk} npw {min(su144.npw)}..{max(su144.npw)} bands {spec144.nbands}")
        f64, e_f = scf_run(gscf, lib, spec144, su144, lib.DEREK_FP64, verbose=True, kbatch=0, graph=graph)
        log(f"# kmesh144 FP64 batched: {f64['steps']} steps, {f64['time_to_scf_s']:.3f} s, E = {f64['etot_ha']:.10f}")
        mx144, e_x = scf_run(gscf, lib, spec144, su144, lib.DEREK_MIXED, verbose=True, kbatch=0, graph=graph,
                           refine_steps=args.refine, autotune=True, gemm_bf16=args.bf16)
        log(f"# kmesh144 mixed batched: {mx144['steps']} steps ({mx144['fp32_steps']} FP32), {mx144['time_to_scf_s']:.3f} s")
        scf_info["kmesh144_fp64_batched"] = f64
        scf_info["kmesh144_mixed_batched"] = dict(mx144, bf16_terms=args.bf16,
                                              dE_vs_fp64_ha=abs(mx144["etot_ha"] - f64["etot_ha"]),
                                              max_deig_vs_fp64_ha=float(np.max(np.abs(e_x - e_f))),
                                              speedup_vs_fp64=round(f64["time_to_scf_s"] / mx144["time_to_scf_s"], 3))
        if args.bf16:
            bx144, e_bx = scf_run(gscf, lib, spec144, su144, lib.DEREK_MIXED, verbose=True, kbatch=0, graph=graph,
                                refine_steps=args.refine, autotune=True, gemm_bf16=0)
            log(f"# kmesh144 mixed batched, FP32 MFMA GEMMs: {bx144['steps']} steps ({bx144['fp32_steps']} FP32), "
                f"{bx144['time_to_scf_s']:.3f} s")
            scf_info["kmesh144_mixed_batched_fp32_mfma"] = dict(
                bx144, bf16_terms=0, dE_vs_fp64_ha=abs(bx144["etot_ha"] - f64["etot_ha"]),
                max_deig_vs_fp64_ha=float(np.max(np.abs(e_bx - e_f))),
                speedup_vs_fp64=round(f64["time_to_scf_s"] / bx144["time_to_scf_s"], 3))

    sweep = None
    if args.items_sweep:
        sweep = {"system": args.items_sweep,
                 "fp64": items_sweep(gscf, lib, config, args.items_sweep, [64, 128, 256, 512, 0], lib.DEREK_FP64),
                 "complex64": items_sweep(gscf, lib, config, args.items_sweep, [64, 128, 256, 512, 0],
                                          lib.DEREK_MIXED)}

    heev_info = {}
    if args.heev:
        hc = lib.Context((8, 8, 8), 1, 4, 1.0, 1)  # a minimal context: the solver needs no basis
        rng = np.random.default_rng(0)
        for n in [int(x) for x in args.heev.split(",") if x.strip()]:
            a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
            A = 0.5 * (a + a.conj().T)
            t0 = time.perf_counter()
            for _ in range(3):
                w, _z = hc.heev(A, n // 2)
            heev_info[str(n)] = {"nev": n // 2, "ms_per_call_with_copies": round((time.perf_counter() - t0) / 3e-3, 3),
                                 "max_eig_err": float(np.abs(w - np.linalg.eigvalsh(A)[:n // 2]).max())}
            log(f"# heev n={n}: {heev_info[str(n)]}")
        hc.close()

    out = {
        "metric": METRIC, "value": round(bands_per_s, 2), "unit": "H|psi> bands/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded wells and wavefunctions)",
        "config": {"workload": "plain FP64 H|psi>: cubic L=20 bohr, Ecut=30 Ha, 108^3 grid, 62669 plane waves, "
                               "80 bands (Gamma), 16 wells + 0.2 Ha cos", "bands": nb, "npw": su.npw[0],
                   "grid": su.dims, "parallelism": "single GPU",
                   "launch_shapes": {"bands_per_wg": tune64["bands_per_wg"],
                                     "items_per_launch": tune64["items_per_launch"],
                                     "fused_yx": tune64["fused"], "fused_density": tune64["fused_density"],
                                     "z_sticks": tune64["z_sticks"], "gemm_rounds": tune64["gemm_rounds"],
                                     "gemm_tile": tune64["gemm_tile"]}},
        "roofline": roofline, "roofline_valu": roofline_valu, "kernels": kernels,
        "hpsi_fp32": {"value": round(bps32, 2), "unit": "bands/s", "ms_per_step": round(ms32, 4),
                      "launch_shapes": {"bands_per_wg": tune32["bands_per_wg"],
                                        "items_per_launch": tune32["items_per_launch"],
                                        "fused_yx": tune32["fused"], "fused_density": tune32["fused_density"],
                                        "z_sticks": tune32["z_sticks"], "gemm_rounds": tune32["gemm_rounds"],
                                        "gemm_tile": tune32["gemm_tile"]},
                      "kernels": kernels32},
        "scf": scf_info, "cpu_baseline": cpu, "cpu_baseline_all_cores": cpu_all, "cpu_baseline_scf": cpu_scf,
        "items_sweep": sweep, "heev": heev_info,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
