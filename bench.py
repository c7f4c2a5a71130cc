"""Benchmark of the ADER-DG element step on one MI355X.

Default workload `bn_ensemble_256`: 256 independent Baer-Nunziato two-phase
solves (9 variables, path-conservative fluxes, subcell limiter on) of a seeded
ensemble: mesh side log-uniform 64..384 on [0,1]^2, degree N in {2,3},
transmissive boundaries, stiffened-gas solid (gamma 4.4, pi 6e8) / ideal gas
(gamma 1.4), circular interface of radius U(0.1,0.3), pressure ratio
log-uniform 1e3..1e4, volume fraction 1e-6 / 1-1e-6; FP64.  The same K steps
of every solve run two ways:
  batched  all 256 solves in one batch (one arena per N, one step graph per
           step for all solves); `value`
  plain    one single-instance context per solve, one after the other
           (graph replay per step); `plain_value`

A step is one pass of the reference time loop body (Euler/main.f90:24-100,
Baer-Nunziato/main.f90): time step, predictor + volume integral, Riemann +
surface + update + DMP, limiter.  `value` = element-DOF updates per second of
the timed steps, a DOF update being one coefficient (v,i,j) of one cell
advanced one step (nVar*(N+1)^2 per cell-step).  Inputs are resident in HBM
before the timed region starts; uploads are not timed.

The line's `config.state` reports what the timed steps did to the solves: the
time step of the last timed step (min / median / max over the solves), the
solves holding non-finite values, the troubled-cell fraction.  The configured
interface (volume fractions 1e-6 / 1-1e-6) collapses under the reference scheme:
its time step falls by 10-20 orders of magnitude within ~30 steps.  The default
run therefore also times `bn_ensemble_256_stable` (`stable`): the same seeds,
meshes, degrees, equation of state, radii and pressure ratios with volume
fractions 0.3 / 0.7, which the reference scheme advances stably.

Workloads `euler_vortex_512_n4` and `euler_quadrant_1024_n3` time one plain
solve; `euler_quadrant_1024_n3_t0.3` runs the quadrant to its end time T=0.3.  `ad_sine_64_n3` runs the linear advection-diffusion problem (periodic
[0,1]^2, 64x64 cells, N=3, c=(1,1), diffusivity 0.01, u0 = sin 2 pi x sin 2 pi y,
CFL 0.9) to T=1 with the reference's viscous-only predictor flux and with the
fused advective + viscous one, and reports each run's L2 error against the exact
decaying solution.

Prints ONE JSON line.  Per-kernel details go to stderr.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ader-dg_amd"))

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix) dense peak
HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the dominant predictor (k_predictor_mfma at N=3, k_predictor_lean at N=4), from two rocprofv3 --pmc
# passes: read = 32/64/128 x TCC_EA0_RDREQ_{32B,64B,128B}_sum (FETCH_SIZE counts a 128-byte
# request as 64 bytes on this GPU), write = WRITE_SIZE KiB x 1024; both checked against the 1 GiB
# read and copy of --probe-only.  profiles/hbm_counter_calibration.md and profiles/hbm_traffic_*.md
# hold the runs.
PREDICTOR_TRAFFIC_BYTES = {"euler_vortex_512_n4": 755.3e6, "euler_quadrant_1024_n3": 2148.1e6,
                           "bn_ensemble_256_n3": 21.24e9}

PLAIN = {
    # name: (ic, N, imax, jmax, xl, xr, bc)
    "euler_vortex_512_n4": ("vortex", 4, 512, 512, (0.0, 0.0), (10.0, 10.0), "periodic"),
    "euler_quadrant_1024_n3": ("quadrant", 3, 1024, 1024, (0.0, 0.0), (1.0, 1.0), "transmissive"),
}
# plain workloads run to an end time: name -> (plain workload, end time)
PLAIN_TO_T = {"euler_quadrant_1024_n3_t0.3": ("euler_quadrant_1024_n3", 0.3)}
# name: (N, cells per side, advection velocity, diffusivity, end time)
AD = {"ad_sine_64_n3": (3, 64, (1.0, 1.0), 0.01, 1.0)}
ENSEMBLES = {"bn_ensemble_256": dict(count=256, seed=2026),
             # the same ensemble on small meshes (side 16..48), where one launch per solve cannot fill the GPU
             "bn_ensemble_256_small": dict(count=256, seed=2026, side=(16, 48)),
             # the same ensemble with volume fractions 0.3 / 0.7 (ics.BN_PHI_STABLE), which advances stably
             "bn_ensemble_256_stable": dict(count=256, seed=2026, phi_eps="stable")}


def ensemble_specs(name, count=None):
    from aderdg import ics
    cfg = dict(ENSEMBLES[name])
    if count:
        cfg["count"] = count
    if cfg.get("phi_eps") == "stable":
        cfg["phi_eps"] = ics.BN_PHI_STABLE
    return ics.bn_ensemble(**cfg)
SERVING = {"serving_4000": dict(count=4000, seed=7, steps=20)}
SERVING_MIX = ("35% Euler isentropic vortex, 35% Euler four-quadrant Riemann problem (end time 20 initial time "
               "steps), 30% Baer-Nunziato stiffened-gas interface (circle, volume fractions 0.3 / 0.7, end time 5 "
               "initial time steps)")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# CPU baselines: the reference Fortran routines built from its sources
# (oracle/_ref, flang -O3), or the oracle's C restatement where they are absent
# ---------------------------------------------------------------------------
def cpu_plain(wl, target_seconds):
    """One process on a 64x64 sample of a plain workload, steps until target_seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import aderdg
    from aderdg import ics
    import pyoracle
    ic, N, I, J, xl, xr, bc = PLAIN[wl]
    si = sj = 64
    uh = ics.initial_uh(ic, aderdg.operators(N)["xi"], si, sj, xl, xr)
    mode = {"periodic": pyoracle.BC_PERIODIC, "transmissive": pyoracle.BC_TRANSMISSIVE}[bc]
    if pyoracle.RefEuler.available(N):
        solver, kind = pyoracle.RefEuler(N, si, sj, xl, xr, mode), "reference"
    else:
        solver, kind = pyoracle.Oracle(N, si, sj, xl, xr, mode), "port"
    solver.set_uh(uh)
    solver.set_scalars(0.0, 0.0, 1e9)
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < target_seconds:
        solver.step()
        steps += 1
    el = time.perf_counter() - t0
    dof = steps * si * sj * 4 * (N + 1) ** 2
    return {"value": dof / el, "unit": "element-DOF updates/s", "cores": 1, "kind": kind,
            "sample": "%s IC, %dx%d cells, N=%d, %s, %d steps in %.1f s, one process" % (ic, si, sj, N, bc, steps, el)}


def cpu_ad(wl, target_seconds):
    """The reference advection-diffusion routines on the workload itself (64x64 is the
    workload's own mesh), both predictor builds in turn, steps until target_seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import aderdg
    import pyoracle
    N, side, c, a, T = AD[wl]
    uh = aderdg.ic_eval(aderdg.ic_sine(1.0, 1.0, 1.0), N, side, side, (0.0, 0.0), (1.0, 1.0), pde="ad")
    steps, el, kinds = 0, 0.0, set()
    for fixed in (False, True):
        if pyoracle.RefAD.available(N, fixed):
            s = pyoracle.RefAD(N, side, side, (0.0, 0.0), (1.0, 1.0), c, a, (3, 3, 3, 3), fixed=fixed)
            kinds.add("reference")
        else:
            s = pyoracle.Oracle(N, side, side, (0.0, 0.0), (1.0, 1.0), 1, pde=2,
                                ad=dict(c=c, a=a, bc=(3, 3, 3, 3), fixed=int(fixed)))
            kinds.add("port")
        s.set_uh(uh)
        s.set_scalars(0.0, 0.0, T)
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < target_seconds / 2:
            s.step()
            steps += 1
        el += time.perf_counter() - t0
    dof = steps * side * side * (N + 1) ** 2
    return {"value": dof / el, "unit": "element-DOF updates/s", "cores": 1,
            "kind": "reference" if kinds == {"reference"} else "port",
            "sample": "the workload's first %d steps (both predictor builds, half the time each) in %.1f s, one process"
                      % (steps, el)}


def _cpu_bn_worker(args):
    """One reference Baer-Nunziato solve of an ensemble member on its own mesh: steps until
    target_seconds have passed (at least one step)."""
    spec, target_seconds = args
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import aderdg
    from aderdg import ics
    import pyoracle
    N, I, J = spec["N"], spec["imax"], spec["jmax"]
    eos = ics.BN_EOS_ENSEMBLE
    uh = ics.initial_uh_bn(aderdg.operators(N)["xi"], I, J, (0.0, 0.0), (1.0, 1.0), eos,
                           radius=spec["radius"], pratio=spec["pratio"], phi_eps=spec["phi_eps"])
    if pyoracle.RefBN.available(N):
        s, kind = pyoracle.RefBN(N, I, J, (0.0, 0.0), (1.0, 1.0), eos), "reference"
    else:
        s, kind = pyoracle.Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), pyoracle.BC_TRANSMISSIVE, pde=1, eos=eos), "port"
    s.set_uh(uh)
    s.set_scalars(0.0, 0.0, 1e9)
    steps, t0 = 0, time.perf_counter()
    while steps == 0 or time.perf_counter() - t0 < target_seconds:
        s.step()
        steps += 1
    return N, I, steps * I * J * 9 * (N + 1) ** 2, time.perf_counter() - t0, kind


def host_cpu():
    """CPU model of this host, and the cores this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


# Reference processes of a CPU baseline: one per core, at most 16 -- the host cores a
# one-GPU job is given on the measurement hosts, whatever the machine has in total.
CPU_PROC_CAP = 16


def cpu_ensemble(specs, target_seconds, procs, seed=5):
    """One reference process per core (the reference is single-threaded), each on one
    ensemble member drawn at random, on that member's own mesh; runs before any GPU use.
    value = the processes' DOF/s summed (the host running `procs` members at once)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(specs), size=min(procs, len(specs)), replace=False)
    sample = [specs[int(k)] for k in pick]
    with mp.get_context("fork").Pool(len(sample)) as pool:
        res = pool.map(_cpu_bn_worker, [(s, target_seconds) for s in sample])
    value = sum(d / e for _, _, d, e, _ in res)
    kinds = {k for *_, k in res}
    model, avail = host_cpu()
    per_n = {}
    for N, _, d, e, _ in res:
        a = per_n.setdefault(str(N), [0.0, 0])
        a[0] += d / e
        a[1] += 1
    return {"value": value, "unit": "element-DOF updates/s", "cores": len(sample), "kind":
            "reference" if kinds == {"reference"} else "port",
            "cpu_model": model, "cores_available": avail, "cores_used": len(sample),
            "per_core_dof_per_s_by_N": {N: v / c for N, (v, c) in per_n.items()},
            "sample": "%d processes, one per core (at most %d: the host cores of a one-GPU job), each one "
                      "randomly drawn ensemble member on its own mesh (sides %s, N %s) for >= %.0f s and at "
                      "least one step" % (len(sample), CPU_PROC_CAP, "/".join(str(s["imax"]) for s in sample),
                                          "/".join(str(s["N"]) for s in sample), target_seconds)}


# ---------------------------------------------------------------------------
def pinned_variant(spec, N):
    """The variant ADERDG_PREDICTOR_VARIANT pins for degree N ("<v>" or "<N>:<v>,..."), or None."""
    spec = spec.strip()
    if not spec:
        return None
    if ":" not in spec:
        return int(spec) if spec.isdigit() else None
    for item in spec.split(","):
        n, _, v = item.partition(":")
        if n.strip() == str(N) and v.strip().isdigit():
            return int(v)
    return None


def predictor_name(pde, N, persistent):
    """Kernel name of the predictor variant in use: the one ADERDG_PREDICTOR_VARIANT pins
    (the library then skips the autotuner), else the autotuner's choice."""
    import aderdg
    names = {0: "k_predictor", 1: "k_predictor_nodes<.., 64>", 2: "k_predictor_nodes<.., 128>",
             3: "k_predictor_nodes<.., 256>", 4: "k_predictor_lean", 5: "k_predictor_mfma"}
    v = pinned_variant(os.environ.get("ADERDG_PREDICTOR_VARIANT", ""), N)
    if v is None or not aderdg.pred_variant_valid(pde, N, v, exact=False):
        t = aderdg.tune_query(pde, N, persistent)
        v = t[0] if t else 0
    return "%s<%sPDE,%d>" % (names[v], "Bn" if pde == "bn" else "Euler", N)


def first_sweep_skipped(pde, N):
    """Operations per cell-step of the reference algorithm that the time-line predictor
    kernels (k_predictor, k_predictor_lean, k_predictor_mfma) do not issue: the first Picard sweep starts from
    the trivial initial guess (every time slice holds luh), so slices 1..N repeat slice 0's
    flux, flux contractions and (Baer-Nunziato) gradients and NCP source, which the kernels
    form once.  Same convention and constants as aderdg_flops_predictor."""
    bn = pde == "bn"
    n, V, cF = N + 1, (9 if bn else 4), (76 if bn else 26)
    per_slice = n * n * cF + 2 * V * n * n * 2 * n
    if bn:
        per_slice += 2 * V * n * n * (2 * n + 1) + n * n * 34
    return (n - 1) * per_slice


PEAKS = None  # measured (FP64 TFLOP/s, HBM read GB/s, HBM copy GB/s) of this run's GPU


def roofline_of(flop_per_cell, cells, ms_per_launch, kernel, traffic=None, skipped_per_cell=0.0):
    """achieved: the reference algorithm's operations (the oracle's op count) per second;
    executed_achieved: the operations the kernel issues per second."""
    achieved = flop_per_cell * cells / (ms_per_launch * 1e-3) / 1e12
    executed = flop_per_cell - skipped_per_cell
    extra = {}
    if PEAKS:
        extra = {"peak_measured": PEAKS[0], "frac_measured": achieved / PEAKS[0],
                 "peak_measured_note": "FP64 FMA rate of eight independent chains per thread on this GPU "
                                       "(aderdg_probe_peaks)"}
    # "mfma" is the compute roofline of the line's contract; its FP64 peak equals the vector ALU's
    # on MI355X.  The matrix-core predictor runs the slice contractions on the matrix cores and the
    # pointwise flux / NCP work on the vector ALU (compute_unit)
    mfma = "mfma" in kernel
    if mfma:
        unit_note = ("FP64 vector ALU for the pointwise flux and NCP terms, FP64 matrix cores "
                     "(v_mfma_f64_4x4x4_4b_f64) for the per-slice 4x4 contractions")
    else:
        unit_note = "FP64 vector ALU (no MFMA: the per-cell contractions are 4x4..5x5 and run on the VALU)"
    return {"bound": "mfma", "compute_unit": "fp64-valu+mfma" if mfma else "fp64-valu", "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", **extra,
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic, "traffic_unit": "bytes/launch",
            "kernel": kernel, "flop_per_cell": flop_per_cell, "cells_per_launch": cells,
            "executed_flop_per_cell": executed,
            "executed_achieved": executed * cells / (ms_per_launch * 1e-3) / 1e12,
            "note": unit_note + "; "
                    "MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s. achieved and frac count the "
                    "reference algorithm's operations; the time-line predictors form the first Picard sweep's "
                    "flux terms once for all time slices, so they issue executed_flop_per_cell"}


def skipped_for(kernel, pde, N):
    return first_sweep_skipped(pde, N) if kernel.startswith(("k_predictor<", "k_predictor_lean<",
                                                             "k_predictor_mfma<")) else 0.0


def algorithmic_bytes(pde, N):
    """HBM bytes per cell-step each kernel must move at least (FP64, ping-pong state):
    ranges reads the cell's first-pass ranges (its neighbours' come from cache) and writes
    the 3x3 ranges; predictor reads uh, writes lduh and the face extrapolations fq/ff; face
    reads fq/ff of both neighbours and writes the fluxes (one x and one y face per cell;
    Baer-Nunziato two fluctuation sides); update reads uh, lduh, four face lines and the
    ranges, writes uh', the next ranges and the time-step candidate."""
    V = 9 if pde == "bn" else 4
    n = N + 1
    sides = 2 if pde == "bn" else 1
    return {"ranges": 8 * (2 * V) + 8 * (2 * V),
            "predictor": 8 * (V * n * n) + 8 * (V * n * n + 2 * 4 * V * n),
            "face": 2 * 8 * (2 * 2 * V * n) + 2 * 8 * (sides * V * n),
            "update": 8 * (2 * V * n * n + 4 * V * n + 2 * V) + 8 * (V * n * n + 2 * V + 1)}


def kernel_table(pde, per_group, peaks):
    """Per arena and kernel: ms per launch, algorithmic GB/s and its fraction of HBM peaks."""
    out = {}
    for N, g in per_group.items():
        ab = algorithmic_bytes(pde, int(N))
        row = {}
        for k in ("predictor", "face", "update", "subcell", "dt", "spread", "ranges"):
            ms = g.get(k)
            if not ms:
                continue
            r = {"ms": ms}
            if k in ab:
                gbs = ab[k] * g["cells"] / (ms * 1e-3) / 1e9
                r.update(algorithmic_bytes_per_cell=ab[k], algorithmic_gbs=gbs, frac_hbm_peak=gbs / HBM_PEAK_GBS)
                if peaks:
                    r["frac_measured_read_peak"] = gbs / peaks[1]
            row[k] = r
        out[str(N)] = row
    return out


def run_plain_to_t(args):
    """A plain workload to its end time (the whole run timed, graph replays); a few
    steps of the final state launched kernel by kernel give the per-kernel times."""
    import aderdg
    from aderdg import ics
    base, tend = PLAIN_TO_T[args.workload]
    ic, N, I, J, xl, xr, bc = PLAIN[base]
    uh0 = ics.initial_uh(ic, aderdg.operators(N)["xi"], I, J, xl, xr)
    s = aderdg.Solver(N, I, J, xl, xr, bc)
    s.upload(uh0)
    if args.warmup:
        s.run(args.warmup, 0.0, tend)
        s.upload(uh0)
    t, steps, dt, troubled, last_log = 0.0, 0, 0.0, 0, 0
    t0 = time.perf_counter()
    while t < tend:
        # graph replays in chunks that stop short of tend (dt changes slowly), then single
        # steps, so that no step runs after the solve has reached its end time
        k = 1 if dt <= 0.0 else max(1, min(500, int(0.9 * (tend - t) / dt)))
        t, dt, troubled = s.run(k, t, tend)
        steps += k
        if steps - last_log >= 1000:
            last_log = steps
            log("t = %.6g after %d steps (%.1f s), nRecompute %d" % (t, steps, time.perf_counter() - t0, troubled))
    el = time.perf_counter() - t0
    log("t = %.6g after %d steps in %.2f s" % (t, steps, el))
    uh = s.download()
    s.upload(uh)
    s.timing(True)
    s.run(5, 0.0, 1e30)
    ms, launches = s.timing_get()
    s.close()
    ncell = I * J
    per = {name: (ms[k] / launches[k] if launches[k] else None) for k, name in enumerate(aderdg.TIMING_SLOTS)}
    log("per-kernel average ms (final state):", json.dumps(per))
    fl_pred, _ = aderdg.flops_per_cell(N)
    return {
        "metric": "element-DOF updates/s (plain single-instance Euler step, to the end time)",
        "value": steps * ncell * 4 * (N + 1) ** 2 / el, "unit": "element-DOF updates/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": el / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (analytic initial condition)",
        "config": {"workload": args.workload, "pde": "euler", "ic": ic, "N": N, "cells": [I, J], "bc": bc,
                   "tend": tend, "seconds_to_tend": el, "last_nRecompute": troubled},
        "roofline": roofline_of(fl_pred, ncell, per["predictor"], predictor_name("euler", N, False),
                                skipped_per_cell=skipped_for(predictor_name("euler", N, False), "euler", N))
        if per["predictor"] else None,
    }


def run_plain(args):
    import aderdg
    from aderdg import ics
    ic, N, I, J, xl, xr, bc = PLAIN[args.workload]
    ops = aderdg.operators(N)
    uh0 = ics.initial_uh(ic, ops["xi"], I, J, xl, xr)
    s = aderdg.Solver(N, I, J, xl, xr, bc)
    s.upload(uh0)
    t, dt, nrec = s.run(args.warmup, 0.0, 1e30) if args.warmup else (0.0, 0.0, 0)
    s.timing(not args.graph)
    t0 = time.perf_counter()
    t1_sim, dt, nrec = s.run(args.steps, t, 1e30)
    el = time.perf_counter() - t0
    ms, launches = s.timing_get()
    s.close()
    ncell = I * J
    value = args.steps * ncell * 4 * (N + 1) ** 2 / el
    per = {name: (ms[k] / launches[k] if launches[k] else None) for k, name in enumerate(aderdg.TIMING_SLOTS)}
    log("per-kernel average ms:", json.dumps(per))
    log("simulated time %.6g after %d steps, last dt %.6g, nRecompute %d" % (t1_sim, args.steps, dt, nrec))
    fl_pred, _ = aderdg.flops_per_cell(N)
    kname = predictor_name("euler", N, False)
    roof = roofline_of(fl_pred, ncell, per["predictor"], kname, PREDICTOR_TRAFFIC_BYTES.get(args.workload),
                       skipped_for(kname, "euler", N)) if per["predictor"] else None
    return {
        "metric": "element-DOF updates/s (plain single-instance Euler step)",
        "value": value, "unit": "element-DOF updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (analytic initial condition)",
        "config": {"workload": args.workload, "pde": "euler", "ic": ic, "N": N, "cells": [I, J], "bc": bc,
                   "cell_updates_per_s": args.steps * ncell / el},
        "roofline": roof,
    }


def run_ad(args):
    """Both predictor fluxes to the end time; `value` over the two runs together."""
    import math
    import numpy as np
    import aderdg
    N, side, c, a, T = AD[args.workload]
    ops = aderdg.operators(N)
    xi, w = ops["xi"], ops["w"]
    uh0 = aderdg.ic_eval(aderdg.ic_sine(1.0, 1.0, 1.0), N, side, side, (0.0, 0.0), (1.0, 1.0), pde="ad")
    ii = np.arange(side)
    X = (ii[None, None, :, None] + xi[:, None, None, None]) / side
    Y = (ii[None, None, None, :] + xi[None, :, None, None]) / side
    avg = lambda q: np.einsum("k,l,klij->ij", w, w, q)
    runs, total_steps, total_s = {}, 0, 0.0
    for name, fixed in (("reference_predictor", 0), ("fused", 1)):
        s = aderdg.Solver(N, side, side, (0.0, 0.0), (1.0, 1.0), "periodic", pde="ad",
                          ad=dict(c=c, a=a, bc=(3, 3, 3, 3), fixed=fixed))
        s.upload(uh0)
        s.run(args.warmup, 0.0, T)  # warm-up on a copy of the start state, then restart
        s.upload(uh0)
        t0 = time.perf_counter()
        t, dt, _ = s.run(1, 0.0, T)
        steps = 1
        while t < T:  # dt is the same every step until the last one is clipped to T
            k = max(1, int((T - t) / dt) - 1)
            t, dt, _ = s.run(k, t, T)
            steps += k
        el = time.perf_counter() - t0
        uh = s.download()
        s.timing(True)
        s.run(20, 0.0, 1e30)
        ms, launches = s.timing_get()
        s.close()
        ex = math.exp(-8 * math.pi ** 2 * a * t) * np.sin(2 * math.pi * (X - c[0] * t)) * np.sin(2 * math.pi * (Y - c[1] * t))
        per = {nm: (ms[k] / launches[k] if launches[k] else None) for k, nm in enumerate(aderdg.TIMING_SLOTS)}
        runs[name] = {"steps": steps, "time": t, "seconds": el, "ms_per_step": el / steps * 1e3,
                      "l2_nodal": float(np.sqrt(np.mean((uh[0] - ex) ** 2))),
                      "l2_cell_average": float(np.sqrt(np.mean((avg(uh[0]) - avg(ex)) ** 2))),
                      "kernel_ms": per}
        log("%s: %d steps to t=%.6g in %.3f s; L2 %.4g; kernels %s"
            % (name, steps, t, el, runs[name]["l2_cell_average"], json.dumps(per)))
        total_steps += steps
        total_s += el
    ncell = side * side
    fl, _ = aderdg.flops_per_cell(N, "ad", fused=True)
    pm = runs["fused"]["kernel_ms"]["predictor"]
    return {
        "metric": "element-DOF updates/s (plain single-instance advection-diffusion step)",
        "value": total_steps * ncell * (N + 1) ** 2 / total_s, "unit": "element-DOF updates/s", "n_gpus": 1,
        "steps": total_steps, "warmup": args.warmup, "ms_per_step": total_s / total_steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (analytic initial condition)",
        "config": {"workload": args.workload, "pde": "advection-diffusion", "N": N, "cells": [side, side],
                   "c": list(c), "diffusivity": a, "tend": T, "bc": "periodic", "runs": runs},
        "roofline": roofline_of(fl, ncell, pm, "k_predictor_ad<%d> (fused flux)" % N) if pm else None,
    }


def ensemble_state(b, ids, specs, steps_done):
    """What the steps so far did to the solves: every solve took steps_done steps and none
    failed; the time step of its last step; solves with non-finite DOFs; troubled cells."""
    import numpy as np
    dts, times, bad_solves, bad_dofs, troubled = [], [], 0, 0, 0
    for i in ids:
        q = b.query(i)
        assert q["steps"] == steps_done and b.error(i) == 0, (i, q, steps_done)
        nf, tr = b.health(i)
        dts.append(q["dt"])
        times.append(q["time"])
        bad_solves += 1 if (nf or not np.isfinite(q["dt"])) else 0
        bad_dofs += nf
        troubled += tr
    cells = sum(s["imax"] * s["jmax"] for s in specs)
    d = np.array(dts)
    fin = d[np.isfinite(d)]
    return {"steps": steps_done, "last_dt_min": float(fin.min()) if fin.size else None,
            "last_dt_median": float(np.median(fin)) if fin.size else None,
            "last_dt_max": float(fin.max()) if fin.size else None,
            "time_median": float(np.median(times)), "solves_nonfinite": bad_solves, "dofs_nonfinite": int(bad_dofs),
            "troubled_fraction": troubled / cells}


def run_ensemble(args, specs):
    import aderdg
    from aderdg import ics
    eos = ics.BN_EOS_ENSEMBLE
    xis = {N: aderdg.operators(N)["xi"] for N in {s["N"] for s in specs}}

    def uh_of(sp):
        return ics.initial_uh_bn(xis[sp["N"]], sp["imax"], sp["jmax"], (0.0, 0.0), (1.0, 1.0), eos,
                                 radius=sp["radius"], pratio=sp["pratio"], phi_eps=sp["phi_eps"])

    K, W, T = args.steps, args.warmup, 2
    dof_step = sum(9 * (s["N"] + 1) ** 2 * s["imax"] * s["jmax"] for s in specs)
    cells = sum(s["imax"] * s["jmax"] for s in specs)

    # plain: one context per solve, one after the other; only the K step replays are timed
    plain_s = None
    if not args.no_plain:
        plain_s = 0.0
        for k, sp in enumerate(specs):
            s = aderdg.Solver(sp["N"], sp["imax"], sp["jmax"], (0.0, 0.0), (1.0, 1.0), "transmissive", pde="bn",
                              eos=eos)
            s.upload(uh_of(sp))
            t, _, _ = s.run(W, 0.0, 1e30) if W else (0.0, 0.0, 0)
            t0 = time.perf_counter()
            s.run(K, t, 1e30)
            plain_s += time.perf_counter() - t0
            s.close()
            if k % 64 == 63:
                log("plain: %d solves done, %.2f s timed" % (k + 1, plain_s))

    # batched: every solve in one batch; W warm-up, K timed graph replays, T evented steps
    b = aderdg.Batch()
    for N in sorted({s["N"] for s in specs}):
        mine = [s for s in specs if s["N"] == N]
        b.reserve("bn", N, max_cells=sum(s["imax"] * s["jmax"] for s in mine), max_solves=len(mine))
    log("batch arenas: %.1f GB reserved" % (b.reserved_bytes() / 1e9))
    ids = [b.add(uh_of(sp), sp["N"], sp["imax"], sp["jmax"], (0.0, 0.0), (1.0, 1.0), bc="transmissive", pde="bn",
                 eos=eos, max_steps=W + K + T) for sp in specs]
    b.sync()
    if W:
        b.step(W)
    b.sync()
    t0 = time.perf_counter()
    b.step(K)
    b.sync()
    batched_s = time.perf_counter() - t0
    state = ensemble_state(b, ids, specs, W + K)
    log("ensemble state after the timed steps:", json.dumps(state))
    b.timing(True)
    b.step(T)
    b.sync()
    per_group = {}
    for N in sorted({s["N"] for s in specs}):
        ms, ln = b.timing_get("bn", N)
        gcells = sum(s["imax"] * s["jmax"] for s in specs if s["N"] == N)
        per_group[N] = dict(cells=gcells, **{name: (ms[k] / ln[k] if ln[k] else None)
                                              for k, name in enumerate(aderdg.TIMING_SLOTS)})
    done = b.finished()
    for i in ids:
        q = b.query(i)
        assert q["steps"] == W + K + T and not q["active"] and b.error(i) == 0, (i, q)
    b.close()
    log("batched per-kernel average ms per arena:", json.dumps(per_group))
    assert len(done) == len(ids), "every solve should have stopped after W+K+T steps"

    value = dof_step * K / batched_s
    dom = max(per_group, key=lambda N: per_group[N]["predictor"] or 0.0)
    fl_pred, _ = aderdg.flops_per_cell(dom, "bn")
    tuned = {N: aderdg.tune_query("bn", N, True) for N in per_group}
    kname = predictor_name("bn", dom, True)
    roof = roofline_of(fl_pred, per_group[dom]["cells"], per_group[dom]["predictor"], kname,
                       PREDICTOR_TRAFFIC_BYTES.get("%s_n%d" % (args.workload, dom)), skipped_for(kname, "bn", dom))
    if len(per_group) > 1 and kname.startswith("k_predictor_mfma"):
        roof["note"] += ("; timed launched alone at full residency (two workgroups per CU); in the step graph, "
                         "beside the other arena's branch, it runs one workgroup per CU")
    out = {
        "metric": "element-DOF updates/s batched vs plain; p50/p99 solve latency at fixed QPS",
        "value": value, "unit": "element-DOF updates/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": batched_s / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded analytic initial conditions)",
        "config": {"workload": args.workload, "pde": "baer-nunziato", "solves": len(specs), "cells": cells,
                   "N": sorted({s["N"] for s in specs}),
                   "side": list(ENSEMBLES[args.workload].get("side", (64, 384))), "bc": "transmissive",
                   "seed": ENSEMBLES[args.workload]["seed"], "dof_per_step": dof_step,
                   "volume_fractions": [specs[0]["phi_eps"], 1.0 - specs[0]["phi_eps"]], "state": state},
        "plain_value": dof_step * K / plain_s if plain_s else None,
        "batched_vs_plain": (plain_s / batched_s) if plain_s else None,
        "roofline": roof,
        "kernels": kernel_table("bn", per_group, PEAKS),
        "predictor_autotune": {str(N): {"variant": aderdg.PREDICTOR_VARIANTS[t[0]],
                                        "ms_per_launch_256x256": dict(zip(aderdg.PREDICTOR_VARIANTS, t[1]))}
                               for N, t in tuned.items() if t},
        "update_autotune": {str(N): {"variant": aderdg.UPDATE_VARIANTS[t[0]],
                                     "ms_per_launch_256x256": dict(zip(aderdg.UPDATE_VARIANTS, t[1]))}
                            for N, t in ((N, aderdg.tune_query("bn", N, True, kernel="update")) for N in per_group)
                            if t},
    }
    return out


# ---------------------------------------------------------------------------
# serving: Poisson arrivals into the continuous-batching server
# ---------------------------------------------------------------------------
SERVING_CELLS = {"euler": 2_000_000, "bn": 1_500_000}  # arena cells per (pde, N)
# Step cap of a request, in planned steps: a safety net only.  A request's end time is its planned
# steps times the time step estimated from its initial state; the Baer-Nunziato interfaces
# accelerate their gas phase, and at N = 4 and on the finest meshes the time step falls 10-20x
# below that estimate, so such requests take that many more steps to reach their end time.
STEP_CAP = 50


def serving_setup(specs):
    import aderdg
    srv = aderdg.Server(steps_per_round=2)
    for pde, N in sorted({(s["pde"], s["N"]) for s in specs}):
        srv.reserve(pde, N, max_cells=SERVING_CELLS[pde], max_solves=512)
    srv.start()
    # a step cap beside each end time: a solve whose time step collapses stops at STEP_CAP x the planned steps
    reqs = [srv.request(aderdg.ic_spec(s["kind"], *s["params"]), s["N"], s["side"], s["side"], s["xl"], s["xr"],
                        s["tend"], bc=s["bc"], pde=s["pde"], eos=s["eos"], max_steps=STEP_CAP * s["steps"])
            for s in specs]
    return srv, reqs


def serve(srv, reqs, rate, seed):
    """Submit reqs at Poisson arrival times of `rate` per second (all at once when rate
    is None) from a producer thread; returns latencies (s), wall seconds, DOF updates."""
    import threading
    import numpy as np
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(reqs))) if rate else np.zeros(len(reqs))
    dof0 = srv.stats()["dof_updates"]
    t0 = time.perf_counter()

    def producer():
        for k, r in enumerate(reqs):
            delay = arrivals[k] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            srv.submit_request(r)

    th = threading.Thread(target=producer)
    th.start()
    lat, capped = [], 0
    while len(lat) < len(reqs):
        for t, l, st, steps, reached in srv.poll(details=True):
            if st != 0:
                raise RuntimeError("request %d failed with %d" % (t, st))
            lat.append(l)
            capped += 0 if reached else 1
            if len(lat) % 500 == 0:
                log("  %d / %d requests done after %.1f s" % (len(lat), len(reqs), time.perf_counter() - t0))
        time.sleep(0.001)
    wall = time.perf_counter() - t0
    th.join()
    if capped:
        log("  %d requests stopped at the step cap before their end time" % capped)
    return np.array(lat), wall, srv.stats()["dof_updates"] - dof0, capped


def run_serving(specs, factors, seed=11):
    import numpy as np
    srv, reqs = serving_setup(specs)
    lat, wall, dof, capped = serve(srv, reqs, None, seed)
    sat = len(reqs) / wall
    log("serving saturation: %d requests in %.2f s = %.1f req/s, %.3g DOF/s" % (len(reqs), wall, sat, dof / wall))
    out = {"requests": len(reqs), "saturation_req_per_s": sat, "saturation_dof_per_s": dof / wall,
           "requests_at_step_cap": capped, "step_cap": "%d x planned steps" % STEP_CAP, "rates": []}
    for f in factors:
        lat, wall, dof, capped = serve(srv, reqs, f * sat, seed + int(100 * f))
        r = {"rate_factor": f, "offered_req_per_s": f * sat, "achieved_req_per_s": len(reqs) / wall,
             "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
             "mean_s": float(lat.mean()), "max_s": float(lat.max()), "dof_per_s": dof / wall}
        log("serving at %.1fx saturation: %s" % (f, json.dumps(r)))
        out["rates"].append(r)
    st = srv.stats()
    srv.close()
    out["peak_admitted_bytes"] = st["peak_bytes"]
    return out


def probe():
    """Measure this GPU's FP64 FMA and HBM stream peaks once (aderdg_probe_peaks)."""
    global PEAKS
    import aderdg
    try:
        PEAKS = aderdg.probe_peaks()
        log("measured peaks: FP64 %.1f TFLOP/s, HBM read %.0f GB/s, copy %.0f GB/s" % PEAKS)
    except aderdg.AderdgError as e:
        log("peak probe failed: %s" % e)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 20 ensemble, 200 plain)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps first (default 2 ensemble, 10 plain)")
    ap.add_argument("--workload", default="bn_ensemble_256", choices=sorted(PLAIN) + sorted(PLAIN_TO_T) + sorted(AD) + sorted(ENSEMBLES) + sorted(SERVING))
    ap.add_argument("--serving-requests", type=int, default=1000,
                    help="ensemble workload: requests of the serving stream timed beside it at 0.5x, 0.8x and 1.0x "
                         "of saturation (0: none)")
    ap.add_argument("--count", type=int, default=None, help="ensemble size (default 256)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-procs", type=int, default=None,
                    help="reference processes (default: one per available core, at most %d)" % CPU_PROC_CAP)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="ensemble: skip the plain (one-by-one) leg")
    ap.add_argument("--no-small", action="store_true",
                    help="ensemble: skip the batched-vs-plain run of the same seeds on small meshes")
    ap.add_argument("--no-stable", action="store_true",
                    help="ensemble: skip the run of the same seeds with volume fractions 0.3 / 0.7")
    ap.add_argument("--graph", action="store_true", help="plain: time graph replay instead of evented launches")
    ap.add_argument("--probe-only", action="store_true",
                    help="only measure the GPU's FP64 and HBM peaks (aderdg_probe_peaks: a 1 GiB stream read and "
                         "a 1 GiB copy of known size, the calibration of the HBM counters) and print them")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported "
              "(independent solves shard across GPUs without communication)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.probe_only:
        probe()
        print(json.dumps({"measured_peaks": PEAKS}), flush=True)
        return
    ens = args.workload in ENSEMBLES
    if args.steps is None:
        args.steps = 20 if ens else 200
    if args.warmup is None:
        args.warmup = 2 if ens else 10

    cpu = None
    if args.workload in SERVING:
        from aderdg import ics
        cfg = dict(SERVING[args.workload])
        if args.count:
            cfg["count"] = args.count
        specs = ics.serving_requests(**cfg)
        sv = run_serving(specs, (0.5, 0.8, 1.0))
        r08 = [r for r in sv["rates"] if r["rate_factor"] == 0.8][0]
        out = {"metric": "p50/p99 solve latency at fixed QPS (serving)", "value": r08["p99_s"], "unit": "s",
               "n_gpus": 1, "steps": cfg["steps"], "warmup": 0, "ms_per_step": None, "higher_is_better": False,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64",
               "data": "synthetic (seeded formula initial conditions)",
               "config": {"workload": args.workload, "requests": cfg["count"], "seed": cfg["seed"],
                          "mix": SERVING_MIX,
                          "side": [32, 512], "N": [2, 3, 4], "steps_per_request": cfg["steps"]},
               "serving": sv}
        print(json.dumps(out), flush=True)
        return
    if ens:
        from aderdg import ics
        specs = ensemble_specs(args.workload, args.count)
        if not args.no_cpu:  # before any GPU use: the workers are forked
            procs = args.cpu_procs or max(1, min(CPU_PROC_CAP, host_cpu()[1]))
            cpu = cpu_ensemble(specs, args.cpu_seconds, procs)
            log("cpu baseline:", json.dumps(cpu))
        probe()
        out = run_ensemble(args, specs)
        if args.workload == "bn_ensemble_256" and not args.no_stable:
            # the same seeds on a state the reference scheme advances stably, batched and plain
            sargs = argparse.Namespace(**dict(vars(args), workload="bn_ensemble_256_stable"))
            st = run_ensemble(sargs, ensemble_specs("bn_ensemble_256_stable", args.count))
            out["stable"] = {k: st[k] for k in ("value", "plain_value", "batched_vs_plain", "ms_per_step", "roofline")}
            out["stable"]["config"] = st["config"]
            log("stable interface:", json.dumps(out["stable"]))
        if args.workload == "bn_ensemble_256" and not args.no_small and not args.no_plain:
            # batched vs plain where batching matters: the same seeds on meshes of side 16..48
            sargs = argparse.Namespace(**dict(vars(args), workload="bn_ensemble_256_small"))
            small = run_ensemble(sargs, ensemble_specs("bn_ensemble_256_small"))
            out["small_meshes"] = {k: small[k] for k in ("value", "plain_value", "batched_vs_plain", "ms_per_step")}
            out["small_meshes"]["config"] = small["config"]
            log("small meshes:", json.dumps(out["small_meshes"]))
        if args.serving_requests:
            cfg = dict(SERVING["serving_4000"], count=args.serving_requests)
            out["serving"] = run_serving(ics.serving_requests(**cfg), (0.5, 0.8, 1.0))
            out["serving"]["config"] = {"requests": cfg["count"], "seed": cfg["seed"], "mix": SERVING_MIX,
                                        "side": [32, 512], "N": [2, 3, 4], "steps_per_request": cfg["steps"]}
    elif args.workload in PLAIN_TO_T:
        probe()
        out = run_plain_to_t(args)
        if not args.no_cpu:
            cpu = cpu_plain(PLAIN_TO_T[args.workload][0], args.cpu_seconds)
    elif args.workload in AD:
        out = run_ad(args)
        if not args.no_cpu:
            cpu = cpu_ad(args.workload, args.cpu_seconds)
    else:
        probe()
        out = run_plain(args)
        if not args.no_cpu:
            cpu = cpu_plain(args.workload, args.cpu_seconds)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    if PEAKS:
        out["measured_peaks"] = {"fp64_tflops": PEAKS[0], "hbm_read_gbs": PEAKS[1], "hbm_copy_gbs": PEAKS[2]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
