"""Benchmark of the ADER-DG element step on one MI355X.

Workload (default): 2-D Euler isentropic vortex, periodic [0,10]^2, 512x512
cells, degree N=4, gamma=1.4, vortex strength 5 at (5,5) on (rho,u,v,p) =
(1,1,1,1), FP64, limiter enabled (untriggered), 200 steps.

A step is one pass of the reference time loop body (Euler/main.f90:24-100):
time step, predictor + volume integral, Riemann + surface + update + DMP,
limiter post-processing.  `value` = element-DOF updates per second over the
timed steps, a DOF update being one coefficient (v,i,j) of one cell advanced
one step (nVar*(N+1)^2 per cell-step).  Inputs are resident in HBM before
the timed region starts.

Prints ONE JSON line.  Per-kernel details go to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ader-dg_amd"))

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix) dense peak
HBM_PEAK_GBS = 8000.0

# HBM bytes per k_predictor launch, from rocprofv3 --pmc FETCH_SIZE (x2.0) and WRITE_SIZE (x1.0),
# scaled by a 256 MiB device copy; profiles/hbm_traffic_euler_vortex_512_n4.md holds the run.
PREDICTOR_TRAFFIC_BYTES = {"euler_vortex_512_n4": 965.7e6}

WORKLOADS = {
    # name: (ic, N, imax, jmax, xl, xr, bc)
    "euler_vortex_512_n4": ("vortex", 4, 512, 512, (0.0, 0.0), (10.0, 10.0), "periodic"),
    "euler_quadrant_1024_n3": ("quadrant", 3, 1024, 1024, (0.0, 0.0), (1.0, 1.0), "transmissive"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, target_seconds):
    """Reference Fortran routines (oracle/_ref, flang -O3, one core) on a bounded sample of the
    same workload: the same IC and boundaries on a 64x64 mesh, steps until target_seconds."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import aderdg
    from aderdg import ics
    import pyoracle
    ic, N, I, J, xl, xr, bc = WORKLOADS[wl]
    si = sj = 64
    xi = aderdg.operators(N)["xi"]
    uh = ics.initial_uh(ic, xi, si, sj, xl, xr)
    mode = {"periodic": pyoracle.BC_PERIODIC, "transmissive": pyoracle.BC_TRANSMISSIVE}[bc]
    if pyoracle.RefEuler.available(N):
        solver, kind = pyoracle.RefEuler(N, si, sj, xl, xr, mode), "reference"
        solver.set_uh(uh)
        solver.set_scalars(0.0, 0.0, 1e9)
    else:
        solver, kind = pyoracle.Oracle(N, si, sj, xl, xr, mode), "port"
        solver.set_uh(uh)
        solver.set_scalars(0.0, 0.0, 1e9)
    steps = 0
    t0 = time.perf_counter()
    while True:
        solver.step()
        steps += 1
        el = time.perf_counter() - t0
        if el >= target_seconds:
            break
    dof = steps * si * sj * 4 * (N + 1) ** 2
    return {"value": dof / el, "unit": "element-DOF updates/s", "cores": 1, "kind": kind,
            "sample": "%s IC, %dx%d cells, N=%d, %s, %d steps in %.1f s, one process" % (ic, si, sj, N, bc, steps, el)}


def copy_probe(nbytes):
    """One device-to-device copy of known size (the HIP runtime's copy kernel): the
    FETCH_SIZE / WRITE_SIZE it reads under rocprofv3 --pmc calibrate the counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy probe: %d bytes device to device" % nbytes)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="euler_vortex_512_n4", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true", help="time graph replay instead of evented launches")
    ap.add_argument("--copy-probe-bytes", type=int, default=0,
                    help="after the run, copy this many bytes device to device (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not supported "
              "(independent solves shard across GPUs without communication)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import aderdg
    from aderdg import ics

    ic, N, I, J, xl, xr, bc = WORKLOADS[args.workload]
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
    dof_per_cell = 4 * (N + 1) ** 2
    value = args.steps * ncell * dof_per_cell / el
    fl_pred, fl_upd = aderdg.flops_per_cell(N)
    per = {name: (ms[k] / launches[k] if launches[k] else None) for k, name in enumerate(aderdg.TIMING_SLOTS)}
    log("per-kernel average ms:", json.dumps(per))
    log("simulated time %.6g after %d steps, last dt %.6g, nRecompute %d" % (t1_sim, args.steps, dt, nrec))
    roofline = None
    if per["predictor"]:
        achieved = fl_pred * ncell / (per["predictor"] * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / FP64_PEAK_TFLOPS, "traffic": PREDICTOR_TRAFFIC_BYTES.get(args.workload),
                    "traffic_unit": "bytes/launch", "kernel": "k_predictor",
                    "flop_per_cell": fl_pred, "cells_per_launch": ncell,
                    "note": "FP64 vector ALU; MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s"}
    out = {
        "metric": "element-DOF updates/s (plain single-instance Euler step)",
        "value": value,
        "unit": "element-DOF updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic initial condition)",
        "config": {"workload": args.workload, "pde": "euler", "ic": ic, "N": N, "cells": [I, J], "bc": bc,
                   "cell_updates_per_s": args.steps * ncell / el},
        "roofline": roofline,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_seconds)
    if args.copy_probe_bytes:
        copy_probe(args.copy_probe_bytes)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
