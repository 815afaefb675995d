"""Benchmark of the explicit structured-block update on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): a 3D block of
256^3 cells on the unit cube, ideal air, FP64 plain form; smooth random-mode
field with a planar Mach-3 normal shock at x = 0.3; periodic y/z, supersonic
inflow west, extrapolated outflow east; MUSCL/van Albada; one step = one
predictor-corrector (RK2) step of gasdynamic_explicit_increment_with_fixed_grid
plus the CFL reduction every 10 steps.  The headline value uses AUSMDV; the
HLLC rate is reported beside it.  Inputs are resident in HBM before timing.

Prints one JSON line.  --gpus N with N > 1 exits: this build runs on one GPU.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per cell-update (one full RK2 step of one cell):
# stage 1 reads U0 and writes U1 (80 B), stage 2 reads U1 and U0 and writes
# U2 (120 B); the CFL pass reads U0 (40 B) once every 10 steps.
ALG_BYTES_PER_CELL_STEP = 200.0
ALG_BYTES_CFL_AMORTISED = 4.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/<name>)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch_avg:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(n_sample=64, steps=2):
    """The CPU oracle (a port of the reference algorithm) on a bounded sample of
    the same workload: the shocked-box field at n_sample^3 cells, `steps` RK2
    steps, all host threads OpenMP gives it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    from eilmer4_amd import IdealGas, workloads
    gas = IdealGas()
    geom, fields, bcs, _ = workloads.shocked_box_3d(gas, n=n_sample)
    U0 = workloads.initial_U(gas, fields)
    b = O.Block(O.Gas.ideal(), 3, geom["ni"], geom["nj"], geom["nk"], geom["dx"], geom["dy"], geom["dz"])
    b.set_config()
    for face, (kind, fs) in bcs.items():
        b.set_bc(int(face), int(kind), fs)
    b.set_U(U0)
    dt_allow, _, _ = b.determine_dt(1.0e-3, 0.5, False)
    b.step(dt_allow)                      # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        rc, _ = b.step(dt_allow)
        assert rc == 0
    el = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=n_sample ** 3 * steps / el / 1e6, unit="Mcell-updates/s", cores=cores, kind="port",
                sample="%d^3 cells of the shocked-box workload, %d RK2 steps (ausmdv), %.1f s" % (n_sample, steps, el))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested --gpus %d)" % args.gpus)
        sys.exit(2)

    import torch
    import eilmer4_amd as e4
    from eilmer4_amd import workloads

    L = e4.load()
    e4.lib.check(L.e4hip_init(0))
    gas = e4.IdealGas()
    log("generating %d^3 shocked-box field" % args.n)
    geom, fields, bcs, _ = workloads.shocked_box_3d(gas, n=args.n)
    U0 = workloads.initial_U(gas, fields)
    del fields
    ncells = geom["ni"] * geom["nj"] * geom["nk"]
    stream = torch.cuda.ExternalStream(L.e4hip_stream())

    results = {}
    for flux in ("ausmdv", "hllc"):
        blk = e4.SFluidBlock(gas, 3, geom["ni"], geom["nj"], geom["nk"], geom["dx"], geom["dy"], geom["dz"])
        for face, (kind, fs) in bcs.items():
            blk.set_bc(face, kind, fs)
        blk.upload(U0)
        cfg = e4.GlobalConfig(flux_calculator=e4.flux_calculator_from_name(flux), cfl_value=0.5)
        sim = e4.Simulation([blk], cfg)
        log("%s: warmup %d steps" % (flux, args.warmup))
        sim.integrate_in_time(target_time=1e30, max_step=args.warmup)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        t0 = time.perf_counter()
        sim.integrate_in_time(target_time=1e30, max_step=args.warmup + args.steps)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ev_ms = ev0.elapsed_time(ev1)
        # per-kernel durations with HIP events around each launch (a separate pass)
        L.e4hip_profile_enable(1)
        sim.integrate_in_time(target_time=1e30, max_step=sim.state.step + 10)
        tot, cnt = C.c_double(0), C.c_long(0)
        e4.lib.check(L.e4hip_profile_read(0, C.byref(tot), C.byref(cnt)))
        stage_ms = tot.value / max(cnt.value, 1)
        e4.lib.check(L.e4hip_profile_read(1, C.byref(tot), C.byref(cnt)))
        cfl_ms = tot.value / max(cnt.value, 1)
        L.e4hip_profile_enable(0)
        results[flux] = dict(wall=wall, ev_ms=ev_ms, stage_ms=stage_ms, cfl_ms=cfl_ms,
                             steps_done=sim.state.step, dt=sim.state.dt_global)
        log("%s: %d steps in %.3f s (events %.3f ms), stage launch %.3f ms, cfl %.3f ms" %
            (flux, args.steps, wall, ev_ms, stage_ms, cfl_ms))
        blk.destroy()

    r = results["ausmdv"]
    ms_per_step = r["wall"] * 1e3 / args.steps
    value = ncells * args.steps / r["wall"] / 1e6
    # dominant kernel: the fused stage kernel; its algorithmic bytes per launch
    # are half of a cell-step's 200 B (two stage launches per step)
    alg_bytes_launch = ALG_BYTES_PER_CELL_STEP / 2.0 * ncells
    achieved = alg_bytes_launch / (r["stage_ms"] * 1e-3) / 1e9
    traffic = read_traffic("stage_kernel_256cube_traffic.txt")
    out = {
        "metric": "Mcell-updates/s plain FP64 vs mixed+learned-EOS at <=1e-5 rel err; % HBM peak",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D single block %d^3, ideal air, FP64 plain form, MUSCL van Albada + ausmdv, "
                               "pc (RK2), shocked random-mode field" % args.n,
                   "cells": ncells, "flux_calculator": "ausmdv", "update": "pc", "precision": "fp64"},
        "hllc": {"value": round(ncells * args.steps / results["hllc"]["wall"] / 1e6, 3),
                 "ms_per_step": round(results["hllc"]["wall"] * 1e3 / args.steps, 4),
                 "stage_launch_ms": round(results["hllc"]["stage_ms"], 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "stage_kernel<32,8,3,ausmdv>", "stage_launch_ms": round(r["stage_ms"], 4),
                     "cfl_launch_ms": round(r["cfl_ms"], 4),
                     "alg_bytes_per_launch": alg_bytes_launch},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
