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


def cpu_baseline(n_sample=128, min_seconds=10.0, max_steps=100):
    """The CPU oracle (a port of the reference algorithm) on a bounded sample of
    the same workload: the shocked-box field at n_sample^3 cells, RK2 steps
    until min_seconds of CPU work (at least 2), all host threads OpenMP gives it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    from eilmer4_amd import IdealGas, workloads
    gas = IdealGas()
    geom, fields, bcs, _ = workloads.shocked_box_3d(gas, n=n_sample)
    U0 = workloads.initial_U(gas, fields)
    b = O.Block(O.Gas.ideal(), 3, geom["ni"], geom["nj"], geom["nk"], geom["dx"], geom["dy"], geom["dz"])
    b.set_config(flux_calc=O.AUSMDV)      # the headline workload's calculator
    for face, (kind, fs) in bcs.items():
        b.set_bc(int(face), int(kind), fs)
    b.set_U(U0)
    dt_allow, _, _ = b.determine_dt(1.0e-3, 0.5, False)
    b.step(dt_allow)                      # warm-up
    t0 = time.perf_counter()
    steps = 0
    while steps < 2 or (time.perf_counter() - t0 < min_seconds and steps < max_steps):
        rc, _ = b.step(dt_allow)
        assert rc == 0
        steps += 1
    el = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=n_sample ** 3 * steps / el / 1e6, unit="Mcell-updates/s", cores=cores, kind="port",
                sample="%d^3 cells of the shocked-box workload, %d RK2 steps (ausmdv), %.1f s" % (n_sample, steps, el))


def tp_workload(e4, workloads, L, n=128, steps=5, warmup=1):
    """The 5-species thermally perfect workload (n^3, AUSMDV, pc), each variant
    timed over `steps` steps at one fixed dt from the same state:
      plain          FP64, the reference's bracketed Newton from the stored T
      fp64+learned   FP64, decode from the learned guess (polish, fallback)
      mixed          Newton polish from the cell/stored T at faces and decode
      mixed+learned  as mixed, decode from the learned guess
    with each variant's difference from plain after the run."""
    import torch
    from eilmer4_amd import learned_eos
    gas = e4.ThermallyPerfectGas.from_json()
    model = learned_eos.GuessModel.load()
    model.install(gas)
    geom, fields, bcs, _ = workloads.thermally_perfect_box_3d(gas, n=n)
    U0, Tg = workloads.initial_U_tp(gas, fields)
    del fields
    out, finals = {}, {}
    dt_fixed = None
    variants = [("plain", "fp64", False), ("fp64+learned", "fp64", True), ("mixed", "mixed", False),
                ("mixed+learned", "mixed", True)]
    for name, precision, learned in variants:
        blk = e4.SFluidBlock(gas, 3, n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
        for face, (kind, fs) in bcs.items():
            blk.set_bc(face, kind, fs)
        sim = e4.Simulation([blk], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                                   learned_eos_guess=learned,
                                                   precision_mode=e4.PrecisionMode[precision]))
        blk.upload(U0, Tg)
        if dt_fixed is None:
            dt_fixed = sim.block_dt(1e-3, False)[0]
        sim.cfg.fixed_time_step = True
        sim.state.dt_global = dt_fixed
        sim.integrate_in_time(target_time=1e30, max_step=warmup)
        torch.cuda.synchronize()
        learned_eos.decode_stats(reset=True)
        learned_eos.face_stats(reset=True)
        t0 = time.perf_counter()
        sim.integrate_in_time(target_time=1e30, max_step=warmup + steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ev, fb = learned_eos.decode_stats(reset=True)
        fev, ffb = learned_eos.face_stats(reset=True)
        finals[name] = blk.download(with_fs=True)
        nside = 2 * 3 * n * n * (n + 1) * steps * 2    # face-state solves: two per face, two stages per step
        out[name] = {
            "value": round(n ** 3 * steps / wall / 1e6, 3), "ms_per_step": round(wall * 1e3 / steps, 3),
            "decode_energy_evals_per_cell_stage": round(ev / (n ** 3 * steps * 2), 3),
            "decode_fallback_fraction": round(fb / (n ** 3 * steps * 2), 5),
            "face_energy_evals_per_solve": round(fev / nside, 3),
            "face_fallback_fraction": round(ffb / nside, 5)}
        blk.destroy()
    Ur, fr = finals["plain"]
    for name, _, _ in variants[1:]:
        U, f = finals[name]
        out[name]["vs_plain"] = {
            "steps": warmup + steps, "speedup": round(out[name]["value"] / out["plain"]["value"], 3),
            "rho_linf_rel": float(np.max(np.abs(U[0] - Ur[0])) / np.max(np.abs(Ur[0]))),
            "T_linf_rel": float(np.max(np.abs(f[2] - fr[2]) / fr[2]))}
    out["restart_decode"] = restart_decode(e4, learned_eos, gas, U0, Tg, n)
    out["workload"] = "3D %d^3 5-species thermally perfect air (N2,O2,NO,N,O), T 300-8000 K, ausmdv, pc" % n
    out["guess_model"] = "7-64-64-64-1 SiLU, BF16 MFMA inference, trained on 2.16 M device Newton states"
    return out


def restart_decode(e4, learned_eos, gas, U0, Tg, n, reps=3):
    """Where the learned guess pays: the decode of a mixture state uploaded
    without its temperature (a restart from conserved quantities only), every
    cell starting from 300 K -- the reference's bracketed solve against the
    learned guess with its FP64 polish.  Device time of the decode kernel
    (HIP events, best of `reps`), energy evaluations per cell, and the two
    results' largest T difference relative to the field's own T."""
    L = e4.load()
    Tflat = np.full_like(Tg, 300.0)
    out, Ts = {}, {}
    for name, learned in (("bracketed", False), ("learned", True)):
        blk = e4.SFluidBlock(gas, 3, n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
        e4.Simulation([blk], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv, learned_eos_guess=learned,
                                             precision_mode=e4.PrecisionMode.fp64))
        best, evals = 1e30, 0
        for _ in range(reps):
            learned_eos.decode_stats(reset=True)
            L.e4hip_profile_enable(1)
            blk.upload(U0, Tflat)
            tot, cnt = C.c_double(0), C.c_long(0)
            e4.lib.check(L.e4hip_profile_read(2, C.byref(tot), C.byref(cnt)))
            L.e4hip_profile_enable(0)
            ev, fb = learned_eos.decode_stats(reset=True)
            best, evals = min(best, tot.value), ev
        Ts[name] = blk.download(with_fs=True)[1][2].copy()
        out[name] = {"decode_ms": round(best, 4), "energy_evals_per_cell": round(evals / n ** 3, 3)}
        blk.destroy()
    out["speedup"] = round(out["bracketed"]["decode_ms"] / out["learned"]["decode_ms"], 3)
    out["T_linf_rel_learned_vs_bracketed"] = float(np.max(np.abs(Ts["learned"] - Ts["bracketed"]) / Ts["bracketed"]))
    out["T_linf_rel_vs_field"] = float(np.max(np.abs(Ts["bracketed"] - Tg) / Tg))
    return out


def c5_workload(e4, workloads, steps=6, warmup=3):
    """The batched multi-block workload (48 blocks: 32 ideal air, 16
    five-species air, random face-matched connections in every orientation
    within each gas class).  Variants, each from the same state at one fixed dt:
      plain                  FP64, reference Newton, one launch per block
      fp64+batched           FP64, each gas class's blocks of a stage in one
                             launch per kernel (the ideal blocks' stage kernel
                             persistent over a tile queue), steps replayed
                             from captured HIP graphs
      mixed+batched          mixed precision (FP32 reconstruction and flux,
                             short face solves, FP64 decode), batched, graphs
      mixed+learned+batched  as mixed+batched with the learned decode guess
    with each variant's speed-up over plain and its difference from plain.
    The warm-up steps include the k-chunk tuning and the graph captures."""
    import torch
    from eilmer4_amd import learned_eos
    ig, tg = e4.IdealGas(), e4.ThermallyPerfectGas.from_json()
    learned_eos.GuessModel.load().install(tg)
    t0 = time.perf_counter()
    descs, conns = workloads.multiblock_c5(ig, tg)
    ncells = sum(d["ni"] * d["nj"] * d["nk"] for d in descs)
    log("C5 workload: %d blocks, %.2f M cells, generated in %.1f s" % (len(descs), ncells / 1e6,
                                                                     time.perf_counter() - t0))
    variants = [("plain", "fp64", False, False), ("fp64+batched", "fp64", True, False),
                ("mixed+batched", "mixed", True, False), ("mixed+learned+batched", "mixed", True, True)]
    out, finals, dt_fixed = {}, {}, None
    for name, prec, batch, learned in variants:
        blks = []
        for d in descs:
            g = ig if d["cls"] == "ideal" else tg
            b = e4.SFluidBlock(g, 3, d["ni"], d["nj"], d["nk"], d["dx"], d["dy"], d["dz"])
            for face, (kind, fs) in d["bcs"].items():
                b.set_bc(face, kind, fs)
            blks.append(b)
        for (a, fa, b, fb, o) in conns:
            e4.connect_blocks(blks[a], fa, blks[b], fb, o)
        sim = e4.Simulation(blks, e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                                  precision_mode=e4.PrecisionMode[prec], batch_blocks=batch,
                                                  use_graphs=batch, learned_eos_guess=learned))
        for b, d in zip(blks, descs):
            b.upload(d["U0"], d["Tguess"])
        if dt_fixed is None:
            dt_fixed = 0.5 * sim.block_dt(1e-3, False)[0]
        sim.cfg.fixed_time_step = True
        sim.state.dt_global = dt_fixed
        sim.integrate_in_time(target_time=1e30, max_step=warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sim.integrate_in_time(target_time=1e30, max_step=warmup + steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        finals[name] = [b.download()[0] for b in blks]
        out[name] = {"value": round(ncells * steps / wall / 1e6, 3), "ms_per_step": round(wall * 1e3 / steps, 3)}
        for b in blks:
            b.destroy()
        log("  C5 %s: %s" % (name, out[name]))
    for name, _, _, _ in variants[1:]:
        err = max(float(np.max(np.abs(a - r)) / np.max(np.abs(r))) for a, r in zip(finals[name], finals["plain"]))
        out[name]["vs_plain"] = {"speedup": round(out[name]["value"] / out["plain"]["value"], 3),
                                 "rho_linf_rel": err, "steps": warmup + steps}
    out["workload"] = ("48 blocks (32 ideal air, 16 five-species air), %.2f M cells, sizes %s..%s, %d face-matched "
                       "connections (orientations 0-3), ausmdv, pc, fixed dt" %
                       (ncells / 1e6, min((d["ni"], d["nj"], d["nk"]) for d in descs),
                        max((d["ni"], d["nj"], d["nk"]) for d in descs), len(conns)))
    return out


def shock_position(U, n):
    """Mean over (j,k) lines of the i-position (cells) of the largest density jump."""
    rho = U[0].reshape(n, n, n)
    jump = np.abs(np.diff(rho, axis=2))
    return float(np.mean(np.argmax(jump, axis=2)) + 0.5)


def accuracy_study(e4, workloads, sizes=(256, 384), checkpoints=(1, 100, 1000), seed=20261203):
    """Mixed precision against plain FP64 on the shocked box (seed 20261203) at
    each size after 1, 100, 1000 steps, both driven with the plain run's dt
    sequence: L-infinity differences of density, momentum and total energy
    relative to each field's maximum, and the shock positions."""
    import torch
    gas = e4.IdealGas()
    out = {}
    for n in sizes:
        log("accuracy study %d^3" % n)
        geom, fields, bcs, _ = workloads.shocked_box_3d(gas, n=n, seed=seed)
        U0 = workloads.initial_U(gas, fields)
        del fields
        runs = {}
        for prec in ("fp64", "mixed"):
            blk = e4.SFluidBlock(gas, 3, n, n, n, geom["dx"], geom["dy"], geom["dz"])
            for face, (kind, fs) in bcs.items():
                blk.set_bc(face, kind, fs)
            sim = e4.Simulation([blk], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                                       precision_mode=e4.PrecisionMode[prec]))
            blk.upload(U0)
            runs[prec] = (blk, sim)
        del U0
        res = []
        step = 0
        t0 = time.perf_counter()
        for cp in checkpoints:
            while step < cp:
                blk64, sim64 = runs["fp64"]
                sim64.apply_config()
                sim64.determine_time_step_size()
                sim64.gasdynamic_explicit_increment_with_fixed_grid()
                sim64.state.step += 1
                blkm, simm = runs["mixed"]
                simm.apply_config()
                simm.state.dt_global = sim64.state.dt_global
                simm.gasdynamic_explicit_increment_with_fixed_grid()
                simm.state.step += 1
                step += 1
            torch.cuda.synchronize()
            Ua = runs["fp64"][0].download()
            Ub = runs["mixed"][0].download()
            def rel(q):
                return float(np.max(np.abs(Ub[q] - Ua[q])) / np.max(np.abs(Ua[q])))
            xa, xb = shock_position(Ua, n), shock_position(Ub, n)
            res.append({"steps": step, "rho_linf_rel": rel(0), "momentum_linf_rel": max(rel(1), rel(2), rel(3)),
                        "energy_linf_rel": rel(4), "shock_x_cells_fp64": round(xa, 4),
                        "shock_x_cells_mixed": round(xb, 4), "shock_shift_cells": round(xb - xa, 6),
                        "time": runs["fp64"][1].state.time})
            log("  %d^3 after %d steps: %s" % (n, step, res[-1]))
            del Ua, Ub
        for blk, _ in runs.values():
            blk.destroy()
        out["%d^3" % n] = res
        log("  %d^3 study took %.1f s" % (n, time.perf_counter() - t0))
    return out


def graph_study(e4, nb=2, n=8, steps=2000):
    """Per-step wall time of many small connected blocks with and without
    graph replay (use_graphs): nb x nb 2D blocks of n x n cells on the unit
    square, Mach-2.3 flow at 45 degrees, supersonic inflow west/south,
    x_order-1 extrapolation north/east, ausmdv, pc, fixed dt."""
    import time
    gas = e4.IdealGas()
    h = 1.0 / (nb * n)
    fs = gas.flowstate(1.0e5, 300.0, (560.0, 560.0, 0.0))
    out = {}
    for graphs in (False, True, False, True):
        blks = {}
        for jb in range(nb):
            for ib in range(nb):
                b = e4.SFluidBlock(gas, 2, n, n, 1, h, h)
                b.set_from_pT(1.0e5, 300.0, 560.0, 560.0, 0.0)
                if ib == 0:
                    b.set_bc(e4.Face.west, e4.BCType.inflow, fs)
                if jb == 0:
                    b.set_bc(e4.Face.south, e4.BCType.inflow, fs)
                if ib == nb - 1:
                    b.set_bc(e4.Face.east, e4.BCType.extrapolate_1)
                if jb == nb - 1:
                    b.set_bc(e4.Face.north, e4.BCType.extrapolate_1)
                blks[(ib, jb)] = b
        for (ib, jb), b in blks.items():
            if ib + 1 < nb:
                e4.connect_blocks(b, e4.Face.east, blks[(ib + 1, jb)], e4.Face.west)
            if jb + 1 < nb:
                e4.connect_blocks(b, e4.Face.north, blks[(ib, jb + 1)], e4.Face.south)
        blocks = list(blks.values())
        sim = e4.Simulation(blocks, e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv, use_graphs=graphs))
        sim.state.dt_global = 1.0e-6
        for _ in range(20):
            sim.gasdynamic_explicit_increment_with_fixed_grid()
        sim.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.gasdynamic_explicit_increment_with_fixed_grid()
        sim.synchronize()
        out["graphs" if graphs else "eager"] = round(1.0e6 * (time.perf_counter() - t0) / steps, 1)
        for b in blocks:
            b.destroy()
    return {"workload": "%d x %d 2D blocks of %d x %d cells, ausmdv, pc, fixed dt" % (nb, nb, n, n),
            "us_per_step": out, "steps": steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--flux", default="ausmdv,hllc,adaptive_hanel_ausmdv",
                    help="comma list of flux calculators to run; the first is the headline")
    ap.add_argument("--precision", default="fp64,mixed,fp64_exact",
                    help="comma list of precision modes; the first is the headline")
    ap.add_argument("--acc-steps", type=int, default=20, help="steps of the mixed-vs-fp64 accuracy check")
    ap.add_argument("--tp-n", type=int, default=128, help="cells per side of the 5-species workload (0: skip)")
    ap.add_argument("--c5", type=int, default=1, help="run the batched multi-block workload (0: skip)")
    ap.add_argument("--accuracy-study", action="store_true",
                    help="only run the mixed-vs-FP64 study (256^3 and 384^3, 1/100/1000 steps)")
    ap.add_argument("--graph-study", action="store_true",
                    help="only time many small blocks with and without graph replay of steps")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM-counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": n * 8, "calibrate_copy_bytes_written": n * 8}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (requested --gpus %d)" % args.gpus)
        sys.exit(2)

    import torch
    import eilmer4_amd as e4
    from eilmer4_amd import workloads

    L = e4.load()
    e4.lib.check(L.e4hip_init(0))
    if args.graph_study:
        print(json.dumps({"graph_study": [graph_study(e4, nb, n) for nb, n in ((2, 8), (4, 16), (8, 32))]}),
              flush=True)
        return
    if args.accuracy_study:
        print(json.dumps({"accuracy_study": accuracy_study(e4, workloads), "workload": "shocked box seed 20261203",
                          "vs": "plain FP64 device run, same dt sequence"}), flush=True)
        return
    gas = e4.IdealGas()
    log("generating %d^3 shocked-box field" % args.n)
    geom, fields, bcs, _ = workloads.shocked_box_3d(gas, n=args.n)
    U0 = workloads.initial_U(gas, fields)
    del fields
    ncells = geom["ni"] * geom["nj"] * geom["nk"]
    stream = torch.cuda.ExternalStream(L.e4hip_stream())

    def make_sim(flux, precision, U):
        blk = e4.SFluidBlock(gas, 3, geom["ni"], geom["nj"], geom["nk"], geom["dx"], geom["dy"], geom["dz"])
        for face, (kind, fs) in bcs.items():
            blk.set_bc(face, kind, fs)
        blk.upload(U)
        cfg = e4.GlobalConfig(flux_calculator=e4.flux_calculator_from_name(flux), cfl_value=0.5,
                              precision_mode=e4.PrecisionMode[precision])
        return blk, e4.Simulation([blk], cfg)

    results = {}
    runs = [(f, p) for p in args.precision.split(",") for f in args.flux.split(",")]
    for flux, precision in runs:
        blk, sim = make_sim(flux, precision, U0)
        log("%s/%s: warmup %d steps" % (flux, precision, args.warmup))
        sim.integrate_in_time(target_time=1e30, max_step=args.warmup)
        torch.cuda.synchronize()
        kc = C.c_int(0)
        e4.lib.check(L.e4hip_block_kchunk(blk.id, C.byref(kc)))   # the autotuner's choice
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
        results[(flux, precision)] = dict(wall=wall, ev_ms=ev_ms, stage_ms=stage_ms, cfl_ms=cfl_ms,
                                          steps_done=sim.state.step, dt=sim.state.dt_global, kchunk=kc.value)
        log("%s/%s: %d steps in %.3f s (events %.3f ms), stage launch %.3f ms, cfl %.3f ms, kchunk %d" %
            (flux, precision, args.steps, wall, ev_ms, stage_ms, cfl_ms, kc.value))
        blk.destroy()

    # accuracy of the mixed form against the plain FP64 form: same U0, same
    # fixed dt, `acc_steps` steps each
    accuracy = {}
    if "mixed" in args.precision.split(",") and args.acc_steps > 0:
        dt_fixed = results[runs[0]]["dt"]
        finals = {}
        for precision in ("fp64", "mixed"):
            blk, sim = make_sim("ausmdv", precision, U0)
            sim.cfg.fixed_time_step = True
            sim.state.dt_global = dt_fixed
            sim.integrate_in_time(target_time=1e30, max_step=args.acc_steps)
            finals[precision] = blk.download()
            blk.destroy()
        a, b = finals["mixed"], finals["fp64"]
        rel = [float(np.max(np.abs(a[q] - b[q])) / np.max(np.abs(b[q]))) for q in range(5)]
        accuracy = {"steps": args.acc_steps, "flux": "ausmdv", "vs": "plain FP64 device run",
                    "rho_linf_rel": rel[0], "momentum_linf_rel": max(rel[1:4]), "energy_linf_rel": rel[4]}
        log("mixed vs fp64 after %d steps: %s" % (args.acc_steps, accuracy))
        del finals, a, b

    head = (args.flux.split(",")[0], args.precision.split(",")[0])
    r = results[head]
    ms_per_step = r["wall"] * 1e3 / args.steps
    value = ncells * args.steps / r["wall"] / 1e6
    # dominant kernel: the fused stage kernel of the headline variant; its
    # algorithmic bytes per launch are half of a cell-step's 200 B (two stage
    # launches per step)
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
        "variants": {"%s/%s" % k: {"value": round(ncells * args.steps / v["wall"] / 1e6, 3),
                                  "ms_per_step": round(v["wall"] * 1e3 / args.steps, 4),
                                  "stage_launch_ms": round(v["stage_ms"], 4), "kchunk": v["kchunk"],
                                  "stage_hbm_frac": round(ALG_BYTES_PER_CELL_STEP / 2.0 * ncells /
                                                          (v["stage_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
                     for k, v in results.items()},
        "mixed_accuracy": accuracy,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "stage_kernel<%s,32,8,3,%s>" % (head[1], head[0]), "stage_launch_ms": round(r["stage_ms"], 4),
                     "cfl_launch_ms": round(r["cfl_ms"], 4), "kchunk": r["kchunk"],
                     "alg_bytes_per_launch": alg_bytes_launch},
    }
    if args.tp_n > 0:
        log("thermally perfect %d^3 workload ..." % args.tp_n)
        out["thermally_perfect"] = tp_workload(e4, workloads, L, n=args.tp_n)
        log("thermally perfect: %s" % out["thermally_perfect"])
    if args.c5:
        out["multiblock_c5"] = c5_workload(e4, workloads)
        log("multi-block: %s" % out["multiblock_c5"])
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
