"""Parity with the oracle at the benchmark workloads' own sizes.

  256^3 shocked box, ideal air (the headline workload): one RK2 step with
      ausmdv and hllc, bit-identical in fp64_exact and within 1e-12 per
      conserved quantity in fp64; 100 RK2 steps with ausmdv in fp64 within
      1e-9, the device driven with the oracle's dt sequence.
  384^3 shocked box (seed 20261203), mixed precision: one RK2 step, density
      L-inf relative error <= 1e-5.
  128^3 five-species air: one RK2 step in fp64, without and with the learned
      temperature guess: T within 1e-10 of the oracle; U within 1e-9 (U takes
      the Newton's last iterate through p and through u's side effect, and the
      solver stops within its own |dT| < 1e-6 K, newton.d, so a different start
      or rounding may stop one iterate apart; observed 1e-10 .. 2e-10).
  48 connected blocks (32 ideal, 16 five-species, 35 M cells): one RK2 step,
      batched, in fp64_exact (ideal blocks bit-identical, mixture blocks within
      1e-10) and in the mixed + learned-guess form (density within 1e-5; the
      device's temperatures within 1e-10 of the oracle's decode of the same
      conserved state).

Errors are max|dU_q| / max|U_q| per conserved quantity (per block).  The
oracle runs on every host thread (OpenMP); these tests take minutes.
"""
import numpy as np
import pytest

import pyoracle as O
import eilmer4_amd as e4
from eilmer4_amd import learned_eos, workloads
from parity_helpers import oracle_block, device_block, device_sim, rel_err

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


@pytest.fixture(scope="module")
def box256(gpu):
    gas = e4.IdealGas()
    print("generating the 256^3 shocked box", flush=True)
    geom, fields, bcs, _ = workloads.shocked_box_3d(gas, n=256)
    U0 = workloads.initial_U(gas, fields)
    return gas, geom, bcs, U0


def _device_step(gas, geom, bcs, U0, flux, precision, dts):
    db = device_block(gas, geom, bcs, U0)
    sim = device_sim([db], flux=flux, precision=precision)
    for dt in dts:
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    U = db.download()
    db.destroy()
    return U


@pytest.mark.parametrize("flux", ["ausmdv", "hllc"])
def test_box256_one_step(box256, flux):
    gas, geom, bcs, U0 = box256
    print("case box256 one step flux=%s" % flux, flush=True)
    fc = e4.flux_calculator_from_name(flux)
    ob = oracle_block(geom, bcs, U0, flux=int(fc))
    dt = ob.determine_dt(1e-3, 0.5, False)[0]
    rc, nbad = ob.step(dt)
    assert rc == 0
    Uo = ob.get_U()
    del ob
    Ue = _device_step(gas, geom, bcs, U0, fc, "fp64_exact", [dt])
    print("  fp64_exact bitwise %s" % np.array_equal(Ue, Uo), flush=True)
    assert np.array_equal(Ue, Uo)
    del Ue
    Uf = _device_step(gas, geom, bcs, U0, fc, "fp64", [dt])
    err = rel_err(Uf, Uo)
    print("  fp64 rel err per quantity %s" % err, flush=True)
    assert err.max() <= 1.0e-12


def test_box256_100_steps_fp64(box256):
    gas, geom, bcs, U0 = box256
    print("case box256 100 steps ausmdv fp64 (oracle dt sequence)", flush=True)
    ob = oracle_block(geom, bcs, U0, flux=O.AUSMDV)
    osim = O.Simulation(ob, cfl_value=0.5)
    db = device_block(gas, geom, bcs, U0)
    sim = device_sim([db], flux=e4.FluxCalculator.ausmdv, precision="fp64")
    for s in range(100):
        osim.run(max_step=s + 1)
        sim.state.dt_global = osim.dt_global
        sim.gasdynamic_explicit_increment_with_fixed_grid()
        if (s + 1) % 20 == 0:
            print("  step %d" % (s + 1), flush=True)
    err = rel_err(db.download(), ob.get_U())
    print("  fp64 after 100 steps: rel err per quantity %s" % err, flush=True)
    db.destroy()
    assert err.max() <= 1.0e-9


def test_box384_mixed_one_step(gpu):
    gas = e4.IdealGas()
    print("case box384 mixed one step (seed 20261203)", flush=True)
    geom, fields, bcs, _ = workloads.shocked_box_3d(gas, n=384, seed=20261203)
    U0 = workloads.initial_U(gas, fields)
    del fields
    ob = oracle_block(geom, bcs, U0, flux=O.AUSMDV)
    dt = ob.determine_dt(1e-3, 0.5, False)[0]
    assert ob.step(dt)[0] == 0
    Uo = ob.get_U()
    del ob
    Um = _device_step(gas, geom, bcs, U0, e4.FluxCalculator.ausmdv, "mixed", [dt])
    err = rel_err(Um, Uo)
    print("  mixed rel err rho %.3e momentum %.3e energy %.3e" % (err[0], err[1:4].max(), err[4]), flush=True)
    assert err[0] <= 1.0e-5


def _oracle_decode_T(og, ob_like, U, Tguess):
    """The oracle's temperature of the conserved state U (its decode, Newton
    from Tguess)."""
    b = O.Block(og, 3, ob_like["ni"], ob_like["nj"], ob_like["nk"], 1.0, 1.0, 1.0)
    b.set_U(U, Tguess)
    return b.get_fs()[2]


@pytest.mark.parametrize("precision,learned", [("fp64", False), ("fp64", True), ("mixed", False), ("mixed", True)])
def test_tp128_one_step(gpu, precision, learned):
    """fp64: U within 1e-9 and T within 1e-10 of the oracle; mixed (FP32
    reconstruction and flux, short face temperature solves, FP64 decode):
    U and T within 1e-5."""
    print("case tp128 one step %s learned=%s" % (precision, learned), flush=True)
    gas = e4.ThermallyPerfectGas.from_json()
    if learned:
        learned_eos.GuessModel.load().install(gas)
    n = 128
    geom, fields, bcs, _ = workloads.thermally_perfect_box_3d(gas, n=n)
    U0, Tg = workloads.initial_U_tp(gas, fields)
    del fields
    og = O.Gas.from_json("therm-perf-5-species-air.json")
    ob = O.Block(og, 3, n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
    ob.set_config(flux_calc=O.AUSMDV)
    for face, (kind, fs) in bcs.items():
        ob.set_bc(int(face), int(kind), fs)
    ob.set_U(U0, Tg)
    dt = ob.determine_dt(1e-3, 0.5, False)[0]
    assert ob.step(dt)[0] == 0
    Uo, fso = ob.get_U(), ob.get_fs()
    del ob
    db = e4.SFluidBlock(gas, 3, n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
    for face, (kind, fs) in bcs.items():
        db.set_bc(face, kind, fs)
    sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv, learned_eos_guess=learned,
                                              precision_mode=e4.PrecisionMode[precision]))
    db.upload(U0, Tg)
    sim.state.dt_global = dt
    learned_eos.decode_stats(reset=True)
    sim.gasdynamic_explicit_increment_with_fixed_grid()
    ev, fb = learned_eos.decode_stats(reset=True)
    U, fs = db.download(with_fs=True)
    db.destroy()
    err = rel_err(U, Uo)
    errT = float(np.max(np.abs(fs[2] - fso[2]) / fso[2]))
    print("  U rel err per quantity %s; T rel err %.3e; decode energy evals per cell-stage %.3f, fallbacks %d"
          % (err, errT, ev / (2.0 * n ** 3), fb), flush=True)
    if precision == "mixed":
        assert err.max() <= 1.0e-5
        assert errT <= 1.0e-5
    else:
        assert err.max() <= 1.0e-9
        assert errT <= 1.0e-10


def _c5_blocks(descs, conns, ig, tg, og_ideal, og_tp):
    oblocks = []
    for d in descs:
        og = og_ideal if d["cls"] == "ideal" else og_tp
        b = O.Block(og, 3, d["ni"], d["nj"], d["nk"], d["dx"], d["dy"], d["dz"])
        b.set_config(flux_calc=O.AUSMDV)
        for face, (kind, fs) in d["bcs"].items():
            b.set_bc(int(face), int(kind), fs)
        oblocks.append(b)
    for (a, fa, b, fb, o) in conns:
        O.connect_blocks(oblocks[a], int(fa), oblocks[b], int(fb), o)
    for b, d in zip(oblocks, descs):
        b.set_U(d["U0"], d["Tguess"])
    return oblocks


def _c5_device(descs, conns, ig, tg, precision, learned, dt):
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
                                              precision_mode=e4.PrecisionMode[precision], batch_blocks=True,
                                              learned_eos_guess=learned))
    for b, d in zip(blks, descs):
        b.upload(d["U0"], d["Tguess"])
    sim.state.dt_global = dt
    sim.gasdynamic_explicit_increment_with_fixed_grid()
    out = [b.download(with_fs=(d["cls"] == "tp")) for b, d in zip(blks, descs)]
    for b in blks:
        b.destroy()
    return out


def test_c5_multiblock_one_step(gpu):
    print("case c5 48 blocks: generating", flush=True)
    ig, tg = e4.IdealGas(), e4.ThermallyPerfectGas.from_json()
    learned_eos.GuessModel.load().install(tg)
    descs, conns = workloads.multiblock_c5(ig, tg)
    og_ideal = O.Gas.ideal()
    og_tp = O.Gas.from_json("therm-perf-5-species-air.json")
    oblocks = _c5_blocks(descs, conns, ig, tg, og_ideal, og_tp)
    dt = 0.5 * min(b.determine_dt(1e-3, 0.5, False)[0] for b in oblocks)
    print("  oracle step (dt %.4e)" % dt, flush=True)
    rc, nbad = O.multi_step(oblocks, dt)
    assert rc == 0
    ref = [(b.get_U(), b.get_fs()) for b in oblocks]
    del oblocks
    print("  device fp64_exact, batched", flush=True)
    dev = _c5_device(descs, conns, ig, tg, "fp64_exact", False, dt)
    worst_tp = 0.0
    for (Uo, fso), d, out in zip(ref, descs, dev):
        if d["cls"] == "ideal":
            assert np.array_equal(out, Uo)
        else:
            U, fs = out
            worst_tp = max(worst_tp, rel_err(U, Uo).max(), float(np.max(np.abs(fs[2] - fso[2]) / fso[2])))
    print("  fp64_exact: ideal blocks bit-identical; mixture blocks max rel err (U, T) %.3e" % worst_tp, flush=True)
    assert worst_tp <= 1.0e-10
    del dev
    print("  device mixed + learned guess, batched", flush=True)
    dev = _c5_device(descs, conns, ig, tg, "mixed", True, dt)
    worst_rho, worst_T = 0.0, 0.0
    for (Uo, fso), d, out in zip(ref, descs, dev):
        U = out if d["cls"] == "ideal" else out[0]
        worst_rho = max(worst_rho, rel_err(U, Uo)[0])
        if d["cls"] == "tp":
            fs = out[1]
            T_dec = _oracle_decode_T(og_tp, d, U, fs[2])
            worst_T = max(worst_T, float(np.max(np.abs(fs[2] - T_dec) / T_dec)))
    print("  mixed+learned: rho rel err %.3e; T vs the oracle's decode of the same U %.3e" % (worst_rho, worst_T),
          flush=True)
    assert worst_rho <= 1.0e-5
    assert worst_T <= 1.0e-10
