"""Batched launches over many blocks against the oracle and against one launch
per block:

  * ideal-gas blocks with more tiles than the GPU holds workgroups at once
    (four connected 64^3 blocks: the stage kernel runs persistent, its
    workgroups taking tiles from the device work queue, largest blocks first):
    bit-identical to the oracle's multi-block step and to the per-block
    launches in fp64_exact, over several steps (each launch must leave the
    queue cleared for the next), and within 1e-12 in fp64;
  * the same under graph replay of steps, step by step against the eager run;
  * thermally perfect blocks of several sizes in one launch per kernel:
    against the oracle (1e-9; mixed: density 1e-5) and bit-identical to the
    per-block launches.
"""
import numpy as np
import pytest

import pyoracle as O
import eilmer4_amd as e4
from eilmer4_amd import workloads
from eilmer4_amd.fvcore import Face
from parity_helpers import rel_err

pytestmark = pytest.mark.gpu


def _ideal_ring(n=64, nb=4, seed=41):
    """nb connected n^3 blocks in a periodic ring along x; per-block seeded fields."""
    gas = e4.IdealGas()
    rng = np.random.default_rng(seed)
    U0s = []
    for _ in range(nb):
        shape = (n, n, n)
        p = 1.0e5 * (1.0 + 0.2 * rng.random(shape))
        T = 300.0 * (1.0 + 0.2 * rng.random(shape))
        v = [150.0 + 80.0 * rng.standard_normal(shape) for _ in range(3)]
        rho, u, _ = gas.rho_u_a_from_pT(p.ravel(), T.ravel())
        U0s.append(gas.encode(rho, u, *[x.ravel() for x in v]))
    return gas, U0s


def _device_ring(gas, U0s, n):
    blks = []
    for U0 in U0s:
        b = e4.SFluidBlock(gas, 3, n, n, n, 0.01, 0.01, 0.01)
        b.upload(U0)
        blks.append(b)
    nb = len(blks)
    for b in range(nb):
        e4.connect_blocks(blks[b], Face.east, blks[(b + 1) % nb], Face.west, 0)
    return blks


def test_persistent_batched_stage_matches_oracle(gpu):
    n, nb = 64, 4
    gas, U0s = _ideal_ring(n, nb)
    print("case persistent batch: %d x %d^3 ideal blocks, fp64_exact, 2 steps vs oracle" % (nb, n), flush=True)
    og = O.Gas.ideal()
    obl = []
    for U0 in U0s:
        ob = O.Block(og, 3, n, n, n, 0.01, 0.01, 0.01)
        ob.set_config(flux_calc=O.AUSMDV)
        ob.set_U(U0)
        obl.append(ob)
    for b in range(nb):
        O.connect_blocks(obl[b], O.EAST, obl[(b + 1) % nb], O.WEST, 0)
    dbl = _device_ring(gas, U0s, n)
    # autotune off: the default chunking gives 4 x 16 x 16 = 1024 tiles, more
    # than the 768 workgroups a persistent launch holds, so the queue is used
    sim = e4.Simulation(dbl, e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                             precision_mode=e4.PrecisionMode.fp64_exact, autotune=False))
    for s in range(2):
        assert O.multi_step(obl, 2.0e-6)[0] == 0
        sim.state.dt_global = 2.0e-6
        sim.gasdynamic_explicit_increment_with_fixed_grid()
        for k, (ob, db) in enumerate(zip(obl, dbl)):
            assert np.array_equal(ob.get_U(), db.download()), (s, k)
    for b in dbl:
        b.destroy()


@pytest.mark.parametrize("precision,graphs", [("fp64_exact", False), ("fp64_exact", True), ("fp64", True)])
def test_persistent_batched_stage_matches_per_block_launches(gpu, precision, graphs):
    """Bit-identical in fp64_exact (no contraction, so the persistent loop's
    code cannot round differently); in fp64 the batched and single-block
    kernels are separate instantiations whose fused multiply-adds the compiler
    may place differently, so they agree to the mode's 1e-12."""
    n, nb = 64, 4
    gas, U0s = _ideal_ring(n, nb, seed=43)
    print("case persistent batch vs per block: %d x %d^3, %s, graphs=%s, 6 steps" % (nb, n, precision, graphs),
          flush=True)
    runs = []
    for batch in (False, True):
        dbl = _device_ring(gas, U0s, n)
        sim = e4.Simulation(dbl, e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                                 precision_mode=e4.PrecisionMode[precision], batch_blocks=batch,
                                                 use_graphs=(graphs and batch), autotune=False))
        steps = []
        for s in range(6):
            sim.state.dt_global = 2.0e-6
            sim.gasdynamic_explicit_increment_with_fixed_grid()
            steps.append([b.download() for b in dbl])
        runs.append(steps)
        for b in dbl:
            b.destroy()
    for s, (a, b) in enumerate(zip(*runs)):
        for k, (ua, ub) in enumerate(zip(a, b)):
            if precision == "fp64_exact":
                assert np.array_equal(ua, ub), (s, k)
            else:
                err = rel_err(ub, ua)
                assert err.max() <= 1.0e-12, (s, k, err)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_mixture_blocks(gpu, precision):
    """Three five-species blocks of different sizes, connected in a chain,
    one launch per kernel per stage against one launch per block (bitwise) and
    against the oracle (1e-9; mixed: density 1e-5)."""
    tg = e4.ThermallyPerfectGas.from_json()
    og = O.Gas.from_json("therm-perf-5-species-air.json")
    sizes = [(12, 10, 8), (12, 10, 14), (12, 10, 6)]
    fields = []
    for m, (ni, nj, nk) in enumerate(sizes):
        geom, f, bcs, _ = workloads.thermally_perfect_box_3d(tg, n=ni, seed=700 + m)
        fields.append(workloads.initial_U_tp(tg, f))
    # the boxes are n^3: take the first nj*nk*ni cells' states in order
    states = []
    for (ni, nj, nk), (U0, Tg) in zip(sizes, fields):
        nc = ni * nj * nk
        states.append((np.ascontiguousarray(U0[:, :nc]), np.ascontiguousarray(Tg[:nc])))
    print("case batched mixture blocks %s, %s" % (sizes, precision), flush=True)

    def device(batch):
        blks = []
        for (ni, nj, nk), (U0, Tg) in zip(sizes, states):
            b = e4.SFluidBlock(tg, 3, ni, nj, nk, 0.01, 0.01, 0.01)
            blks.append(b)
        for m in range(len(blks) - 1):
            e4.connect_blocks(blks[m], Face.top, blks[m + 1], Face.bottom, 0)
        sim = e4.Simulation(blks, e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                                  precision_mode=e4.PrecisionMode[precision], batch_blocks=batch))
        for b, (U0, Tg) in zip(blks, states):
            b.upload(U0, Tg)
        return blks, sim

    obl = []
    for (ni, nj, nk), (U0, Tg) in zip(sizes, states):
        ob = O.Block(og, 3, ni, nj, nk, 0.01, 0.01, 0.01)
        ob.set_config(flux_calc=O.AUSMDV)
        obl.append(ob)
    for m in range(len(obl) - 1):
        O.connect_blocks(obl[m], O.TOP, obl[m + 1], O.BOTTOM, 0)
    for ob, (U0, Tg) in zip(obl, states):
        ob.set_U(U0, Tg)
    dt = 0.5 * min(ob.determine_dt(1.0e-3, 0.5, False)[0] for ob in obl)
    for _ in range(3):
        assert O.multi_step(obl, dt)[0] == 0
    finals = []
    for batch in (False, True):
        blks, sim = device(batch)
        sim.cfg.fixed_time_step = True
        for _ in range(3):
            sim.state.dt_global = dt
            sim.gasdynamic_explicit_increment_with_fixed_grid()
        finals.append([b.download() for b in blks])
        for b in blks:
            b.destroy()
    for k, (a, b, ob) in enumerate(zip(finals[0], finals[1], obl)):
        assert np.array_equal(a, b), k
        err = rel_err(b, ob.get_U())
        print("  block %d rel err %s" % (k, err), flush=True)
        if precision == "mixed":
            assert err[0] <= 1.0e-5
        else:
            assert err.max() <= 1.0e-9
