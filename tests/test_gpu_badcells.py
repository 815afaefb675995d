"""Invalid cells and the low-temperature decode fall-back against the oracle.

The reference counts invalid cells (count_invalid_cells, fluidblock.d:433-499
with FlowState.check_data, flowstate.d:345-388) after every stage and stops
the step with "Too many bad cells" when a block's count exceeds
max_invalid_cells (simcore.d:2480-2493, 2694-2696, 2897-2899); a decode that
fails (non-positive density) stops it at once; U[0] keeps the state before
the step because the final swap (simcore.d:2936-2941) never happens.  A decode
whose internal energy is not positive re-encodes the cell at T = 200 K
(fvcell.d:685-697) and the step goes on.

The cases are a cold two-stream expansion (T0 a little above the 1.01 K
floor, the halves moving apart at V), where the time step decides what
happens: which stage exceeds the limit, whether a stage-1 fall-back is
followed by a stage-2 decode failure, or whether the step ends normally with
eight re-encoded cells.  Device and oracle run the same U0 and dt in
fp64_exact; per-stage counts, the status and U after the step must agree
exactly.
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle as O
import eilmer4_amd as e4
from eilmer4_amd import workloads
from parity_helpers import oracle_block, device_block, device_sim

pytestmark = pytest.mark.gpu


def expansion(T0, V, ni=32, nj=4, p0=100.0):
    shape = (1, nj, ni)
    x = (np.arange(ni) + 0.5) / ni
    vx = np.broadcast_to(np.where(x < 0.5, -V, V), shape)
    fields = dict(p=np.full(shape, p0), T=np.full(shape, T0), vx=vx, vy=np.zeros(shape), vz=np.zeros(shape))
    geom = dict(dims=2, ni=ni, nj=nj, nk=1, dx=1.0 / ni, dy=0.25 / nj, dz=1.0)
    bcs = {e4.Face.west: (e4.BCType.extrapolate, None), e4.Face.east: (e4.BCType.extrapolate, None),
           e4.Face.north: (e4.BCType.slip_wall, None), e4.Face.south: (e4.BCType.slip_wall, None)}
    return geom, fields, bcs


def device_step(blocks, dt):
    """One e4hip_explicit_increment through the C ABI: (status, per-stage counts)."""
    L = e4.load()
    ids = (C.c_int * len(blocks))(*[b.id for b in blocks])
    nbad = (C.c_int * 3)()
    rc = L.e4hip_explicit_increment(ids, len(blocks), dt, nbad)
    return rc, list(nbad)


# oracle status -> device status: -1 decode failure -> E4HIP_ERR_DECODE, -2 too many -> E4HIP_ERR_BAD_CELLS
STATUS = {0: 0, -1: -5, -2: -4}

CASES = [
    # (T0, V, dt multiple, max_invalid_cells, scheme, what the oracle does)
    (1.05, 10.0, 2.5, 0, "pc", "stage 1 over the limit"),
    (1.05, 10.0, 2.5, 8, "pc", "stage 1 within the limit, stage 2 over it"),
    (2.0, 30.0, 3.5, 0, "pc", "stage-1 low-T fall-back, stage-2 decode failure"),
    (1.05, 20.0, 3.0, 0, "pc", "low-T fall-back, the step completes"),
    (1.05, 20.0, 3.5, 0, "euler", "low-T fall-back, euler"),
    (1.05, 10.0, 3.0, 0, "euler", "euler, stage 1 over the limit"),
    (1.05, 20.0, 1.0, 0, "midpoint", "no bad cells"),
]


@pytest.mark.parametrize("T0,V,m,maxinv,scheme,what", CASES)
def test_bad_cells_per_stage_vs_oracle(gpu, T0, V, m, maxinv, scheme, what):
    print("case badcells T0=%g V=%g m=%g max=%d scheme=%s (%s)" % (T0, V, m, maxinv, scheme, what), flush=True)
    gas = e4.IdealGas()
    geom, fields, bcs = expansion(T0, V)
    U0 = workloads.initial_U(gas, fields)
    us = e4.update_scheme_from_name(scheme)
    ob = oracle_block(geom, bcs, U0, flux=O.AUSMDV, scheme=int(us), max_invalid_cells=maxinv)
    dt = ob.determine_dt(1e-3, 0.5, False)[0] * m
    rc_o, nbad_o = ob.step(dt)
    db = device_block(gas, geom, bcs, U0)
    device_sim([db], flux=e4.FluxCalculator.ausmdv, scheme=us, max_invalid_cells=maxinv, precision="fp64_exact")
    rc_d, nbad_d = device_step([db], dt)
    Ud = db.download()
    print("  oracle rc %d counts %s; device rc %d counts %s" % (rc_o, nbad_o, rc_d, nbad_d), flush=True)
    assert rc_d == STATUS[rc_o]
    assert nbad_d == nbad_o
    if rc_o == 0:
        assert np.array_equal(Ud, ob.get_U())
        if "fall-back" in what:
            # the re-encoded cells: T = 200 K exactly in the oracle's flow state
            fs_o = ob.get_fs()
            assert int(np.sum(fs_o[2] == 200.0)) == 8
    else:
        # the step stopped: U[0] is the state before it, bit for bit
        assert np.array_equal(Ud, U0)
        assert np.array_equal(ob.get_U(0), U0)
    db.destroy()


def test_stopped_step_keeps_u0_of_batched_blocks(gpu):
    """Four batched 2D blocks (one of them driven invalid): the step stops
    after stage 1 and every block keeps its U0; the next step with a sound
    time step then runs from those states and matches a fresh run."""
    print("case badcells: batched blocks, stop then continue", flush=True)
    gas = e4.IdealGas()
    geom, fields, bcs = expansion(1.05, 10.0)
    U0 = workloads.initial_U(gas, fields)
    runs = []
    for stop_first in (True, False):
        blocks = [device_block(gas, geom, bcs, U0) for _ in range(4)]
        sim = device_sim(blocks, flux=e4.FluxCalculator.ausmdv, precision="fp64")
        ob = oracle_block(geom, bcs, U0, flux=O.AUSMDV)
        dt_ok = ob.determine_dt(1e-3, 0.5, False)[0]
        if stop_first:
            rc, nbad = device_step(blocks, 2.5 * dt_ok)
            assert rc == -4 and nbad[0] > 0 and nbad[1] == 0 and nbad[2] == 0, (rc, nbad)
            for b in blocks:
                assert np.array_equal(b.download(), U0)
        rc, nbad = device_step(blocks, dt_ok)
        assert rc == 0 and nbad == [0, 0, 0]
        runs.append([b.download() for b in blocks])
        for b in blocks:
            b.destroy()
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64", "mixed"])
def test_stopped_step_keeps_mixture_state(gpu, precision):
    """A thermally perfect block driven invalid by a far too large step: the
    status is an error and both U[0] and the flow state (the primitive array,
    T included) are those before the step."""
    print("case badcells: mixture block, %s" % precision, flush=True)
    gas = e4.ThermallyPerfectGas.from_json()
    n = 8
    geom, fields, bcs, _ = workloads.thermally_perfect_box_3d(gas, n=n)
    U0, Tg = workloads.initial_U_tp(gas, fields)
    db = e4.SFluidBlock(gas, 3, n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
    for face, (kind, fs) in bcs.items():
        db.set_bc(face, kind, fs)
    sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                              precision_mode=e4.PrecisionMode[precision]))
    db.upload(U0, Tg)
    Ub, fsb = db.download(with_fs=True)
    dt = sim.block_dt(1e-3, False)[0]
    rc, nbad = device_step([db], 200.0 * dt)
    print("  rc %d counts %s" % (rc, nbad), flush=True)
    assert rc in (-4, -5)
    Ua, fsa = db.download(with_fs=True)
    assert np.array_equal(Ua, Ub) and np.array_equal(fsa, fsb)
    rc, nbad = device_step([db], 0.5 * dt)
    assert rc == 0 and nbad == [0, 0, 0]
    db.destroy()


def test_bad_cells_per_block_vs_oracle(gpu):
    """Three unconnected blocks in one step (a benign one, two cold expansions
    of different heights): the step stops after stage 1 and the per-block
    counts (e4hip_last_step_bad_cells) equal each block's own count in the
    oracle; the later stages report 0, and the Simulation raises the
    reference's exception."""
    print("case badcells: per-block counts, 3 blocks", flush=True)
    gas = e4.IdealGas()
    cases = [expansion(300.0, 10.0), expansion(1.05, 10.0), expansion(1.05, 10.0, nj=8)]
    U0s = [workloads.initial_U(gas, f) for _, f, _ in cases]
    obs = [oracle_block(g, b, U0, flux=O.AUSMDV) for (g, _, b), U0 in zip(cases, U0s)]
    dt = 2.5 * obs[1].determine_dt(1e-3, 0.5, False)[0]
    expect = []
    for ob in obs:
        rc, nbad = ob.step(dt)
        expect.append(nbad[0])
    print("  oracle stage-1 counts per block %s" % expect, flush=True)
    assert expect[0] == 0 and expect[1] > 0 and expect[2] > expect[1]
    blocks = [device_block(gas, g, b, U0) for (g, _, b), U0 in zip(cases, U0s)]
    sim = device_sim(blocks, flux=e4.FluxCalculator.ausmdv, precision="fp64_exact")
    sim.cfg.fixed_time_step = True
    sim.state.dt_global = dt
    with pytest.raises(e4.FlowSolverException):
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    got = sim.last_bad_cells_per_block
    print("  device per block %s" % got, flush=True)
    assert [c[0] for c in got] == expect
    assert all(c[1] == 0 and c[2] == 0 for c in got)
    assert sim.last_bad_cells == [sum(expect), 0, 0]
    L = e4.load()
    wrong = (C.c_int * 6)()
    assert L.e4hip_last_step_bad_cells(2, wrong) == -1      # E4HIP_ERR_ARG: not the last call's block count
    for b, U0 in zip(blocks, U0s):
        assert np.array_equal(b.download(), U0)
        b.destroy()
