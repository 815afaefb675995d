"""The flux calculators that read the shock indicator -- hanel, and
adaptive_hanel_ausmdv (the reference's default: Hanel at shock points with
small shear, AUSMDV elsewhere, the detector run after the ghost-cell fill of
every stage) -- on the device against the oracle, with the bars of the other
parity tests: bit-identical in fp64_exact, 1e-12-level in fp64, density
<= 1e-5 in mixed; mixtures within 1e-10."""
import numpy as np
import pytest

import pyoracle as O
import eilmer4_amd as e4
from eilmer4_amd import workloads
from parity_helpers import run_pair, rel_err

pytestmark = pytest.mark.gpu

SHOCK_FLUXES = ["hanel", "adaptive_hanel_ausmdv"]


def check(precision, Ud, Uo, tol):
    err = rel_err(Ud, Uo)
    bitwise = bool(np.array_equal(Ud, Uo))
    print("  %s: rel err per quantity %s bitwise %s" % (precision, err, bitwise), flush=True)
    assert np.all(np.isfinite(Ud))
    if precision == "fp64_exact":
        assert bitwise
    elif precision == "mixed":
        assert err[0] <= 1.0e-5
    else:
        assert err.max() <= tol


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64", "mixed"])
@pytest.mark.parametrize("flux", SHOCK_FLUXES)
def test_shocked_box(gpu, flux, precision):
    """The 3D shocked box (Mach-3 normal shock plus random modes), 5 steps on
    the oracle's dt sequence."""
    geom, fields, bcs, _ = workloads.shocked_box_3d(e4.IdealGas(), n=24)
    print("case shock box n=24 flux=%s precision=%s steps=5" % (flux, precision), flush=True)
    Uo, Ud, _, _, _ = run_pair(geom, fields, bcs, 5, flux=flux, precision=precision, same_dt=True)
    check(precision, Ud, Uo, 1.0e-11)
    if flux == "adaptive_hanel_ausmdv" and precision == "fp64_exact":
        # the shock is detected: the result is not plain AUSMDV's
        _, Ua, _, _, _ = run_pair(geom, fields, bcs, 5, flux="ausmdv", precision=precision, same_dt=True)
        assert not np.array_equal(Ud, Ua)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64"])
def test_planar_bump_2d_adaptive(gpu, precision):
    """The 2D Mach-2 bump workload with the reference's default calculator,
    50 steps, each side with its own dt policy (inflow, outflow, slip walls)."""
    geom, fields, bcs, _ = workloads.planar_bump_2d(e4.IdealGas())
    print("case bump2d adaptive steps=50 precision=%s" % precision, flush=True)
    Uo, Ud, dto, dtd, _ = run_pair(geom, fields, bcs, 50, flux="adaptive_hanel_ausmdv", precision=precision)
    check(precision, Ud, Uo, 1.0e-10)


def _state(gas, shape, seed):
    rng = np.random.default_rng(seed)
    p = 1.0e5 * (1.0 + 0.4 * rng.random(shape))
    T = 300.0 * (1.0 + 0.4 * rng.random(shape))
    v = [300.0 * (rng.random(shape) - 0.5) for _ in range(3)]
    rho, u, _ = gas.rho_u_a_from_pT(p.ravel(), T.ravel())
    return gas.encode(rho, u, *[x.ravel() for x in v])


@pytest.mark.parametrize("orient", [0, 1])
def test_connected_blocks_adaptive(gpu, orient):
    """Two connected blocks with compressive random velocities: the ghost
    cells across the connection carry the other block's previous indicator."""
    gas = e4.IdealGas()
    n = 10
    UA, UB = _state(gas, (n, n, n), 21), _state(gas, (n, n, n), 22)
    fc = int(e4.FluxCalculator.adaptive_hanel_ausmdv)
    oa, ob = (O.Block(O.Gas.ideal(), 3, n, n, n, 0.1, 0.1, 0.1) for _ in range(2))
    for b, U in ((oa, UA), (ob, UB)):
        b.set_config(flux_calc=fc)
        b.set_U(U)
    O.connect_blocks(oa, O.EAST, ob, O.NORTH, orient)
    da, db = (e4.SFluidBlock(gas, 3, n, n, n, 0.1, 0.1, 0.1) for _ in range(2))
    e4.connect_blocks(da, e4.Face.east, db, e4.Face.north, orient)
    sim = e4.Simulation([da, db], e4.GlobalConfig(flux_calculator=fc, precision_mode=e4.PrecisionMode.fp64_exact))
    da.upload(UA)
    db.upload(UB)
    print("case connected adaptive orient=%d" % orient, flush=True)
    for _ in range(3):
        assert O.multi_step([oa, ob], 1.0e-5)[0] == 0
        sim.state.dt_global = 1.0e-5
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    for o_, d_ in ((oa, da), (ob, db)):
        Uo, Ud = o_.get_U(), d_.download()
        print("  rel err", rel_err(Ud, Uo), flush=True)
        assert np.array_equal(Ud, Uo)
    da.destroy()
    db.destroy()


def test_batched_adaptive_matches_per_block(gpu):
    """Batched and per-block launches with the adaptive calculator over a
    small multi-block set: bit-identical."""
    ig, tg = e4.IdealGas(), e4.ThermallyPerfectGas.from_json()
    descs, conns = workloads.multiblock_c5(ig, tg, n_ideal=6, n_tp=2, target_cells=1.5e5)
    res = {}
    fc = e4.FluxCalculator.adaptive_hanel_ausmdv
    for batch in (True, False):
        blks = []
        for d in descs:
            b = e4.SFluidBlock(ig if d["cls"] == "ideal" else tg, 3, d["ni"], d["nj"], d["nk"], d["dx"], d["dy"],
                               d["dz"])
            for face, (kind, fs) in d["bcs"].items():
                b.set_bc(face, kind, fs)
            blks.append(b)
        for (a, fa, b, fb, o) in conns:
            e4.connect_blocks(blks[a], fa, blks[b], fb, o)
        sim = e4.Simulation(blks, e4.GlobalConfig(flux_calculator=fc, precision_mode=e4.PrecisionMode.fp64_exact,
                                                  batch_blocks=batch))
        for b, d in zip(blks, descs):
            b.upload(d["U0"], d["Tguess"])
        for _ in range(3):
            sim.state.dt_global = 1.0e-7
            sim.gasdynamic_explicit_increment_with_fixed_grid()
        res[batch] = [b.download() for b in blks]
        for b in blks:
            b.destroy()
    print("case batched adaptive %d blocks" % len(descs), flush=True)
    for Ua, Ub in zip(res[True], res[False]):
        assert np.array_equal(Ua, Ub)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64"])
def test_thermally_perfect_adaptive(gpu, precision):
    """Five-species air (T 300-8000 K) with the adaptive calculator, 4 steps."""
    gas = e4.ThermallyPerfectGas.from_json()
    n = 12
    geom, fields, bcs, _ = workloads.thermally_perfect_box_3d(gas, n=n)
    U0, Tg = workloads.initial_U_tp(gas, fields)
    fc = int(e4.FluxCalculator.adaptive_hanel_ausmdv)
    og = O.Gas.from_json("therm-perf-5-species-air.json")
    ob = O.Block(og, 3, n, n, n, geom["dx"], geom["dy"], geom["dz"])
    ob.set_config(flux_calc=fc)
    for face, (kind, fs) in bcs.items():
        ob.set_bc(int(face), int(kind), fs)
    ob.set_U(U0, Tg)
    db = e4.SFluidBlock(gas, 3, n, n, n, geom["dx"], geom["dy"], geom["dz"])
    for face, (kind, fs) in bcs.items():
        db.set_bc(face, kind, fs)
    sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=fc, precision_mode=e4.PrecisionMode[precision]))
    db.upload(U0, Tg)
    print("case tp adaptive n=%d precision=%s steps=4" % (n, precision), flush=True)
    dt, _, _ = ob.determine_dt(1.0e-3, 0.5, False)
    for _ in range(4):
        assert ob.step(dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    err = rel_err(db.download(), ob.get_U())
    print("  U rel err", err, flush=True)
    assert err.max() <= 1.0e-10
    db.destroy()
