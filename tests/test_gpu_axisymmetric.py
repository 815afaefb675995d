"""2D axisymmetric blocks (x the axis, y the radius) on the device against the
oracle: per-radian volumes and face areas from the reference's 2D cell and
face geometry, and the pressure source of the radial momentum.  Bars:
fp64_exact bit-identical for the ideal gas, fp64 1e-11, mixed density 1e-5;
mixtures 1e-10."""
import numpy as np
import pytest

import pyoracle as O
import eilmer4_amd as e4
from eilmer4_amd.fvcore import BCType
from parity_helpers import oracle_block, device_block, device_sim, rel_err

pytestmark = pytest.mark.gpu


def _field_2d(gas, ni, nj, seed):
    rng = np.random.default_rng(seed)
    shape = (1, nj, ni)
    p = 1.0e5 * (1.0 + 0.3 * rng.random(shape))
    T = 300.0 * (1.0 + 0.3 * rng.random(shape))
    rho, u, _ = gas.rho_u_a_from_pT(p.ravel(), T.ravel())
    vx = 600.0 + 200.0 * (rng.random(shape) - 0.5)
    vy = 200.0 * (rng.random(shape) - 0.5)
    return gas.encode(rho, u, vx.ravel(), vy.ravel(), np.zeros(rho.size))


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64", "mixed"])
@pytest.mark.parametrize("flux", ["ausmdv", "adaptive_hanel_ausmdv"])
@pytest.mark.parametrize("y0", [0.0, 0.2])
def test_axisymmetric_ideal(gpu, y0, flux, precision):
    gas = e4.IdealGas()
    ni, nj = 24, 16
    geom = dict(dims=2, ni=ni, nj=nj, nk=1, dx=1.0 / ni, dy=0.5 / nj, dz=1.0)
    rho, u, a = gas.rho_u_a_from_pT(np.array([1.0e5]), np.array([300.0]))
    fs = np.array([rho[0], 1.0e5, 300.0, u[0], a[0], 700.0, 0.0, 0.0])
    # the axis (y0 = 0) is a slip wall, as in the reference's axisymmetric cases
    bcs = {O.WEST: (BCType.inflow, fs), O.EAST: (BCType.outflow_simple_flux, None),
           O.NORTH: (BCType.slip_wall_1, None), O.SOUTH: (BCType.slip_wall, None)}
    U0 = _field_2d(gas, ni, nj, 61)
    fc = e4.flux_calculator_from_name(flux)
    ob = oracle_block(geom, bcs, U0, flux=int(fc))
    ob.set_axisymmetric(0.0, y0)
    db = device_block(gas, geom, bcs, U0)
    db.set_axisymmetric(0.0, y0)
    sim = device_sim([db], flux=fc, precision=precision)
    osim = O.Simulation(ob)
    print("case axisymmetric 2D %dx%d y0=%g flux=%s precision=%s steps=10" % (ni, nj, y0, flux, precision),
          flush=True)
    for s in range(10):
        osim.run(max_step=s + 1)
        sim.state.dt_global = osim.dt_global
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    Uo, Ud = ob.get_U(), db.download()
    db.destroy()
    err = rel_err(Ud, Uo)
    print("  rel err %s bitwise %s" % (err, np.array_equal(Ud, Uo)), flush=True)
    assert np.all(np.isfinite(Ud))
    if precision == "fp64_exact":
        assert err.max() <= 1e-13          # the wall flux of the north face uses pow()
    elif precision == "mixed":
        assert err[0] <= 1.0e-5
    else:
        assert err.max() <= 1.0e-11


def test_axisymmetric_ideal_bitwise_without_wall_flux(gpu):
    """With ghost-cell boundaries only, exact FP64 is bit-identical."""
    gas = e4.IdealGas()
    ni, nj = 20, 12
    geom = dict(dims=2, ni=ni, nj=nj, nk=1, dx=1.0 / ni, dy=0.5 / nj, dz=1.0)
    bcs = {O.WEST: (BCType.extrapolate, None), O.EAST: (BCType.extrapolate_1, None),
           O.NORTH: (BCType.slip_wall, None), O.SOUTH: (BCType.slip_wall, None)}
    U0 = _field_2d(gas, ni, nj, 62)
    ob = oracle_block(geom, bcs, U0)
    ob.set_axisymmetric(0.0, 0.0)
    db = device_block(gas, geom, bcs, U0)
    db.set_axisymmetric(0.0, 0.0)
    sim = device_sim([db], precision="fp64_exact")
    osim = O.Simulation(ob)
    for s in range(10):
        osim.run(max_step=s + 1)
        sim.state.dt_global = osim.dt_global
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    assert np.array_equal(db.download(), ob.get_U())
    db.destroy()


def test_axisymmetric_thermally_perfect(gpu):
    gas = e4.ThermallyPerfectGas.from_json()
    og = O.Gas.from_json("therm-perf-5-species-air.json")
    ni, nj = 16, 12
    rng = np.random.default_rng(63)
    shape = (1, nj, ni)
    Y = np.array([0.7, 0.2, 0.05, 0.03, 0.02])
    Yf = np.broadcast_to(Y[:, None], (5, ni * nj)).copy()
    T = 1500.0 * (1.0 + 0.2 * rng.random(shape))
    p = 1.0e5 * (1.0 + 0.3 * rng.random(shape))
    rho = (p.ravel() / (gas.Rmix(Yf) * T.ravel())).reshape(shape)
    from eilmer4_amd import workloads
    fields = dict(p=p, T=T, rho=rho, vx=400.0 + 100.0 * rng.random(shape), vy=100.0 * (rng.random(shape) - 0.5),
                  vz=np.zeros(shape), Y=Yf)
    U0, Tg = workloads.initial_U_tp(gas, fields)
    bcs = {e4.Face.west: BCType.extrapolate, e4.Face.east: BCType.outflow_simple_flux,
           e4.Face.north: BCType.slip_wall, e4.Face.south: BCType.slip_wall}
    dx, dy = 1.0 / ni, 0.5 / nj
    ob = O.Block(og, 2, ni, nj, 1, dx, dy)
    ob.set_config(flux_calc=O.AUSMDV)
    db = e4.SFluidBlock(gas, 2, ni, nj, 1, dx, dy, 1.0)
    for face, kind in bcs.items():
        ob.set_bc(int(face), int(kind), None)
        db.set_bc(face, kind, None)
    ob.set_axisymmetric(0.0, 0.05)
    db.set_axisymmetric(0.0, 0.05)
    ob.set_U(U0, Tg)
    sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                              precision_mode=e4.PrecisionMode.fp64_exact))
    db.upload(U0, Tg)
    print("case tp axisymmetric 2D %dx%d" % (ni, nj), flush=True)
    dt, _, _ = ob.determine_dt(1.0e-3, 0.5, False)
    for _ in range(4):
        assert ob.step(dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    err = rel_err(db.download(), ob.get_U())
    print("  U rel err per quantity", err, flush=True)
    assert err.max() <= 1.0e-10
    db.destroy()
