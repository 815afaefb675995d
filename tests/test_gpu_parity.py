"""Device path vs the CPU oracle on identical inputs (through the C ABI).

Bars, per conserved quantity q, as max|dU_q| / max|U_q|:
  fp64_exact  bit-identical to the oracle (reference roundings)
  fp64        <= 1e-12 after one step, <= 1e-9 after 100 steps
  mixed       density L-inf relative error <= 1e-5 (FP32 reconstruction and flux)
"""
import numpy as np
import pytest

import eilmer4_amd as e4
from eilmer4_amd import workloads
from parity_helpers import run_pair, rel_err, device_block, device_sim

pytestmark = pytest.mark.gpu

TOL_1 = 1.0e-12
TOL_100 = 1.0e-9
TOL_MIXED_RHO = 1.0e-5


def small_box(n=24, seed=workloads.SEED_BOX_3D):
    return workloads.shocked_box_3d(e4.IdealGas(), n=n, seed=seed)


def check(precision, Ud, Uo, tol):
    err = rel_err(Ud, Uo)
    bitwise = bool(np.array_equal(Ud, Uo))
    print("  %s: rel err per quantity %s bitwise %s" % (precision, err, bitwise), flush=True)
    assert np.all(np.isfinite(Ud))
    if precision == "fp64_exact":
        assert bitwise
    elif precision == "mixed":
        assert err[0] <= TOL_MIXED_RHO
    else:
        assert err.max() <= tol
    return err


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64", "mixed"])
@pytest.mark.parametrize("flux", ["ausmdv", "hllc"])
def test_box3d_one_step(gpu, flux, precision):
    geom, fields, bcs, _ = small_box(24)
    print("case box3d n=24 flux=%s precision=%s steps=1" % (flux, precision), flush=True)
    Uo, Ud, dto, dtd, _ = run_pair(geom, fields, bcs, 1, flux=flux, precision=precision)
    assert dto[0] == dtd[0]
    check(precision, Ud, Uo, TOL_1)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64", "mixed"])
@pytest.mark.parametrize("flux", ["ausmdv", "hllc"])
def test_box3d_100_steps(gpu, flux, precision):
    geom, fields, bcs, _ = small_box(20)
    print("case box3d n=20 flux=%s precision=%s steps=100" % (flux, precision), flush=True)
    Uo, Ud, dto, dtd, cps = run_pair(geom, fields, bcs, 100, flux=flux, precision=precision, same_dt=True)
    check(precision, Ud, Uo, TOL_100)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64"])
def test_planar_bump_2d_200_steps(gpu, precision):
    """The 2D 128x64 Mach-2 bump workload, 200 steps, ausmdv, pc, CFL 0.5,
    each side with its own dt policy."""
    gas = e4.IdealGas()
    geom, fields, bcs, _ = workloads.planar_bump_2d(gas)
    print("case bump2d 128x64 steps=200 precision=%s" % precision, flush=True)
    Uo, Ud, dto, dtd, cps = run_pair(geom, fields, bcs, 200, check_every=1, precision=precision)
    check(precision, cps[0][2], cps[0][1], TOL_1)
    check(precision, Ud, Uo, TOL_100)
    assert np.max(np.abs(dto - dtd) / dto) <= TOL_100


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64"])
@pytest.mark.parametrize("scheme", ["euler", "midpoint", "classic_rk3", "tvd_rk3", "denman_rk3"])
def test_update_schemes(gpu, scheme, precision):
    geom, fields, bcs, _ = small_box(16, seed=7)
    print("case box3d n=16 scheme=%s precision=%s steps=3" % (scheme, precision), flush=True)
    Uo, Ud, dto, dtd, _ = run_pair(geom, fields, bcs, 3, scheme=scheme, cfl=0.3, precision=precision)
    check(precision, Ud, Uo, TOL_1 * 10)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64"])
@pytest.mark.parametrize("order,limiter,clipping", [(1, True, True), (2, False, True), (2, True, False)])
def test_reconstruction_options(gpu, order, limiter, clipping, precision):
    geom, fields, bcs, _ = small_box(16, seed=11)
    print("case options order=%d limiter=%d clipping=%d precision=%s" % (order, limiter, clipping, precision),
          flush=True)
    Uo, Ud, _, _, _ = run_pair(geom, fields, bcs, 2, order=order, limiter=limiter, clipping=clipping, cfl=0.3,
                               precision=precision)
    check(precision, Ud, Uo, TOL_1 * 10)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64", "mixed"])
def test_ragged_sizes_and_all_bcs(gpu, precision):
    """Block sizes that are not multiples of the tile, and every face kind on
    one block: inflow W, extrapolate E, slip walls N/S, periodic T/B."""
    gas = e4.IdealGas()
    ni, nj, nk = 37, 13, 11
    rng = np.random.default_rng(3)
    shape = (nk, nj, ni)
    p = 1.0e5 * (1.0 + 0.2 * rng.random(shape))
    T = 300.0 * (1.0 + 0.2 * rng.random(shape))
    vx = 600.0 + 50.0 * rng.standard_normal(shape)
    vy = 50.0 * rng.standard_normal(shape)
    vz = 50.0 * rng.standard_normal(shape)
    fields = dict(p=p, T=T, vx=vx, vy=vy, vz=vz)
    geom = dict(dims=3, ni=ni, nj=nj, nk=nk, dx=0.01, dy=0.013, dz=0.007)
    inflow = gas.flowstate(1.0e5, 300.0, (700.0, 10.0, -5.0))
    bcs = {e4.Face.west: (e4.BCType.inflow, inflow), e4.Face.east: (e4.BCType.extrapolate, None),
           e4.Face.north: (e4.BCType.slip_wall, None), e4.Face.south: (e4.BCType.slip_wall, None),
           e4.Face.top: (e4.BCType.periodic, None), e4.Face.bottom: (e4.BCType.periodic, None)}
    print("case ragged 37x13x11 all-bcs precision=%s" % precision, flush=True)
    for flux in ("ausmdv", "hllc"):
        Uo, Ud, dto, dtd, _ = run_pair(geom, fields, bcs, 5, flux=flux, precision=precision, same_dt=True)
        check(precision, Ud, Uo, TOL_1 * 10)


def test_sod_shock_tube_on_device(gpu):
    """The reference's 3D Sod integration test (examples/eilmer/3D/sod-shock-tube/
    sg/sod.lua, sod-test.rb:20-98) run on the device: 75 +- 3 steps to 0.6 ms
    and the probe values within 1 %."""
    gas = e4.IdealGas()
    ni, nj, nk = 100, 2, 2
    b = e4.SFluidBlock(gas, 3, ni, nj, nk, 1.0 / ni, 0.1 / nj, 0.1 / nk)
    x = (np.arange(ni) + 0.5) / ni
    p = np.broadcast_to(np.where(x < 0.5, 1.0e5, 1.0e4), (nk, nj, ni))
    T = np.broadcast_to(np.where(x < 0.5, 348.4, 278.8), (nk, nj, ni))
    b.set_from_pT(p, T, 0.0, 0.0, 0.0)
    sim = e4.Simulation([b], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,   # sod.lua
                                             max_time=0.6e-3, max_step=600))
    steps = sim.integrate_in_time()
    U, fs = b.download(with_fs=True)
    fs = fs.reshape(8, nk, nj, ni)
    print("sod steps", steps)
    assert abs(steps - 75) < 3
    for xp, ref in ((0.78, dict(rho=0.2647, p=30.2e3, T=398.0, u=293.0)),
                    (0.6, dict(rho=0.4271, p=30.2e3, T=247.0, u=293.0))):
        i = int(xp * ni)
        got = dict(rho=fs[0, 0, 0, i], p=fs[1, 0, 0, i], T=fs[2, 0, 0, i], u=fs[5, 0, 0, i])
        print(xp, got)
        for key in ref:
            assert abs(got[key] - ref[key]) / ref[key] < 1.0e-2
    b.destroy()


def test_upload_rejects_invalid_state(gpu):
    gas = e4.IdealGas()
    b = e4.SFluidBlock(gas, 3, 8, 8, 8, 0.1, 0.1, 0.1)
    U = workloads.initial_U(gas, dict(p=np.full((8, 8, 8), 1e5), T=np.full((8, 8, 8), 300.0),
                                      vx=np.zeros((8, 8, 8)), vy=np.zeros((8, 8, 8)), vz=np.zeros((8, 8, 8))))
    U[0, 17] = -1.0
    with pytest.raises(e4.E4hipError) as ei:
        b.upload(U)
    assert ei.value.code == -5
    b.destroy()


def test_too_many_bad_cells_reported(gpu):
    """A time step far beyond the CFL limit drives cells invalid; the library
    reports it as the reference does (FlowSolverException), without faulting."""
    geom, fields, bcs, _ = small_box(12, seed=5)
    gas = e4.IdealGas()
    U0 = workloads.initial_U(gas, fields)
    db = device_block(gas, geom, bcs, U0)
    sim = device_sim([db])
    sim.state.dt_global = 1.0e-2     # ~ 1000x the stable step
    sim.cfg.fixed_time_step = True
    with pytest.raises(e4.FlowSolverException):
        sim.integrate_in_time(target_time=1.0, max_step=3)
    db.destroy()
