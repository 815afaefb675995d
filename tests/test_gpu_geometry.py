"""Blocks built from their grid on the device (e4hip_block_create_geometry,
gen_kernels.h) against the oracle's restatement of
compute_primary_cell_geometric_data and the general branch of
convective_flux_phase0 (sfluidblock.d:650-842, 2189-2345):

  * curvilinear 3D and 2D blocks (warped faces, stretched and skewed cells)
    with inflow, simple-outflow-flux, slip-wall and extrapolated faces, for
    ausmdv, hllc and adaptive_hanel_ausmdv: bit-identical to the oracle in
    fp64_exact, within 1e-12 per conserved quantity in fp64, density within
    1e-5 in the mixed form;
  * two curvilinear blocks joined by a full-face connection (ghost lengths
    from the neighbour's cells) and a periodic block: bit-identical;
  * a five-species mixture on a curvilinear grid: within 1e-10;
  * the CFL signal: dt_allow equal to the oracle's;
  * the reference's 3D ramp case (ramp-test.rb:21-53) on the device:
    862 +- 3 steps, shock angle 57 +- 1 degrees, and the final state
    bit-identical to the oracle's in fp64_exact.
"""
import numpy as np
import pytest

import pyoracle as O
import eilmer4_amd as e4
import ramp_case as R
from eilmer4_amd.fvcore import BCType, Face
from parity_helpers import rel_err
from test_oracle_geometry import curvy_grid, _flow

pytestmark = pytest.mark.gpu

RGAS = 8.31451 / 0.02896
OFACE = {Face.north: O.NORTH, Face.east: O.EAST, Face.south: O.SOUTH, Face.west: O.WEST, Face.top: O.TOP,
         Face.bottom: O.BOTTOM}


def _inflow(gas):
    p, T, vel = 9.0e4, 320.0, (420.0, 35.0, -20.0)
    rho, u, a = gas.rho_u_a_from_pT(np.array([p]), np.array([T]))
    return [rho[0], p, T, u[0], a[0], vel[0], vel[1], vel[2]]


def _mixed_bcs(dims):
    bcs = {Face.west: BCType.inflow, Face.east: BCType.outflow_simple_flux, Face.north: BCType.slip_wall,
           Face.south: BCType.extrapolate}
    if dims == 3:
        bcs[Face.top] = BCType.slip_wall
        bcs[Face.bottom] = BCType.extrapolate
    return bcs


def _oracle_from_grid(g, U0, bcs, fs_in, flux, scheme=O.PC, gas=None, Tg=None):
    b = O.Block.from_grid(gas or O.Gas.ideal(), g)
    b.set_config(flux_calc=int(flux), update_scheme=int(scheme))
    for f, kind in bcs.items():
        b.set_bc(OFACE[f], int(kind), fs_in if kind == BCType.inflow else None)
    if Tg is None:
        b.set_U(U0)
    else:
        b.set_U(U0, Tg)
    return b


def _device_from_grid(gas, g, U0, bcs, fs_in, Tg=None):
    b = e4.SFluidBlock.from_grid(gas, g)
    for f, kind in bcs.items():
        b.set_bc(f, kind, fs_in if kind == BCType.inflow else None)
    b.upload(U0, Tg)
    return b


def _U0(gas, shape, seed):
    p, T, v = _flow(shape, seed)
    rho, u, _ = gas.rho_u_a_from_pT(p.ravel(), T.ravel())
    return gas.encode(rho, u, *[x.ravel() for x in v])


def _run(dims, flux, precision, steps=3, seed=31):
    gas = e4.IdealGas()
    g = curvy_grid(10, 8, 6, seed=seed)
    shape = (6, 8, 10)
    if dims == 2:
        g = g[0].copy()
        g[..., 2] = 0.0
        shape = (1, 8, 10)
    U0 = _U0(gas, shape, seed + 1)
    bcs = _mixed_bcs(dims)
    fs_in = _inflow(gas)
    fc = e4.flux_calculator_from_name(flux)
    ob = _oracle_from_grid(g, U0, bcs, fs_in, fc)
    db = _device_from_grid(gas, g, U0, bcs, fs_in)
    sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=fc, precision_mode=e4.PrecisionMode[precision]))
    dt = 2.0e-6
    for _ in range(steps):
        assert ob.step(dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    Uo, Ud = ob.get_U(), db.download()
    db.destroy()
    return Uo, Ud


@pytest.mark.parametrize("dims", [3, 2])
@pytest.mark.parametrize("flux", ["ausmdv", "hllc", "adaptive_hanel_ausmdv"])
def test_curvilinear_block_exact(gpu, dims, flux):
    print("case curvilinear %dD %s fp64_exact" % (dims, flux), flush=True)
    Uo, Ud = _run(dims, flux, "fp64_exact")
    assert np.array_equal(Uo, Ud), rel_err(Ud, Uo)


@pytest.mark.parametrize("dims", [3, 2])
def test_curvilinear_block_fp64_and_mixed(gpu, dims):
    print("case curvilinear %dD ausmdv fp64 / mixed" % dims, flush=True)
    Uo, Ud = _run(dims, "ausmdv", "fp64")
    err = rel_err(Ud, Uo)
    print("  fp64 rel err %s" % err, flush=True)
    assert err.max() <= 1.0e-12
    Uo, Um = _run(dims, "ausmdv", "mixed")
    err = rel_err(Um, Uo)
    print("  mixed rel err %s" % err, flush=True)
    assert err[0] <= 1.0e-5


@pytest.mark.parametrize("orient_case", ["east-west", "north-south"])
def test_connected_curvilinear_blocks_exact(gpu, orient_case):
    """Two blocks of one curvilinear grid joined across i = 5 (or j = 4): the
    ghost cells take the neighbour's cells' own lengths."""
    print("case connected curvilinear %s" % orient_case, flush=True)
    gas = e4.IdealGas()
    g = curvy_grid(10, 8, 6, seed=41)
    shape = (6, 8, 10)
    p, T, v = _flow(shape, 42)
    if orient_case == "east-west":
        parts = [(g[:, :, :6], (slice(None), slice(None), slice(0, 5))),
                 (g[:, :, 5:], (slice(None), slice(None), slice(5, None)))]
        fa, fb = Face.east, Face.west
    else:
        parts = [(g[:, :5, :], (slice(None), slice(0, 4), slice(None))),
                 (g[:, 4:, :], (slice(None), slice(4, None), slice(None)))]
        fa, fb = Face.north, Face.south
    oblocks, dblocks = [], []
    fs_in = _inflow(gas)
    for gp, sl in parts:
        rho, u, _ = gas.rho_u_a_from_pT(p[sl].ravel(), T[sl].ravel())
        U0 = gas.encode(rho, u, *[x[sl].ravel() for x in v])
        bcs = {Face.west: BCType.inflow} if (orient_case == "north-south" or gp is parts[0][0]) else {}
        ob = _oracle_from_grid(gp, U0, bcs, fs_in, O.AUSMDV)
        db = _device_from_grid(gas, gp, U0, bcs, fs_in)
        oblocks.append(ob)
        dblocks.append(db)
    O.connect_blocks(oblocks[0], OFACE[fa], oblocks[1], OFACE[fb], 0)
    e4.connect_blocks(dblocks[0], fa, dblocks[1], fb, 0)
    sim = e4.Simulation(dblocks, e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                                 precision_mode=e4.PrecisionMode.fp64_exact))
    dt = 2.0e-6
    for _ in range(3):
        assert O.multi_step(oblocks, dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    for n, (ob, db) in enumerate(zip(oblocks, dblocks)):
        Uo, Ud = ob.get_U(), db.download()
        if not np.array_equal(Uo, Ud):
            d = np.abs(Ud - Uo)
            q, c = np.unravel_index(np.argmax(d / np.abs(Uo).max(axis=1, keepdims=True)), d.shape)
            ni, nj = db.ni, db.nj
            print("block %d: rel err %s; largest at q=%d cell (i,j,k)=(%d,%d,%d), %d cells differ"
                  % (n, rel_err(Ud, Uo), q, c % ni, (c // ni) % nj, c // (ni * nj),
                     int(np.any(d > 0, axis=0).sum())), flush=True)
        assert np.array_equal(Uo, Ud)
    for db in dblocks:
        db.destroy()


def test_periodic_curvilinear_block_exact(gpu):
    print("case periodic curvilinear block", flush=True)
    gas = e4.IdealGas()
    g = curvy_grid(8, 6, 6, seed=51, amp=0.1)
    # make the k-faces periodic-compatible in shape (any geometry is accepted)
    U0 = _U0(gas, (6, 6, 8), 52)
    bcs = {Face.top: BCType.periodic, Face.bottom: BCType.periodic, Face.west: BCType.inflow,
           Face.east: BCType.extrapolate}
    fs_in = _inflow(gas)
    ob = _oracle_from_grid(g, U0, bcs, fs_in, O.HLLC)
    db = _device_from_grid(gas, g, U0, bcs, fs_in)
    sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.hllc,
                                              precision_mode=e4.PrecisionMode.fp64_exact))
    for _ in range(3):
        assert ob.step(2.0e-6)[0] == 0
        sim.state.dt_global = 2.0e-6
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    assert np.array_equal(ob.get_U(), db.download())
    db.destroy()


def test_curvilinear_cfl_matches_oracle(gpu):
    print("case curvilinear CFL", flush=True)
    gas = e4.IdealGas()
    for dims, stringent in ((3, False), (3, True), (2, False)):
        g = curvy_grid(9, 7, 5, seed=61)
        shape = (5, 7, 9)
        if dims == 2:
            g = g[0].copy()
            g[..., 2] = 0.0
            shape = (1, 7, 9)
        U0 = _U0(gas, shape, 62)
        ob = _oracle_from_grid(g, U0, {}, None, O.AUSMDV)
        ob.set_stringent_cfl(stringent)
        db = _device_from_grid(gas, g, U0, {}, None)
        sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                                  precision_mode=e4.PrecisionMode.fp64_exact,
                                                  stringent_cfl=stringent))
        da, cm = sim.block_dt(1.0e-6, False)
        oa = ob.determine_dt(1.0e-6, 0.5, False)
        print("  dims %d stringent %s: device %.17g %.17g oracle %.17g %.17g" % (dims, stringent, da, cm, oa[0], oa[1]),
              flush=True)
        assert da == oa[0] and cm == oa[1]
        db.destroy()


def test_mixture_on_curvilinear_grid(gpu):
    print("case five-species mixture on a curvilinear grid", flush=True)
    gas = e4.ThermallyPerfectGas.from_json()
    og = O.Gas.from_json("therm-perf-5-species-air.json")
    g = curvy_grid(8, 6, 5, seed=71)
    shape = (5, 6, 8)
    rng = np.random.default_rng(72)
    T = 1500.0 + 3000.0 * rng.random(shape)
    p = 5.0e4 * (1.0 + rng.random(shape))
    v = [300.0 * (rng.random(shape) - 0.5) for _ in range(3)]
    Y = np.array([0.7, 0.2, 0.04, 0.03, 0.03])[:, None] * np.ones((1, T.size))
    ob0 = O.Block.from_grid(og, g)
    ob0.set_from_pT(p, T, *v, massf=Y)
    U0, Tg = ob0.get_U(), ob0.get_fs()[2]
    bcs = {Face.west: BCType.extrapolate, Face.east: BCType.outflow_simple_flux, Face.north: BCType.slip_wall}
    for precision, tol in (("fp64_exact", 1.0e-10), ("fp64", 1.0e-10)):
        ob = _oracle_from_grid(g, U0, bcs, None, O.AUSMDV, gas=og, Tg=Tg)
        db = _device_from_grid(gas, g, U0, bcs, None, Tg=Tg)
        sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                                  precision_mode=e4.PrecisionMode[precision]))
        for _ in range(2):
            assert ob.step(1.0e-6)[0] == 0
            sim.state.dt_global = 1.0e-6
            sim.gasdynamic_explicit_increment_with_fixed_grid()
        U, fs = db.download(with_fs=True)
        err = rel_err(U, ob.get_U())
        errT = float(np.max(np.abs(fs[2] - ob.get_fs()[2]) / ob.get_fs()[2]))
        print("  %s: U rel err %s, T rel err %.3e" % (precision, err, errT), flush=True)
        assert err.max() <= tol and errT <= tol
        db.destroy()


def _ramp_device(precision):
    gas = e4.IdealGas()
    g0, g1 = R.grids()
    b0, b1 = e4.SFluidBlock.from_grid(gas, g0), e4.SFluidBlock.from_grid(gas, g1)
    b0.set_bc(Face.west, BCType.inflow, _fs_pT(gas, R.INFLOW["p"], R.INFLOW["T"], R.INFLOW["vel"]))
    b1.set_bc(Face.east, BCType.outflow_simple_flux)
    e4.connect_blocks(b0, Face.east, b1, Face.west, 0)
    cfg = e4.GlobalConfig(flux_calculator=e4.FluxCalculator.adaptive_hanel_ausmdv,
                          gasdynamic_update_scheme=e4.GasdynamicUpdate.euler, cfl_value=0.5,
                          dt_init=R.DT_INIT, dt_max=1.0e-3, max_time=R.MAX_TIME, max_step=1000,
                          precision_mode=e4.PrecisionMode[precision])
    sim = e4.Simulation([b0, b1], cfg)
    U0s = []
    for b, g in ((b0, g0), (b1, g1)):
        ob = O.Block.from_grid(O.Gas.ideal(), g)
        ob.set_from_pT(R.INITIAL["p"], R.INITIAL["T"], 0.0, 0.0, 0.0)
        U0s.append(ob.get_U())
        b.upload(U0s[-1])
    steps = sim.integrate_in_time(target_time=R.MAX_TIME, max_step=1000)
    out = [b.download(with_fs=True) for b in (b0, b1)]
    for b in (b0, b1):
        b.destroy()
    return steps, out, U0s


def _fs_pT(gas, p, T, vel):
    rho, u, a = gas.rho_u_a_from_pT(np.array([p]), np.array([T]))
    return [rho[0], p, T, u[0], a[0], vel[0], vel[1], vel[2]]


def test_ramp_case_on_device(gpu):
    """ramp-test.rb:21-53 on the device: 862 +- 3 steps, shock angle 57 +- 1
    degrees, mean deviation below 2 mm; in fp64_exact the same steps and the
    same final state as the oracle started from the same conserved state."""
    print("case ramp fp64_exact", flush=True)
    steps, ((U0d, fs0), (U1d, fs1)), U0s = _ramp_device("fp64_exact")
    g0, g1 = R.grids()
    p = np.concatenate([fs0[1].reshape(40, 4, 10), fs1[1].reshape(40, 4, 30)], axis=2)
    pos = np.concatenate([R.cell_centres(g0), R.cell_centres(g1)], axis=2)
    angle, dev = R.shock_angle(pos, p)
    print("  device: %d steps, shock angle %.3f deg, deviation %.4f m" % (steps, angle, dev), flush=True)
    assert abs(steps - 862) < 3
    assert abs(angle - 57.0) < 1.0
    assert dev < 0.002
    # the oracle from the same conserved state, with the same dt policy
    gas = O.Gas.ideal()
    ob = [O.Block.from_grid(gas, g) for g in (g0, g1)]
    for b, U in zip(ob, U0s):
        b.set_config(flux_calc=O.ADAPTIVE_HANEL_AUSMDV, update_scheme=O.EULER)
    ob[0].set_bc(O.WEST, O.BC_INFLOW, O.flowstate_pT(gas, R.INFLOW["p"], R.INFLOW["T"], R.INFLOW["vel"]))
    ob[1].set_bc(O.EAST, O.BC_OUTFLOW_SIMPLE_FLUX)
    O.connect_blocks(ob[0], O.EAST, ob[1], O.WEST, 0)
    for b, U in zip(ob, U0s):
        b.set_U(U)
    dt, t, step = R.DT_INIT, 0.0, 0
    while t < R.MAX_TIME and step < 1000:
        if step % 10 == 0:
            da = min(b.determine_dt(dt, 0.5, step > 0)[0] for b in ob)
            if step == 0:
                da = min(R.DT_INIT, da)
            dt = da if da <= dt else min(dt * 1.5, da, 1.0e-3)
        assert O.multi_step(ob, dt)[0] == 0
        t += dt
        step += 1
    print("  oracle: %d steps" % step, flush=True)
    assert step == steps
    assert np.array_equal(ob[0].get_U(), U0d) and np.array_equal(ob[1].get_U(), U1d)


def test_ramp_case_fp64(gpu):
    print("case ramp fp64", flush=True)
    steps, ((U0d, fs0), (U1d, fs1)), _ = _ramp_device("fp64")
    g0, g1 = R.grids()
    p = np.concatenate([fs0[1].reshape(40, 4, 10), fs1[1].reshape(40, 4, 30)], axis=2)
    pos = np.concatenate([R.cell_centres(g0), R.cell_centres(g1)], axis=2)
    angle, dev = R.shock_angle(pos, p)
    print("  device fp64: %d steps, shock angle %.3f deg" % (steps, angle), flush=True)
    assert abs(steps - 862) < 3
    assert abs(angle - 57.0) < 1.0


def test_grid_blocks_refuse_what_they_lack(gpu):
    gas = e4.IdealGas()
    g = curvy_grid(5, 4, 4, seed=81)
    b = e4.SFluidBlock.from_grid(gas, g)
    with pytest.raises(e4.E4hipError):
        b.set_bc(Face.north, BCType.slip_wall_1)
    with pytest.raises(e4.E4hipError):
        b.set_bc(Face.east, BCType.extrapolate_1)
    u = e4.SFluidBlock(gas, 3, 5, 4, 4, 0.1, 0.1, 0.1)
    with pytest.raises(e4.E4hipError):
        e4.connect_blocks(b, Face.east, u, Face.west, 0)
    b.destroy()
    u.destroy()
