"""Boundary flavours that change face fluxes on the device against the oracle:
WallBC_WithSlip1 (no ghost data: one-sided reconstruction, wall flux),
OutFlowBC_SimpleFlux (boundary outflow flux) and x_order-1 extrapolation,
through the stage kernel plus the boundary kernel that finishes the cells
beside such faces.  Bars: fp64_exact bit-identical except where the wall
flux's pow() enters (the device's and the host's pow may differ in the last
bit: 1e-13 there), fp64 1e-11, mixed density 1e-5."""
import numpy as np
import pytest

import pyoracle as O
import eilmer4_amd as e4
from eilmer4_amd.fvcore import BCType
from parity_helpers import oracle_block, device_block, device_sim, rel_err

pytestmark = pytest.mark.gpu


def _field(gas, shape, seed, vmean=(0.0, 0.0, 0.0), vamp=300.0):
    rng = np.random.default_rng(seed)
    p = 1.0e5 * (1.0 + 0.4 * rng.random(shape))
    T = 300.0 * (1.0 + 0.4 * rng.random(shape))
    v = [vmean[d] + vamp * (rng.random(shape) - 0.5) for d in range(3)]
    rho, u, _ = gas.rho_u_a_from_pT(p.ravel(), T.ravel())
    return gas.encode(rho, u, *[x.ravel() for x in v])


def _run(dims, n, bcs, U0, steps, flux, precision, clipping=True):
    gas = e4.IdealGas()
    ni, nj, nk = n
    geom = dict(dims=dims, ni=ni, nj=nj, nk=nk, dx=1.0 / ni, dy=1.0 / nj, dz=1.0 / nk)
    fc = e4.flux_calculator_from_name(flux)
    ob = oracle_block(geom, bcs, U0, flux=int(fc), clipping=int(clipping))
    db = device_block(gas, geom, bcs, U0)
    dsim = device_sim([db], flux=fc, clipping=clipping, precision=precision)
    osim = O.Simulation(ob)
    for s in range(steps):
        osim.run(max_step=s + 1)
        dsim.state.dt_global = osim.dt_global
        dsim.gasdynamic_explicit_increment_with_fixed_grid()
        dsim.state.step += 1
    Uo, Ud = ob.get_U(), db.download()
    db.destroy()
    return Uo, Ud


def _check(precision, Ud, Uo, wall):
    err = rel_err(Ud, Uo)
    bitwise = bool(np.array_equal(Ud, Uo))
    print("  %s: rel err %s bitwise %s" % (precision, err, bitwise), flush=True)
    assert np.all(np.isfinite(Ud))
    if precision == "fp64_exact":
        if wall:
            assert err.max() <= 1.0e-13
        else:
            assert bitwise
    elif precision == "mixed":
        assert err[0] <= 1.0e-5
    else:
        assert err.max() <= 1.0e-11


CASES = {
    # name: (bcs, mean velocity, wall flux involved)
    "slip1_box": ({f: (BCType.slip_wall_1, None) for f in range(6)}, (0.0, 0.0, 0.0), True),
    "slip1_west_east": ({O.WEST: (BCType.slip_wall_1, None), O.EAST: (BCType.slip_wall_1, None),
                         O.NORTH: (BCType.periodic, None), O.SOUTH: (BCType.periodic, None),
                         O.TOP: (BCType.slip_wall, None), O.BOTTOM: (BCType.extrapolate, None)},
                        (150.0, 0.0, 0.0), True),
    "simple_flux": ({O.EAST: (BCType.outflow_simple_flux, None), O.WEST: (BCType.extrapolate, None),
                     O.NORTH: (BCType.outflow_simple_flux, None), O.SOUTH: (BCType.slip_wall, None),
                     O.TOP: (BCType.outflow_simple_flux, None), O.BOTTOM: (BCType.outflow_simple_flux, None)},
                    (200.0, -120.0, 60.0), False),
    "extrapolate_1": ({O.EAST: (BCType.extrapolate_1, None), O.WEST: (BCType.extrapolate_1, None),
                       O.NORTH: (BCType.extrapolate_1, None), O.SOUTH: (BCType.slip_wall, None),
                       O.TOP: (BCType.periodic, None), O.BOTTOM: (BCType.periodic, None)},
                      (100.0, 50.0, 0.0), False),
    "mixed_flavours": ({O.EAST: (BCType.outflow_simple_flux, None), O.WEST: (BCType.slip_wall_1, None),
                        O.NORTH: (BCType.extrapolate_1, None), O.SOUTH: (BCType.slip_wall_1, None),
                        O.TOP: (BCType.outflow_simple_flux, None), O.BOTTOM: (BCType.slip_wall_1, None)},
                       (80.0, -40.0, 30.0), True),
}


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64", "mixed"])
@pytest.mark.parametrize("case", list(CASES))
def test_boundaries_3d(gpu, case, precision):
    bcs, vmean, wall = CASES[case]
    n = (11, 9, 7)
    U0 = _field(e4.IdealGas(), (n[2], n[1], n[0]), 31, vmean)
    print("case %s 3D %s precision=%s steps=4" % (case, n, precision), flush=True)
    Uo, Ud = _run(3, n, bcs, U0, 4, "ausmdv", precision)
    _check(precision, Ud, Uo, wall)


@pytest.mark.parametrize("flux", ["hllc", "adaptive_hanel_ausmdv"])
def test_boundaries_fluxes(gpu, flux):
    """The boundary kernel's interior faces use the same flux calculator,
    shock indicators included (a fast compression into a wall without ghost
    data flags the wall cell by the one-sided detector test)."""
    bcs, vmean, wall = CASES["mixed_flavours"]
    n = (12, 10, 8)
    U0 = _field(e4.IdealGas(), (n[2], n[1], n[0]), 32, (-250.0, 0.0, 0.0), vamp=500.0)
    print("case mixed_flavours flux=%s fp64_exact steps=4" % flux, flush=True)
    Uo, Ud = _run(3, n, bcs, U0, 4, flux, "fp64_exact")
    _check("fp64_exact", Ud, Uo, wall)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64"])
def test_boundaries_no_clipping(gpu, precision):
    """Without extrema clipping the wall face extrapolates the two cells next
    to it (interp_l0r2 / interp_l2r0) instead of copying the wall cell."""
    bcs, vmean, wall = CASES["mixed_flavours"]
    n = (10, 8, 6)
    U0 = _field(e4.IdealGas(), (n[2], n[1], n[0]), 33, vmean)
    print("case mixed_flavours clipping off precision=%s" % precision, flush=True)
    Uo, Ud = _run(3, n, bcs, U0, 4, "ausmdv", precision, clipping=False)
    _check(precision, Ud, Uo, wall)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64"])
def test_boundaries_2d(gpu, precision):
    """The 2D block (C1-like): WallBC_WithSlip1 on north and south,
    OutFlowBC_Simple on the east, supersonic inflow on the west."""
    gas = e4.IdealGas()
    ni, nj = 40, 16
    U0 = _field(gas, (1, nj, ni), 34, (600.0, 0.0, 0.0), vamp=200.0)
    rho, u, a = gas.rho_u_a_from_pT(np.array([1.0e5]), np.array([300.0]))
    fs = np.array([rho[0], 1.0e5, 300.0, u[0], a[0], 700.0, 0.0, 0.0])
    bcs = {O.WEST: (BCType.inflow, fs), O.EAST: (BCType.outflow_simple_flux, None),
           O.NORTH: (BCType.slip_wall_1, None), O.SOUTH: (BCType.slip_wall_1, None)}
    print("case 2D %dx%d slip1 + simple flux precision=%s steps=20" % (ni, nj, precision), flush=True)
    Uo, Ud = _run(2, (ni, nj, 1), bcs, U0, 20, "ausmdv", precision)
    _check(precision, Ud, Uo, True)


def test_boundaries_batched(gpu):
    """Two blocks in one batched launch, each with one-sided boundaries: the
    boundary kernel runs per block after the batched stage kernel."""
    gas = e4.IdealGas()
    n = 9
    fc = int(e4.FluxCalculator.ausmdv)
    Us = [_field(gas, (n, n, n), 40 + s, (60.0, 0.0, 0.0)) for s in range(2)]
    obs, dbs = [], []
    geom = dict(dims=3, ni=n, nj=n, nk=n, dx=0.1, dy=0.1, dz=0.1)
    bcs = [{O.WEST: (BCType.slip_wall_1, None), O.EAST: (BCType.outflow_simple_flux, None)},
           {O.NORTH: (BCType.extrapolate_1, None), O.BOTTOM: (BCType.slip_wall_1, None)}]
    for s in range(2):
        obs.append(oracle_block(geom, bcs[s], Us[s], flux=fc))
        dbs.append(device_block(gas, geom, bcs[s], Us[s]))
    dsim = device_sim(dbs, flux=e4.FluxCalculator.ausmdv, precision="fp64_exact")   # batch_blocks on by default
    for _ in range(3):
        dt = min(b.determine_dt(1e-3, 0.5, False)[0] for b in obs)
        assert O.multi_step(obs, dt)[0] == 0
        dsim.state.dt_global = dt
        dsim.gasdynamic_explicit_increment_with_fixed_grid()
    for s in range(2):
        err = rel_err(dbs[s].download(), obs[s].get_U())
        print("  block %d rel err %s" % (s, err), flush=True)
        assert err.max() <= 1.0e-13
        dbs[s].destroy()


# ------------------------------------------------------------------ mixtures

@pytest.mark.parametrize("precision", ["fp64_exact", "fp64", "mixed"])
@pytest.mark.parametrize("flux", ["ausmdv", "adaptive_hanel_ausmdv"])
def test_boundaries_thermally_perfect(gpu, flux, precision):
    """5-species air with every new flavour on one block: walls without ghost
    data on three faces (the boundary-face kernel's wall flux with gamma =
    Cp/Cv of the mixture state, and the one-sided faces after them), simple
    outflow flux, and x_order-1 ghosts (stored per face, T by the EOS from the
    previous ghost T).  Bar: 1e-10 (the logarithm differs between the device
    and host libraries); mixed (FP32 reconstruction and flux): density 1e-5."""
    from test_gpu_tp import setup_pair
    bcs = {e4.Face.west: (BCType.slip_wall_1, None), e4.Face.east: (BCType.outflow_simple_flux, None),
           e4.Face.north: (BCType.extrapolate_1, None), e4.Face.south: (BCType.slip_wall_1, None),
           e4.Face.top: (BCType.extrapolate_1, None), e4.Face.bottom: (BCType.slip_wall_1, None)}
    print("case tp new bcs n=10 flux=%s precision=%s steps=3" % (flux, precision), flush=True)
    ob, db, sim = setup_pair(10, flux=flux, precision=precision, bcs=bcs, seed=77)
    dt, _, _ = ob.determine_dt(1.0e-3, 0.5, False)
    for _ in range(3):
        assert ob.step(dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    err = rel_err(db.download(), ob.get_U())
    print("  U rel err per quantity", err, flush=True)
    assert np.all(np.isfinite(err))
    if precision == "mixed":
        assert err[0] <= 1.0e-5
    else:
        assert err.max() <= 1.0e-10
    db.destroy()


def test_boundaries_thermally_perfect_no_clipping(gpu):
    """Without clipping the wall faces' states are extrapolated and go
    through the mixture EOS before the wall flux.  A mild field (T 1500-1800 K,
    uniform composition): on the 300-8000 K workload the unclipped wall
    extrapolation drives the reference's own algorithm to invalid cells."""
    from test_gpu_tp import TOL
    gas = e4.ThermallyPerfectGas.from_json()
    n = 10
    rng = np.random.default_rng(5)
    shape = (n, n, n)
    Y = np.array([0.7, 0.2, 0.05, 0.03, 0.02])
    Yf = np.broadcast_to(Y[:, None], (5, n ** 3)).copy()
    T = 1500.0 * (1.0 + 0.2 * rng.random(shape))
    p = 1.0e5 * (1.0 + 0.3 * rng.random(shape))
    rho = (p.ravel() / (gas.Rmix(Yf) * T.ravel())).reshape(shape)
    fields = dict(p=p, T=T, rho=rho, vx=200.0 * (rng.random(shape) - 0.5), vy=200.0 * (rng.random(shape) - 0.5),
                  vz=200.0 * (rng.random(shape) - 0.5), Y=Yf)
    from eilmer4_amd import workloads
    U0, Tg = workloads.initial_U_tp(gas, fields)
    bcs = {e4.Face.west: BCType.slip_wall_1, e4.Face.east: BCType.slip_wall_1,
           e4.Face.north: BCType.outflow_simple_flux, e4.Face.south: BCType.extrapolate_1,
           e4.Face.top: BCType.periodic, e4.Face.bottom: BCType.periodic}
    ob = O.Block(O.Gas.from_json("therm-perf-5-species-air.json"), 3, n, n, n, 0.1, 0.1, 0.1)
    ob.set_config(flux_calc=O.AUSMDV, extrema_clipping=0)
    db = e4.SFluidBlock(gas, 3, n, n, n, 0.1, 0.1, 0.1)
    for face, kind in bcs.items():
        ob.set_bc(int(face), int(kind), None)
        db.set_bc(face, kind, None)
    ob.set_U(U0, Tg)
    sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv, extrema_clipping=False,
                                              precision_mode=e4.PrecisionMode.fp64_exact))
    db.upload(U0, Tg)
    print("case tp new bcs clipping off n=10", flush=True)
    dt, _, _ = ob.determine_dt(1.0e-3, 0.5, False)
    for _ in range(3):
        assert ob.step(dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    err = rel_err(db.download(), ob.get_U())
    print("  U rel err per quantity", err, flush=True)
    assert err.max() <= TOL
    db.destroy()


# ---------------------------------------------------------- static profiles

def _profile_file(gas, path, seed):
    """Profile points on a 5 x 4 lattice in (y, z) at x = -0.05, states
    varying over it, every third point a shock point."""
    from eilmer4_amd import flowfiles as ff
    rng = np.random.default_rng(seed)
    yy, zz = np.meshgrid((np.arange(5) + 0.5) / 5, (np.arange(4) + 0.5) / 4, indexing="xy")
    y, z = yy.ravel(), zz.ravel()
    m = y.size
    p = 1.0e5 * (1.0 + 0.3 * rng.random(m))
    T = 300.0 * (1.0 + 0.3 * rng.random(m))
    rho, u, a = gas.rho_u_a_from_pT(p, T)
    zero = np.zeros(m)
    v = {"pos.x": np.full(m, -0.05), "pos.y": y, "pos.z": z, "volume": np.full(m, 1e-3), "rho": rho,
         "vel.x": 500.0 + 200.0 * rng.random(m), "vel.y": 50.0 * (rng.random(m) - 0.5),
         "vel.z": 50.0 * (rng.random(m) - 0.5), "p": p, "a": a, "mu": zero, "k": zero, "mu_t": zero, "k_t": zero,
         "S": (np.arange(m) % 3 == 0).astype(float), "tke": zero, "omega": np.ones(m), "massf[0]-air": np.ones(m),
         "u": u, "T": T}
    fl = ff.FlowSolution(ff.flow_variable_list(gas), v, 3, m, 1, 1)
    ff.write_profile(path, fl, range(m))
    return ff.FlowProfile(path, gas)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64"])
@pytest.mark.parametrize("flux", ["ausmdv", "adaptive_hanel_ausmdv"])
def test_static_profile_inflow(gpu, tmp_path, flux, precision):
    """InFlowBC_StaticProfile on the west face: the ghost states of the
    nearest profile points (shock indicators included) on the device and in
    the oracle."""
    from eilmer4_amd import flowfiles as ff
    gas = e4.IdealGas()
    n = (10, 8, 6)
    prof = _profile_file(gas, tmp_path / "inflow.dat", 51)
    geom = dict(dims=3, ni=n[0], nj=n[1], nk=n[2], dx=1.0 / n[0], dy=1.0 / n[1], dz=1.0 / n[2])
    probe = e4.SFluidBlock(gas, 3, *n, geom["dx"], geom["dy"], geom["dz"])
    states = ff.profile_ghost_states(prof, probe, int(e4.Face.west))
    probe.destroy()
    assert states.shape == (2, n[1] * n[2], 10)
    bcs = {O.WEST: (BCType.inflow_profile, states), O.EAST: (BCType.outflow_simple_flux, None),
           O.NORTH: (BCType.extrapolate, None), O.SOUTH: (BCType.slip_wall, None),
           O.TOP: (BCType.periodic, None), O.BOTTOM: (BCType.periodic, None)}
    U0 = _field(gas, (n[2], n[1], n[0]), 52, (550.0, 0.0, 0.0), vamp=100.0)
    print("case static profile 3D %s flux=%s precision=%s steps=4" % (n, flux, precision), flush=True)
    Uo, Ud = _run(3, n, bcs, U0, 4, flux, precision)
    _check(precision, Ud, Uo, False)


def test_static_profile_inflow_thermally_perfect(gpu):
    """The mixture path reads the stored ghost states through the same
    per-face buffers as its x_order-1 ghosts."""
    from test_gpu_tp import setup_pair
    gas = e4.ThermallyPerfectGas.from_json()
    n = 10
    rng = np.random.default_rng(53)
    Y = np.array([0.74, 0.2, 0.03, 0.02, 0.01])
    states = np.zeros((2, n * n, 9 + 5))
    for g in range(2 * n * n):
        fs = gas.flowstate(5.0e4 * (1 + 0.2 * rng.random()), 1500.0 * (1 + 0.2 * rng.random()),
                           (900.0, 20.0 * rng.random(), 0.0), Y)
        states.reshape(-1, 14)[g, :8] = fs[:8]
        states.reshape(-1, 14)[g, 8] = float(g % 4 == 0)
        states.reshape(-1, 14)[g, 9:] = fs[8:]
    bcs = {e4.Face.west: (BCType.inflow_profile, states), e4.Face.east: (BCType.extrapolate, None),
           e4.Face.north: (BCType.slip_wall, None), e4.Face.south: (BCType.slip_wall, None),
           e4.Face.top: (BCType.periodic, None), e4.Face.bottom: (BCType.periodic, None)}
    print("case tp static profile n=10", flush=True)
    ob, db, sim = setup_pair(10, flux="adaptive_hanel_ausmdv", precision="fp64", bcs=bcs, seed=54)
    dt, _, _ = ob.determine_dt(1.0e-3, 0.5, False)
    for _ in range(3):
        assert ob.step(dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    err = rel_err(db.download(), ob.get_U())
    print("  U rel err per quantity", err, flush=True)
    assert err.max() <= 1.0e-10
    db.destroy()
