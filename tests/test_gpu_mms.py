"""User-defined source terms and ghost cells on the device: the reference's
manufactured-solution smoke test (Euler case 1: 2 x 2 blocks of 8 x 8 cells,
UserDefinedBC south/west, x_order-1 extrapolation north/east, no limiter or
clipping, stringent CFL, to t = 60 ms) through the package's host API, against
the reference's known answers (about 2620 steps; density error norms within
1e-3) and against the oracle's final state; then source terms on a 3D
ideal-gas block and on a mixture block against the oracle.  Bars: fp64_exact
bit-identical (the oracle starting, like the device, from the decoded initial
state), fp64 1e-9, mixed density 1e-5; mixtures the thermally perfect
tolerance."""
import numpy as np
import pytest

import pyoracle as O
import eilmer4_amd as e4
from eilmer4_amd.fvcore import BCType
import mms_case as mms
from parity_helpers import rel_err

pytestmark = pytest.mark.gpu


def _device_mms(precision):
    gas = e4.IdealGas(mMass=mms.MMASS, gamma=mms.GAMMA)
    layout, h = mms.block_layout(16, 2, 2)
    blks = {}
    def ghost_cell(t, x, y, z):
        return dict(zip(("p", "T", "velx", "vely"), mms.analytic_state(x, y)[1:]))

    for (ib, jb, ni, nj, origin) in layout:
        b = e4.SFluidBlock(gas, 2, ni, nj, 1, h, h)
        b.set_from_pT(1.0e5, 1.0e5 / mms.RGAS, 800.0, 800.0, 0.0)
        e4.udf.set_source_terms(b, lambda t, cell: dict(zip(("mass", "momentum_x", "momentum_y", "total_energy"),
                                                            mms.source_terms(cell["x"], cell["y"]))), origin=origin)
        if ib == 1:
            b.set_bc(e4.Face.east, e4.bc.OutFlowBC_SimpleExtrapolate(xOrder=1))
        if jb == 1:
            b.set_bc(e4.Face.north, e4.bc.OutFlowBC_SimpleExtrapolate(xOrder=1))
        if ib == 0:
            b.set_bc(e4.Face.west, e4.bc.UserDefinedBC(ghost_cell, origin=origin))
        if jb == 0:
            b.set_bc(e4.Face.south, e4.bc.UserDefinedBC(ghost_cell, origin=origin))
        blks[(ib, jb)] = b
    for (ib, jb), b in blks.items():
        if ib == 0:
            e4.connect_blocks(b, e4.Face.east, blks[(1, jb)], e4.Face.west)
        if jb == 0:
            e4.connect_blocks(b, e4.Face.north, blks[(ib, 1)], e4.Face.south)
    order = [blks[(ib, jb)] for (ib, jb, *_r) in layout]
    cfg = e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv, gasdynamic_update_scheme=e4.GasdynamicUpdate.pc,
                          apply_limiter=False, extrema_clipping=False, stringent_cfl=True, cfl_value=0.5,
                          dt_init=1.0e-6, dt_max=1.0e-3, max_time=60.0e-3, max_step=10 ** 6,
                          precision_mode=e4.PrecisionMode[precision])
    sim = e4.Simulation(order, cfg)
    steps = sim.integrate_in_time()
    return steps, order, layout, h


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64", "mixed"])
def test_mms_case1_known_answer(gpu, precision):
    print("case mms euler case 1, 16x16 as 2x2 blocks, precision=%s, to t=60 ms" % precision, flush=True)
    steps, blocks, layout, h = _device_mms(precision)
    fs = [b.download(with_fs=True)[1] for b in blocks]
    norms = mms.density_norms([f[0] for f in fs], layout, h)
    print("  steps %d norms %s" % (steps, norms), flush=True)
    assert abs(steps - mms.CASE1_STEPS) < 3
    for k, ref in mms.CASE1_NORMS.items():
        assert abs(norms[k] - ref) / ref < 1e-3, (k, norms[k], ref)
    # against the oracle's run of the same case
    osteps, onorms, oblocks = mms.run_oracle(O, decode_initial=True)
    assert steps == osteps
    Ud = np.concatenate([b.download() for b in blocks], axis=1)
    Uo = np.concatenate([b.get_U() for b in oblocks], axis=1)
    err = rel_err(Ud, Uo)
    print("  vs oracle: rel err %s bitwise %s" % (err, np.array_equal(Ud, Uo)), flush=True)
    for b in blocks:
        b.destroy()
    if precision == "fp64_exact":
        assert np.array_equal(Ud, Uo)
    elif precision == "fp64":
        assert err.max() <= 1e-9
    else:
        assert err[0] <= 1e-5


def _random_sources(nq, n, seed, scale):
    rng = np.random.default_rng(seed)
    return scale[:, None] * (rng.random((nq, n)) - 0.5)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64"])
def test_source_terms_ideal_3d(gpu, precision):
    """Random per-cell sources on a closed 3D block (slip walls), three pc
    steps at the oracle's dt."""
    gas = e4.IdealGas()
    n = 12
    rng = np.random.default_rng(3)
    shape = (n, n, n)
    p = 1.0e5 * (1.0 + 0.2 * rng.random(shape))
    T = 300.0 * (1.0 + 0.2 * rng.random(shape))
    vx = 300.0 * (rng.random(shape) - 0.5)
    rho, u, _ = gas.rho_u_a_from_pT(p.ravel(), T.ravel())
    U0 = gas.encode(rho, u, vx.ravel(), np.zeros(rho.size), np.zeros(rho.size))
    Q = _random_sources(5, n ** 3, 4, np.array([50.0, 2.0e4, 2.0e4, 2.0e4, 5.0e7]))
    ob = O.Block(O.Gas.ideal(), 3, n, n, n, 0.01, 0.01, 0.01)
    ob.set_config(flux_calc=O.AUSMDV)
    db = e4.SFluidBlock(gas, 3, n, n, n, 0.01, 0.01, 0.01)
    for face in e4.Face:
        ob.set_bc(int(face), int(BCType.slip_wall), None)
        db.set_bc(face, BCType.slip_wall, None)
    ob.set_U(U0)
    db.upload(U0)
    ob.set_source_terms(Q)
    db.set_source_terms(Q)
    sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                              precision_mode=e4.PrecisionMode[precision]))
    print("case source terms ideal 3D n=12 precision=%s" % precision, flush=True)
    dt, _, _ = ob.determine_dt(1.0e-3, 0.5, False)
    for _ in range(3):
        assert ob.step(dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    Ud, Uo = db.download(), ob.get_U()
    err = rel_err(Ud, Uo)
    print("  rel err %s bitwise %s" % (err, np.array_equal(Ud, Uo)), flush=True)
    # removing them restores the source-free path
    db.set_source_terms(None)
    db.destroy()
    if precision == "fp64_exact":
        assert np.array_equal(Ud, Uo)
    else:
        assert err.max() <= 1e-11


def test_source_terms_thermally_perfect(gpu):
    from test_gpu_tp import TOL
    from eilmer4_amd import workloads
    gas = e4.ThermallyPerfectGas.from_json()
    n = 8
    rng = np.random.default_rng(9)
    shape = (n, n, n)
    Y = np.array([0.7, 0.2, 0.05, 0.03, 0.02])
    Yf = np.broadcast_to(Y[:, None], (5, n ** 3)).copy()
    T = 1500.0 * (1.0 + 0.2 * rng.random(shape))
    p = 1.0e5 * (1.0 + 0.3 * rng.random(shape))
    rho = (p.ravel() / (gas.Rmix(Yf) * T.ravel())).reshape(shape)
    fields = dict(p=p, T=T, rho=rho, vx=200.0 * (rng.random(shape) - 0.5), vy=np.zeros(shape), vz=np.zeros(shape),
                  Y=Yf)
    U0, Tg = workloads.initial_U_tp(gas, fields)
    nq = U0.shape[0]
    scale = np.array([5.0, 1.0e4, 1.0e4, 1.0e4, 5.0e7] + [1.0] * (nq - 5))
    Q = _random_sources(nq, n ** 3, 10, scale)
    Q[5:] = 0.0
    Q[5] = Q[0]          # the mass source as nitrogen, so the species sum keeps the mass
    ob = O.Block(O.Gas.from_json("therm-perf-5-species-air.json"), 3, n, n, n, 0.1, 0.1, 0.1)
    ob.set_config(flux_calc=O.AUSMDV)
    db = e4.SFluidBlock(gas, 3, n, n, n, 0.1, 0.1, 0.1)
    for face in e4.Face:
        ob.set_bc(int(face), int(BCType.slip_wall), None)
        db.set_bc(face, BCType.slip_wall, None)
    ob.set_U(U0, Tg)
    db.upload(U0, Tg)
    ob.set_source_terms(Q)
    db.set_source_terms(Q)
    sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                              precision_mode=e4.PrecisionMode.fp64_exact))
    print("case source terms thermally perfect n=8", flush=True)
    dt, _, _ = ob.determine_dt(1.0e-3, 0.5, False)
    for _ in range(3):
        assert ob.step(dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    err = rel_err(db.download(), ob.get_U())
    print("  U rel err per quantity", err, flush=True)
    db.destroy()
    assert err.max() <= TOL


def test_stringent_cfl_matches_oracle(gpu):
    gas = e4.IdealGas()
    ni, nj = 12, 9
    rng = np.random.default_rng(12)
    shape = (1, nj, ni)
    p = 1.0e5 * (1.0 + 0.2 * rng.random(shape))
    T = 300.0 * (1.0 + 0.2 * rng.random(shape))
    vx, vy = 400.0 * rng.random(shape), 100.0 * rng.random(shape)
    rho, u, _ = gas.rho_u_a_from_pT(p.ravel(), T.ravel())
    U0 = gas.encode(rho, u, vx.ravel(), vy.ravel(), np.zeros(rho.size))
    ob = O.Block(O.Gas.ideal(), 2, ni, nj, 1, 0.03, 0.01)
    ob.set_config(flux_calc=O.AUSMDV)
    ob.set_U(U0)
    db = e4.SFluidBlock(gas, 2, ni, nj, 1, 0.03, 0.01)
    db.upload(U0)
    for stringent in (False, True):
        ob.set_stringent_cfl(stringent)
        sim = e4.Simulation([db], e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv,
                                                  stringent_cfl=stringent))
        dt_o = ob.determine_dt(1.0e-9, 0.5, False)[0]
        dt_d = sim.block_dt(1.0e-9, False)[0]
        print("case stringent_cfl=%s dt oracle %.17g device %.17g" % (stringent, dt_o, dt_d), flush=True)
        assert dt_d == dt_o
    db.destroy()
