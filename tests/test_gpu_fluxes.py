"""The flux calculators beyond ausmdv, hanel and hllc on the device -- efm,
ausm_plus_up, adaptive_efm_ausmdv and roe (fluxcalc.d:626-1037, 1202-1371) --
against the oracle, on every kernel family that evaluates face fluxes:

  uniform ideal-gas blocks (the fused stage kernel and the boundary-cell
      kernel), single and batched;
  blocks built from their grid (per-face frames);
  thermally perfect five-species blocks (face kernels of the mixture).

Bars: fp64_exact bit-identical where the calculator uses only +, -, *, /,
sqrt (ausm_plus_up, roe); efm evaluates exp(-x^2), whose last bit can differ
between the device math library and glibc, so efm and adaptive_efm_ausmdv are
held to 1e-12 per conserved quantity in every FP64 mode.  fp64 <= 1e-12,
mixed density <= 1e-5, mixtures <= 1e-10 (the logarithm of the enthalpy).
"""
import numpy as np
import pytest

import pyoracle as O
import eilmer4_amd as e4
from eilmer4_amd import workloads
from eilmer4_amd.fvcore import BCType, Face
from parity_helpers import run_pair, rel_err

pytestmark = pytest.mark.gpu

NEW = ["efm", "ausm_plus_up", "adaptive_efm_ausmdv", "roe"]
BITWISE = {"ausm_plus_up", "roe"}


def _check(flux, precision, Ud, Uo, tol=1.0e-12):
    err = rel_err(Ud, Uo)
    bitwise = bool(np.array_equal(Ud, Uo))
    print("  %s %s: rel err %s bitwise %s" % (flux, precision, err, bitwise), flush=True)
    assert np.all(np.isfinite(Ud))
    if precision == "fp64_exact" and flux in BITWISE:
        assert bitwise
    elif precision == "mixed":
        assert err[0] <= 1.0e-5
    else:
        assert err.max() <= tol


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64", "mixed"])
@pytest.mark.parametrize("flux", NEW)
def test_uniform_box(gpu, flux, precision):
    """16^3 shocked box (inflow, outflow, periodic faces), 3 RK2 steps."""
    geom, fields, bcs, _ = workloads.shocked_box_3d(e4.IdealGas(), n=16, seed=23)
    print("case box16 flux=%s precision=%s" % (flux, precision), flush=True)
    Uo, Ud, _, _, _ = run_pair(geom, fields, bcs, 3, flux=flux, precision=precision, same_dt=True)
    _check(flux, precision, Ud, Uo)


@pytest.mark.parametrize("flux", NEW)
def test_uniform_planar_2d_with_walls(gpu, flux):
    """2D 128x64 Mach-2 bump (inflow, outflow, slip walls), 20 steps in
    fp64_exact, each side with its own dt policy."""
    gas = e4.IdealGas()
    geom, fields, bcs, _ = workloads.planar_bump_2d(gas)
    print("case bump2d flux=%s" % flux, flush=True)
    Uo, Ud, dto, dtd, _ = run_pair(geom, fields, bcs, 20, flux=flux, precision="fp64_exact")
    _check(flux, "fp64_exact", Ud, Uo, tol=1.0e-11)


@pytest.mark.parametrize("flux", NEW)
def test_one_sided_boundaries(gpu, flux):
    """WallBC_WithSlip1 and OutFlowBC_SimpleFlux faces: the boundary-cell
    kernel evaluates the interior faces beside them with the calculator."""
    gas = e4.IdealGas()
    geom, fields, bcs, _ = workloads.shocked_box_3d(gas, n=12, seed=29)
    bcs = dict(bcs)
    bcs[e4.Face.north] = (e4.BCType.slip_wall_1, None)
    bcs[e4.Face.south] = (e4.BCType.slip_wall_1, None)
    bcs[e4.Face.east] = (e4.BCType.outflow_simple_flux, None)
    print("case one-sided flux=%s" % flux, flush=True)
    Uo, Ud, _, _, _ = run_pair(geom, fields, bcs, 3, flux=flux, precision="fp64_exact", same_dt=True)
    # the wall flux uses pow: not bitwise for any calculator
    err = rel_err(Ud, Uo)
    print("  rel err", err, flush=True)
    assert err.max() <= 1.0e-12


@pytest.mark.parametrize("flux", NEW)
def test_batched_blocks(gpu, flux):
    """Four connected 12^3 blocks in one batched launch per stage, against the
    oracle's multi-block step."""
    gas = e4.IdealGas()
    og = O.Gas.ideal()
    fc = e4.flux_calculator_from_name(flux)
    rng = np.random.default_rng(31)
    n = 12
    obl, dbl = [], []
    for b in range(4):
        p = 1.0e5 * (1.0 + 0.2 * rng.random((n, n, n)))
        T = 300.0 * (1.0 + 0.2 * rng.random((n, n, n)))
        v = [80.0 * rng.standard_normal((n, n, n)) for _ in range(3)]
        rho, u, _ = gas.rho_u_a_from_pT(p.ravel(), T.ravel())
        U0 = gas.encode(rho, u, *[x.ravel() for x in v])
        ob = O.Block(og, 3, n, n, n, 0.01, 0.01, 0.01)
        ob.set_config(flux_calc=int(fc))
        ob.set_U(U0)
        db = e4.SFluidBlock(gas, 3, n, n, n, 0.01, 0.01, 0.01)
        db.upload(U0)
        obl.append(ob)
        dbl.append(db)
    # a ring of blocks along x, periodic through the last one
    for b in range(4):
        c = (b + 1) % 4
        O.connect_blocks(obl[b], O.EAST, obl[c], O.WEST, 0)
        e4.connect_blocks(dbl[b], Face.east, dbl[c], Face.west, 0)
    sim = e4.Simulation(dbl, e4.GlobalConfig(flux_calculator=fc, precision_mode=e4.PrecisionMode.fp64_exact))
    print("case batched 4x12^3 flux=%s" % flux, flush=True)
    for _ in range(3):
        assert O.multi_step(obl, 2.0e-6)[0] == 0
        sim.state.dt_global = 2.0e-6
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    for ob, db in zip(obl, dbl):
        _check(flux, "fp64_exact", db.download(), ob.get_U())
        db.destroy()


@pytest.mark.parametrize("flux", NEW)
def test_curvilinear_block(gpu, flux):
    """A curvilinear 3D block (warped faces, per-face frames), 3 steps."""
    import test_gpu_geometry as G
    print("case curvilinear flux=%s" % flux, flush=True)
    Uo, Ud = G._run(3, flux, "fp64_exact")
    _check(flux, "fp64_exact", Ud, Uo)
    Uo, Ud = G._run(3, flux, "fp64")
    _check(flux, "fp64", Ud, Uo)


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64"])
@pytest.mark.parametrize("flux", NEW)
def test_thermally_perfect_block(gpu, flux, precision):
    """Five-species air, 12^3, 3 steps: species upwinded by the mass flux,
    efm's Cv and roe's gamma from the mixture at the face temperature."""
    from test_gpu_tp import setup_pair
    print("case tp n=12 flux=%s precision=%s" % (flux, precision), flush=True)
    ob, db, sim = setup_pair(12, flux=flux, precision=precision)
    # ausm_plus_up's low-Mach scaling (dissipation over fa, Liou 2006 eqs.
    # 72-75) shrinks its explicit limit: CFL 0.1 at this box's Mach numbers
    dt, _, _ = ob.determine_dt(1.0e-3, 0.1 if flux == "ausm_plus_up" else 0.5, False)
    sim.cfg.fixed_time_step = True
    for _ in range(3):
        assert ob.step(dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    err = rel_err(db.download(), ob.get_U())
    print("  rel err", err, flush=True)
    assert err.max() <= 1.0e-10
    db.destroy()


def test_refused_calculators(gpu):
    """hlle is MHD only (fluxcalc.d:1197) and adaptive_hlle_roe takes it near
    shocks: the device path refuses both."""
    gas = e4.IdealGas()
    b = e4.SFluidBlock(gas, 3, 8, 8, 8, 0.1, 0.1, 0.1)
    for fc in (e4.FluxCalculator.hlle, e4.FluxCalculator.adaptive_hlle_roe):
        with pytest.raises(e4.FlowSolverException):
            e4.Simulation([b], e4.GlobalConfig(flux_calculator=fc))
    b.destroy()
