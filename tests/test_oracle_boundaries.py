"""Boundary flavours of the oracle that change the face fluxes rather than only
the ghost cells: WallBC_WithSlip1 (no ghost-cell data; one-sided reconstruction
and the wall flux of fluxcalc.d:130-298), OutFlowBC_SimpleFlux (the boundary
face's flux from the interior cell, boundary_flux_effect.d:367-546) and
OutFlowBC_SimpleExtrapolate with x_order 1 (extrapolate_copy.d:102-186).
Each check is a closed form or an invariant that a slip in the restatement
(sign, side, index, exponent) breaks.  CPU only."""
import numpy as np
import pytest
from scipy.optimize import brentq

import pyoracle as O

G = 1.4
RGAS = 8.31451 / 0.02896


def state(p, T, vel, gas=None):
    return O.flowstate_pT(gas or O.Gas.ideal(), p, T, vel)


# ---------------------------------------------------------------- wall flux

def test_wall_flux_at_rest_is_the_gas_pressure():
    """Gas at rest against a wall: the Riemann-invariant estimate gives back p
    (tmp = p^((g-1)/2g)), no mass or energy crosses, momentum flux p n."""
    g = O.Gas.ideal()
    rng = np.random.default_rng(11)
    for _ in range(50):
        st = state(rng.uniform(1e3, 1e6), rng.uniform(200, 3000), (0.0, rng.uniform(-300, 300), 0.0), g)
        for left in (True, False):
            F = O.wall_flux(g, left, st)
            assert F[0] == 0.0 and F[4] == 0.0 and F[2] == 0.0 and F[3] == 0.0
            assert abs(F[1] - st[1]) / st[1] < 1e-14


def test_wall_flux_rarefaction_closed_form():
    """Gas leaving the wall: the isentropic wave gives
    p* = p (1 - (g-1)/2 |v|/a)^(2g/(g-1)) (Toro 2009, eq. 4.86 with u* = 0),
    and the floor 0.1 Pa once the gas separates (|v| >= 2a/(g-1))."""
    g = O.Gas.ideal()
    rng = np.random.default_rng(12)
    for _ in range(50):
        p, T = rng.uniform(1e4, 1e6), rng.uniform(200, 2000)
        a = np.sqrt(G * RGAS * T)
        v = rng.uniform(0.0, 0.9) * 2 * a / (G - 1)
        exact = p * (1 - 0.5 * (G - 1) * v / a) ** (2 * G / (G - 1))
        Fl = O.wall_flux(g, True, state(p, T, (v, 0, 0), g))       # wall on the left, gas moving +x
        Fr = O.wall_flux(g, False, state(p, T, (-v, 0, 0), g))     # wall on the right, gas moving -x
        assert abs(Fl[1] - exact) / exact < 1e-10
        assert abs(Fr[1] - exact) / exact < 1e-10
    a = np.sqrt(G * RGAS * 300.0)
    assert O.wall_flux(g, True, state(1e5, 300.0, (2.5 * a / (G - 1), 0, 0), g))[1] == 0.1


def exact_reflected_shock_pressure(p, rho, v):
    """Toro 2009 eq. 4.6-4.7: the pressure behind the shock that brings gas
    moving into a wall at speed v to rest, f(p*) = v."""
    A, B = 2.0 / ((G + 1) * rho), (G - 1) / (G + 1) * p
    return brentq(lambda ps: (ps - p) * np.sqrt(A / (ps + B)) - v, p, 1e4 * p, xtol=1e-12 * p)


def test_wall_flux_compression_matches_the_exact_shock():
    """Gas driven into the wall: the secant iteration (1 % increment tolerance)
    lands on the exact reflected-shock pressure; ten times the gas pressure
    caps it."""
    g = O.Gas.ideal()
    rng = np.random.default_rng(13)
    for _ in range(40):
        p, T = rng.uniform(1e4, 1e6), rng.uniform(200, 2000)
        rho = p / (RGAS * T)
        a = np.sqrt(G * RGAS * T)
        v = rng.uniform(0.2, 1.5) * a
        ex = min(exact_reflected_shock_pressure(p, rho, v), 10 * p)
        Fl = O.wall_flux(g, True, state(p, T, (-v, 0, 0), g))
        Fr = O.wall_flux(g, False, state(p, T, (v, 0, 0), g))
        assert Fl[1] == Fr[1]
        assert abs(Fl[1] - ex) / ex < 2e-3, (v / a, Fl[1] / ex)
    # the cap, where the exact reflected shock is stronger (gas at 2.5 and 4
    # times its sound speed); the reference's iteration itself leaves the
    # physical range for much stronger impacts (6a gives a negative pressure),
    # which the restatement keeps
    for m in (2.5, 4.0):
        assert O.wall_flux(g, True, state(1e5, 300.0, (-m * np.sqrt(G * RGAS * 300), 0, 0), g))[1] == 1e6


def test_wall_flux_thermally_perfect_uses_cp_over_cv():
    """For the mixture the wave relations use gamma = Cp/Cv of the state
    (gas_model.d:162); a calorically frozen check: at rest p* = p, and for a
    weak rarefaction the isentropic form with that gamma."""
    g = O.Gas.from_json("therm-perf-5-species-air.json")
    Y = np.array([0.7, 0.2, 0.05, 0.03, 0.02])
    st = O.flowstate_pT(g, 2e5, 2500.0, (0.0, 0.0, 0.0), Y)
    F = O.wall_flux(g, True, st)
    assert abs(F[1] - 2e5) / 2e5 < 1e-13
    assert np.all(F[5:] == 0.0) and F[0] == 0.0
    Cp = sum(Y[i] * g.eval_Cp(i, 2500.0) for i in range(5))
    Rm = st[1] / (st[0] * st[2])
    gam = Cp / (Cp - Rm)
    v = 0.1 * st[4]
    st2 = O.flowstate_pT(g, 2e5, 2500.0, (v, 0.0, 0.0), Y)
    exact = 2e5 * (1 - 0.5 * (gam - 1) * v / st2[4]) ** (2 * gam / (gam - 1))
    assert abs(O.wall_flux(g, True, st2)[1] - exact) / exact < 1e-10


# ---------------------------------------------------------- outflow flux

def test_outflow_flux_is_the_physical_flux_or_a_closed_face():
    g = O.Gas.ideal()
    rng = np.random.default_rng(14)
    for _ in range(50):
        vel = rng.uniform(-800, 800, 3)
        st = state(rng.uniform(1e4, 1e6), rng.uniform(200, 2000), vel, g)
        rho, p, u = st[0], st[1], st[3]
        E = rho * (u + 0.5 * np.dot(vel, vel))
        phys = np.array([rho * vel[0], rho * vel[0] * vel[0] + p, rho * vel[0] * vel[1], rho * vel[0] * vel[2],
                         vel[0] * (E + p)])
        closed = np.array([0.0, p, 0.0, 0.0, 0.0])
        for outsign in (1, -1):
            F = O.outflow_flux(g, outsign, st)
            want = phys if outsign * vel[0] > 0 else closed
            assert np.max(np.abs(F - want)) / np.abs(want).max() < 1e-14


# ------------------------------------------------- one-sided reconstruction

def test_one_sided_reconstruction_scalars():
    """l2r1/l1r2 (onedinterp.d:165-227): linear data are reproduced on both
    sides; the full side equals l2r2's; at an extremum the linear side takes
    the cell value; clipping keeps both sides inside the two cell values.
    Wall-side extrapolation (linear_extrap_prepare): 1.5 q0 - 0.5 q1."""
    rng = np.random.default_rng(15)
    for _ in range(100):
        a, b = rng.uniform(-5, 5, 2)
        q = a + b * np.arange(4)
        qL, qR = O.interp_onesided_scalar(0, q[:3])
        assert abs(qL - (a + 1.5 * b)) < 1e-12 and abs(qR - (a + 1.5 * b)) < 1e-12
        qL, qR = O.interp_onesided_scalar(1, q[:3])
        assert abs(qL - (a + 0.5 * b)) < 1e-12 and abs(qR - (a + 0.5 * b)) < 1e-12
        # the full side is l2r2's
        q4 = rng.uniform(-5, 5, 4)
        l2r2 = O.interp_l2r2_scalar([1.0] * 4, q4)
        assert O.interp_onesided_scalar(0, q4[:3])[0] == l2r2[0]
        assert O.interp_onesided_scalar(1, q4[1:])[1] == l2r2[1]
        qL, qR = O.interp_onesided_scalar(0, q4[:3])
        lo, hi = min(q4[1], q4[2]), max(q4[1], q4[2])
        assert lo <= qL <= hi and lo <= qR <= hi
        if (q4[1] - q4[0]) * (q4[2] - q4[1]) < 0:
            assert qR == q4[2]
        else:
            assert qR == 0.5 * q4[1] + 0.5 * q4[2]
        qL, qR = O.interp_onesided_scalar(1, q4[1:])
        if (q4[2] - q4[1]) * (q4[3] - q4[2]) < 0:
            assert qL == q4[1]
        assert O.interp_onesided_scalar(2, q4[:2]) == 1.5 * q4[0] - 0.5 * q4[1]


# -------------------------------------------------------------- blocks

def box(ni, nj, nk, rng, bc, calc=O.AUSMDV, clip=1, v=50.0, amp=0.05):
    g = O.Gas.ideal()
    b = O.Block(g, 3, ni, nj, nk, 1.0 / ni, 1.0 / nj, 1.0 / nk)
    b.set_config(flux_calc=calc, extrema_clipping=clip)
    for f, t in bc.items():
        b.set_bc(f, t)
    shape = (nk, nj, ni)
    p = 1e5 * (1 + amp * rng.standard_normal(shape))
    T = 300 * (1 + amp * rng.standard_normal(shape))
    vel = [v * rng.standard_normal(shape) for _ in range(3)]
    b.set_from_pT(p, T, *vel)
    return b, (p, T, vel)


@pytest.mark.parametrize("clip", [1, 0])
def test_slip1_box_conserves_mass_and_energy(clip):
    """A box closed by WallBC_WithSlip1 on all six faces: the wall flux carries
    no mass and no energy, so their totals change only by summation rounding."""
    rng = np.random.default_rng(16)
    b, _ = box(10, 8, 6, rng, {f: O.BC_SLIP_WALL_1 for f in range(6)}, clip=clip)
    U0 = b.get_U().sum(axis=1)
    for _ in range(10):
        assert b.step(2e-6)[0] == 0
    U1 = b.get_U().sum(axis=1)
    assert abs(U1[0] - U0[0]) / U0[0] < 1e-14
    assert abs(U1[4] - U0[4]) / U0[4] < 1e-14


def test_slip1_gas_at_rest_stays_at_rest():
    rng = np.random.default_rng(17)
    b, _ = box(8, 8, 8, rng, {f: O.BC_SLIP_WALL_1 for f in range(6)}, v=0.0, amp=0.0)
    U0 = b.get_U()
    for _ in range(5):
        assert b.step(1e-5)[0] == 0
    U = b.get_U()
    assert np.array_equal(U[0], U0[0]) and np.array_equal(U[4], U0[4])
    assert np.abs(U[1:4]).max() < 1e-12 * 1e5 * 1e-5 / 0.125


@pytest.mark.parametrize("kind", [O.BC_SLIP_WALL_1, O.BC_OUTFLOW_SIMPLE_FLUX, O.BC_EXTRAPOLATE_1])
@pytest.mark.parametrize("clip", [1, 0])
def test_mirror_symmetry_of_one_sided_boundaries(kind, clip):
    """The boundary on the west face and the x-mirrored field with the boundary
    on the east face give x-mirrored results: the left and right forms
    (l1r2 / l2r1, left / right wall, outsign -1 / +1) are mirror images."""
    rng = np.random.default_rng(18)
    ni, nj, nk = 9, 4, 3
    a, (p, T, vel) = box(ni, nj, nk, rng, {O.WEST: kind, O.EAST: O.BC_EXTRAPOLATE}, clip=clip, v=150.0)
    b = O.Block(O.Gas.ideal(), 3, ni, nj, nk, 1.0 / ni, 1.0 / nj, 1.0 / nk)
    b.set_config(flux_calc=O.AUSMDV, extrema_clipping=clip)
    b.set_bc(O.EAST, kind)
    b.set_bc(O.WEST, O.BC_EXTRAPOLATE)
    fl = lambda f: np.ascontiguousarray(f[:, :, ::-1])
    b.set_from_pT(fl(p), fl(T), -fl(vel[0]), fl(vel[1]), fl(vel[2]))
    for _ in range(4):
        assert a.step(2e-6)[0] == 0 and b.step(2e-6)[0] == 0
    Ua = a.get_U().reshape(5, nk, nj, ni)
    Ub = b.get_U().reshape(5, nk, nj, ni)
    for q, sgn in ((0, 1), (1, -1), (2, 1), (3, 1), (4, 1)):
        ref = sgn * Ua[q][:, :, ::-1]
        assert np.max(np.abs(Ub[q] - ref)) <= 1e-13 * np.abs(Ua[q]).max()


def test_simple_flux_equals_extrapolation_for_supersonic_outflow():
    """Supersonic outflow: the extrapolated ghost makes both sides of the
    boundary face the cell's state, whose AUSMDV flux is the physical flux the
    simple outflow flux writes directly."""
    rng = np.random.default_rng(19)
    bc_a = {O.EAST: O.BC_OUTFLOW_SIMPLE_FLUX, O.WEST: O.BC_EXTRAPOLATE}
    bc_b = {O.EAST: O.BC_EXTRAPOLATE, O.WEST: O.BC_EXTRAPOLATE}
    a, (p, T, vel) = box(12, 3, 3, rng, bc_a, v=10.0)
    b, _ = box(12, 3, 3, np.random.default_rng(19), bc_b, v=10.0)
    shape = (3, 3, 12)
    for blk in (a, b):
        blk.set_from_pT(p, T, np.full(shape, 900.0) + vel[0], vel[1], vel[2])
    for _ in range(5):
        assert a.step(1e-6)[0] == 0 and b.step(1e-6)[0] == 0
    Ua, Ub = a.get_U(), b.get_U()
    assert np.max(np.abs(Ua - Ub) / np.abs(Ub).max(axis=1, keepdims=True)) < 1e-12


def test_simple_flux_closes_the_face_against_inflow():
    """Flow into the block through an OutFlowBC_SimpleFlux face: the face
    carries only pressure, so no mass enters through it."""
    rng = np.random.default_rng(20)
    a, (p, T, vel) = box(8, 3, 3, rng, {O.EAST: O.BC_OUTFLOW_SIMPLE_FLUX, O.WEST: O.BC_SLIP_WALL}, v=5.0)
    shape = (3, 3, 8)
    a.set_from_pT(p, T, np.full(shape, -100.0) + vel[0], vel[1], vel[2])
    m0 = a.get_U()[0].sum()
    for _ in range(4):
        assert a.step(1e-6)[0] == 0
    assert abs(a.get_U()[0].sum() - m0) / m0 < 1e-14


def test_extrapolate_x_order_1_ghosts():
    """x_order 1: the ghost cells continue the two cells inside linearly in
    rho, u and velocity; T and p follow from the EOS; the sound speed and
    shock indicator are the ones the cell was constructed with (p = 100 kPa,
    T = 300 K, fvcell.d:170), which the extrapolation does not write."""
    g = O.Gas.ideal()
    ni, nj, nk = 6, 3, 3
    b = O.Block(g, 3, ni, nj, nk, 0.1, 0.1, 0.1)
    b.set_config(flux_calc=O.AUSMDV)
    for f in range(6):
        b.set_bc(f, O.BC_EXTRAPOLATE_1)
    x = np.arange(ni, dtype=float)
    shape = (nk, nj, ni)
    T = np.broadcast_to(300.0 + 8.0 * x, shape)
    p = np.broadcast_to(1e5 + 512.0 * x, shape)
    b.set_from_pT(p, T, np.broadcast_to(16.0 * x, shape), 4.0, -2.0)
    fs, S = b.get_fs_full()
    inner = b.get_fs()
    rho = inner[0].reshape(shape)[1, 1]
    u = inner[3].reshape(shape)[1, 1]
    gi = fs[:, 3, 3]                      # the (j, k) = (1, 1) line, slots i = 0 .. ni+3
    assert gi[0, 1] == 2 * rho[0] - rho[1] and gi[0, 0] == 2 * gi[0, 1] - rho[0]
    assert gi[0, ni + 2] == 2 * rho[-1] - rho[-2] and gi[0, ni + 3] == 2 * gi[0, ni + 2] - rho[-1]
    assert gi[3, ni + 2] == 2 * u[-1] - u[-2]
    assert gi[5, ni + 2] == 2 * 16.0 * (ni - 1) - 16.0 * (ni - 2)
    cv = RGAS / (G - 1)
    assert gi[2, ni + 2] == gi[3, ni + 2] * (1.0 / cv)
    a_init = np.sqrt(G * RGAS * 300.0)
    assert abs(gi[4, 0] - a_init) < 1e-12 * a_init and gi[4, ni + 3] == gi[4, 0]
    assert S[3, 3, 0] == 0 and S[3, 3, ni + 3] == 0


def test_slip1_detector_uses_the_wall_relative_velocity():
    """Without ghost data the boundary face of PJ_ShockDetector has one cell
    (shockdetectors.d:39-54): gas running into the wall at more than 0.3 of
    its sound speed marks the cell, gas leaving it does not."""
    g = O.Gas.ideal()
    ni = 8
    b = O.Block(g, 3, ni, 2, 2, 0.1, 0.1, 0.1)
    b.set_config(flux_calc=O.ADAPTIVE_HANEL_AUSMDV)
    b.set_bc(O.WEST, O.BC_SLIP_WALL_1)
    b.set_bc(O.EAST, O.BC_SLIP_WALL_1)
    a = np.sqrt(G * RGAS * 300.0)
    vx = np.where(np.arange(ni) < ni // 2, -0.5 * a, -0.2 * a)
    b.set_from_pT(1e5, 300.0, np.broadcast_to(vx, (2, 2, ni)), 0.0, 0.0)
    assert b.step(1e-7)[0] == 0
    S = b.get_S().reshape(2, 2, ni)[0, 0]
    assert S[0] == 1 and S[-1] == 0


# --------------------------------------------------------------- axisymmetric

def axi_block(ni, nj, y0, bc, rng=None, v=(300.0, 0.0), amp=0.0, calc=O.AUSMDV):
    g = O.Gas.ideal()
    b = O.Block(g, 2, ni, nj, 1, 1.0 / ni, 0.5 / nj)
    b.set_config(flux_calc=calc)
    for f, t in bc.items():
        b.set_bc(f, t)
    b.set_axisymmetric(0.0, y0)
    shape = (1, nj, ni)
    rng = rng or np.random.default_rng(0)
    p = 1e5 * (1 + amp * rng.standard_normal(shape))
    T = 300 * (1 + amp * rng.standard_normal(shape))
    b.set_from_pT(p, T, v[0] + 50 * amp * rng.standard_normal(shape), v[1] + 50 * amp * rng.standard_normal(shape), 0.0)
    return b


def test_axisymmetric_uniform_flow_along_the_axis():
    """Uniform flow parallel to the axis: the pressure on a cell's outer and
    inner faces (areas dx y_N and dx y_S per radian) is balanced by the source
    p areaxy / V (add_inviscid_source_vector), so nothing changes beyond
    rounding; without the source the cells would accelerate outwards."""
    b = axi_block(12, 8, 0.25, {O.EAST: O.BC_PERIODIC, O.WEST: O.BC_PERIODIC, O.NORTH: O.BC_SLIP_WALL,
                                O.SOUTH: O.BC_SLIP_WALL})
    U0 = b.get_U()
    for _ in range(10):
        assert b.step(1e-5)[0] == 0
    U = b.get_U()
    assert np.max(np.abs(U[0] - U0[0])) / U0[0].max() < 1e-13
    assert np.max(np.abs(U[2])) < 1e-13 * 1e5 * 10 * 1e-5 / (0.5 / 8)


def test_axisymmetric_closed_box_conserves_mass_and_energy():
    """Per-radian volumes: sum of V rho and V rho E over a box closed by slip
    walls is conserved to rounding (radial momentum is not: the walls push)."""
    rng = np.random.default_rng(7)
    b = axi_block(10, 10, 0.0, {f: O.BC_SLIP_WALL for f in range(4)}, rng, v=(0.0, 0.0), amp=0.05)
    # per-radian cell volumes of the uniform grid, y from 0: dx (y_{j+1}^2 - y_j^2) / 2
    dx, dy = 0.1, 0.05
    yj = dy * np.arange(11)
    vol = np.broadcast_to((dx * 0.5 * (yj[1:] ** 2 - yj[:-1] ** 2))[:, None], (10, 10)).ravel()
    U0 = b.get_U()
    for _ in range(10):
        assert b.step(2e-6)[0] == 0
    U = b.get_U()
    for q in (0, 4):
        assert abs((vol * U[q]).sum() - (vol * U0[q]).sum()) / (vol * U0[q]).sum() < 1e-14


def test_axisymmetric_far_from_the_axis_is_planar():
    """Far from the axis (y0 = 1e4, cells 0.06 high) the per-radian geometry is
    a scaled planar one and the source is small: one step agrees with the
    planar block to O(dy / y0)."""
    rng = np.random.default_rng(8)
    bc = {O.EAST: O.BC_EXTRAPOLATE, O.WEST: O.BC_EXTRAPOLATE, O.NORTH: O.BC_SLIP_WALL, O.SOUTH: O.BC_SLIP_WALL}
    a = axi_block(8, 8, 1.0e4, bc, rng, amp=0.1)
    g = O.Gas.ideal()
    pl = O.Block(g, 2, 8, 8, 1, 1.0 / 8, 0.5 / 8)
    pl.set_config(flux_calc=O.AUSMDV)
    for f, t in bc.items():
        pl.set_bc(f, t)
    pl.set_U(a.get_U())
    U0 = a.get_U()
    assert a.step(2e-6)[0] == 0 and pl.step(2e-6)[0] == 0
    dA, dP = a.get_U() - U0, pl.get_U() - U0
    rel = np.max(np.abs(dA - dP)) / np.max(np.abs(dP))
    assert 0.0 < rel < 1e-4, rel
