"""The oracle on the reference's manufactured-solution smoke test (Euler,
case 1) and the pieces it exercises: user-defined source terms, user-defined
ghost cells, x_order-1 extrapolation without limiter or clipping, the
stringent CFL signal."""
import numpy as np
import pytest

import mms_case as mms


def test_hand_derived_sources_match_sympy():
    """The divergence of the Euler fluxes differentiated by hand equals the
    one a computer-algebra system gives (make_lua_files.py's route)."""
    sp = pytest.importorskip("sympy")
    x, y = sp.symbols("x y")
    c = mms.SUPERSONIC
    pi = sp.pi
    for case in (1, 3):
        S = 1 if case == 1 else sp.exp(-16 * ((x - 0.5) ** 2 + (y - 0.5) ** 2))
        rho = c["rho0"] + S * (c["rhox"] * sp.sin(c["arhox"] * pi * x) + c["rhoy"] * sp.cos(c["arhoy"] * pi * y)
                               + c["rhoxy"] * sp.cos(c["arhoxy"] * pi * x * y))
        u = c["u0"] + S * (c["ux"] * sp.sin(c["aux"] * pi * x) + c["uy"] * sp.cos(c["auy"] * pi * y)
                           + c["uxy"] * sp.cos(c["auxy"] * pi * x * y))
        v = c["v0"] + S * (c["vx"] * sp.cos(c["avx"] * pi * x) + c["vy"] * sp.sin(c["avy"] * pi * y)
                           + c["vxy"] * sp.cos(c["avxy"] * pi * x * y))
        p = c["p0"] + S * (c["px"] * sp.cos(c["apx"] * pi * x) + c["py"] * sp.sin(c["apy"] * pi * y)
                           + c["pxy"] * sp.sin(c["apxy"] * pi * x * y))
        et = p / rho / (mms.GAMMA - 1) + u * u / 2 + v * v / 2
        exprs = [sp.diff(rho * u, x) + sp.diff(rho * v, y),
                 sp.diff(rho * u * u + p, x) + sp.diff(rho * u * v, y),
                 sp.diff(rho * v * u, x) + sp.diff(rho * v * v + p, y),
                 sp.diff(rho * u * et + p * u, x) + sp.diff(rho * v * et + p * v, y)]
        fns = [sp.lambdify((x, y), e, "numpy") for e in exprs]
        rng = np.random.default_rng(7 + case)
        px_, py_ = rng.uniform(-0.1, 1.1, 50), rng.uniform(-0.1, 1.1, 50)
        mine = mms.source_terms(px_, py_, case)
        for q in range(4):
            ref = np.broadcast_to(fns[q](px_, py_), px_.shape)
            scale = np.max(np.abs(ref))
            assert np.max(np.abs(mine[q] - ref)) <= 1e-12 * scale, (case, q)


def test_mms_case1_reproduces_reference_norms(oracle):
    """mms-euler-test.rb: about 2620 steps and the density error norms at
    t = 60 ms within 1e-3 of the reference's; the oracle lands within 1e-9."""
    steps, norms, _ = mms.run_oracle(oracle)
    assert abs(steps - mms.CASE1_STEPS) < 3
    for k, ref in mms.CASE1_NORMS.items():
        assert abs(norms[k] - ref) / ref < 1e-9, (k, norms[k], ref)


def test_mms_source_terms_drive_the_discrete_steady_state(oracle):
    """Without the source terms the same run drifts far from the manufactured
    field: the sources are what holds it (a dropped or mis-signed source term
    fails this and the norm test)."""
    po = oracle
    layout, h = mms.block_layout(8, 1, 1)
    (_, _, ni, nj, origin), = layout
    gas = po.Gas.ideal(mMass=mms.MMASS, gamma=mms.GAMMA)
    errs = {}
    for with_src in (True, False):
        b = po.Block(gas, 2, ni, nj, 1, h, h)
        b.set_config(flux_calc=po.AUSMDV, update_scheme=po.PC, interp_order=2, apply_limiter=0, extrema_clipping=0)
        j, i = np.meshgrid(np.arange(nj), np.arange(ni), indexing="ij")
        xc, yc = (i.ravel() + 0.5) * h, (j.ravel() + 0.5) * h
        rho, p, T, vx, vy = mms.analytic_state(xc, yc)
        b.set_from_pT(*(a.reshape(b.shape) for a in (p, T, vx, vy)), 0.0)
        if with_src:
            b.set_source_terms(mms.cell_sources(ni, nj, h, origin))
        b.set_bc(mms.FACE_EAST, po.BC_EXTRAPOLATE_1)
        b.set_bc(mms.FACE_NORTH, po.BC_EXTRAPOLATE_1)
        b.set_bc(mms.FACE_WEST, po.BC_INFLOW_PROFILE, mms.ghost_states_for_face(ni, nj, h, mms.FACE_WEST, origin))
        b.set_bc(mms.FACE_SOUTH, po.BC_INFLOW_PROFILE, mms.ghost_states_for_face(ni, nj, h, mms.FACE_SOUTH, origin))
        for _ in range(400):
            rc, _ = b.step(2.0e-5)
            assert rc == 0
        errs[with_src] = np.max(np.abs(b.get_fs()[0] - rho))
    assert errs[True] < 0.02, errs
    assert errs[False] > 5 * errs[True], errs


def test_mms_second_order_convergence(oracle):
    """The density error falls at second order with grid refinement (the
    order-of-accuracy study the smoke test's norms come from)."""
    l2 = {}
    for n in (8, 16, 32):
        _, norms, _ = mms.run_oracle(oracle, ncells=n)
        l2[n] = norms["L2"]
    order = np.log(l2[16] / l2[32]) / np.log(2.0)
    assert 1.7 < order < 2.5, l2


def test_source_terms_add_to_dUdt(oracle):
    """A constant source on a uniform closed box (slip walls: zero flux
    divergence) changes U by dt * Q exactly in one Euler step."""
    po = oracle
    gas = po.Gas.ideal()
    b = po.Block(gas, 3, 4, 3, 2, 0.1, 0.1, 0.1)
    b.set_config(flux_calc=po.AUSMDV, update_scheme=po.EULER)
    b.set_from_pT(1.0e5, 300.0, 0.0, 0.0, 0.0)
    U0 = b.get_U()
    Q = np.zeros((b.nq, b.ncells))
    Q[0], Q[1], Q[4] = 0.5, -3.0, 1.0e4
    b.set_source_terms(Q)
    dt = 1.0e-6
    rc, _ = b.step(dt)
    assert rc == 0
    U1 = b.get_U()
    np.testing.assert_allclose(U1 - U0, dt * Q, rtol=1e-9, atol=1e-12 * np.abs(U0).max())


def test_stringent_cfl_signal(oracle):
    """fvcell.d:1642-1650: (max normal speed + a) / min length, against the
    per-direction maximum; closed form on a uniform state with dx != dy."""
    po = oracle
    gas = po.Gas.ideal()
    b = po.Block(gas, 2, 6, 5, 1, 0.02, 0.05)
    b.set_config(flux_calc=po.AUSMDV, update_scheme=po.PC)
    b.set_from_pT(1.0e5, 300.0, 100.0, 400.0, 0.0)
    a = np.sqrt(1.4 * (8.31451 / 0.02896) * 300.0)
    dt_plain, _, s_plain = b.determine_dt(1e-9, 0.5, False)
    b.set_stringent_cfl(True)
    dt_str, _, s_str = b.determine_dt(1e-9, 0.5, False)
    np.testing.assert_allclose(s_plain, max((100.0 + a) / 0.02, (400.0 + a) / 0.05), rtol=1e-12)
    np.testing.assert_allclose(s_str, (400.0 + a) / 0.02, rtol=1e-12)
    assert dt_str < dt_plain
