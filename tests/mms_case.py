"""Manufactured solution of the Euler equations used by the reference's 2D
manufactured-solution smoke tests (Roy, Nelson, Smith and Ober, "Verification
of Euler/Navier-Stokes codes using the method of manufactured solutions",
Int. J. Numer. Meth. Fluids 44 (2004) 599-620, Appendix A).

The case set-up follows examples/eilmer/2D/manufactured-solution/sg/smoke-tests:
  mms.lua                  16 x 16 cells on the unit square as 2 x 2 blocks,
                           south/west UserDefinedBC (ghost cells from the
                           analytic field), north/east OutFlowBC_SimpleExtrapolate
                           xOrder=1, ausmdv, pc, x_order 2, no limiter, no
                           extrema clipping, cfl 0.5, stringent_cfl, dt_init
                           1e-6, max_time 60 ms, udf_source_terms
  analytic_solution.py     the fields and their constants (cases 1 and 3:
                           supersonic; case 3 with the Gaussian envelope S)
  constants.txt            R = 287, gamma = 1.4
  make_lua_files.py        sources = divergence of the Euler fluxes (mu = k = 0)
  mms-euler-test.rb        the known answers: about 2620 steps and the density
                           error norms L1, L2, Linf at t = 60 ms, to 1e-3.

The sources are differentiated here by hand (product and chain rules) rather
than with a computer-algebra system; a test checks them against sympy.
"""
import numpy as np

L = 1.0
RGAS = 287.0
GAMMA = 1.4
# the ideal gas of the case as the gas model builds it: R = R_universal / mMass
MMASS = 8.31451 / RGAS
R_MODEL = 8.31451 / MMASS
CV_MODEL = R_MODEL / (GAMMA - 1.0)

# the reference's known answers for case 1 (mms-euler-test.rb)
CASE1_STEPS = 2620
CASE1_NORMS = {"L1": 5.992287970340e-04, "L2": 8.134693405237e-04, "Linf": 2.746228919538e-03}

# supersonic constants (analytic_solution.py, cases 1 and 3)
SUPERSONIC = dict(rho0=1.0, rhox=0.15, rhoy=-0.1, rhoxy=0.0, arhox=1.0, arhoy=0.5, arhoxy=0.0,
                  u0=800.0, ux=50.0, uy=-30.0, uxy=0.0, aux=1.5, auy=0.6, auxy=0.0,
                  v0=800.0, vx=-75.0, vy=40.0, vxy=0.0, avx=0.5, avy=2.0 / 3, avxy=0.0,
                  p0=1.0e5, px=0.2e5, py=0.5e5, pxy=0.0, apx=2.0, apy=1.0, apxy=0.0)


def _term(c, fn, arg, a, x, y):
    """c * fn(a pi w) with w = x/L, y/L or x y / L^2, and its x and y derivatives."""
    w = {"x": x / L, "y": y / L, "xy": x * y / (L * L)}[arg]
    dw_dx = {"x": 1.0 / L, "y": 0.0, "xy": y / (L * L)}[arg]
    dw_dy = {"x": 0.0, "y": 1.0 / L, "xy": x / (L * L)}[arg]
    th = a * np.pi * w
    if fn == "sin":
        f, df = np.sin(th), np.cos(th)
    else:
        f, df = np.cos(th), -np.sin(th)
    return c * f, c * df * a * np.pi * dw_dx, c * df * a * np.pi * dw_dy


def _envelope(case, x, y):
    if case in (1, 2):
        one = np.ones_like(np.asarray(x, dtype=np.float64))
        return one, 0.0 * one, 0.0 * one
    S = np.exp(-16.0 * ((x - L / 2) ** 2 + (y - L / 2) ** 2) / (L * L))
    return S, S * (-32.0 * (x - L / 2) / (L * L)), S * (-32.0 * (y - L / 2) / (L * L))


def fields(x, y, case=1):
    """rho, u, v, p at (x, y) and their x and y derivatives: {name: (f, f_x, f_y)}."""
    c = SUPERSONIC
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    S, Sx, Sy = _envelope(case, x, y)
    spec = {
        "rho": (c["rho0"], [(c["rhox"], "sin", "x", c["arhox"]), (c["rhoy"], "cos", "y", c["arhoy"]),
                            (c["rhoxy"], "cos", "xy", c["arhoxy"])]),
        "u": (c["u0"], [(c["ux"], "sin", "x", c["aux"]), (c["uy"], "cos", "y", c["auy"]),
                        (c["uxy"], "cos", "xy", c["auxy"])]),
        "v": (c["v0"], [(c["vx"], "cos", "x", c["avx"]), (c["vy"], "sin", "y", c["avy"]),
                        (c["vxy"], "cos", "xy", c["avxy"])]),
        "p": (c["p0"], [(c["px"], "cos", "x", c["apx"]), (c["py"], "sin", "y", c["apy"]),
                        (c["pxy"], "sin", "xy", c["apxy"])]),
    }
    out = {}
    for name, (f0, terms) in spec.items():
        g = gx = gy = 0.0
        for t in terms:
            v, vx, vy = _term(*t, x, y)
            g, gx, gy = g + v, gx + vx, gy + vy
        out[name] = (f0 + S * g, Sx * g + S * gx, Sy * g + S * gy)
    return out


def source_terms(x, y, case=1):
    """Steady Euler sources: the divergence of the inviscid fluxes of the
    manufactured field (mass, x-momentum, y-momentum, total energy)."""
    F = fields(x, y, case)
    r, rx, ry = F["rho"]
    u, ux, uy = F["u"]
    v, vx, vy = F["v"]
    p, px, py = F["p"]
    fmass = rx * u + r * ux + ry * v + r * vy
    fxmom = rx * u * u + 2.0 * r * u * ux + px + ry * u * v + r * uy * v + r * u * vy
    fymom = rx * u * v + r * ux * v + r * u * vx + ry * v * v + 2.0 * r * v * vy + py
    # rho * et = p / (gamma - 1) + rho (u^2 + v^2) / 2
    E = p / (GAMMA - 1.0) + 0.5 * r * (u * u + v * v)
    Ex = px / (GAMMA - 1.0) + 0.5 * rx * (u * u + v * v) + r * (u * ux + v * vx)
    Ey = py / (GAMMA - 1.0) + 0.5 * ry * (u * u + v * v) + r * (u * uy + v * vy)
    fe = ux * (E + p) + u * (Ex + px) + vy * (E + p) + v * (Ey + py)
    return fmass, fxmom, fymom, fe


def analytic_state(x, y, case=1):
    """rho, p, T, vel.x, vel.y of the reference solution (T = p / (rho R))."""
    F = fields(x, y, case)
    rho, p = F["rho"][0], F["p"][0]
    return rho, p, p / (rho * RGAS), F["u"][0], F["v"][0]


def ghost_state(x, y, case=1):
    """The flow state the reference's UserDefinedBC gives a ghost cell at
    (x, y): p and T from the field, rho and u from update_thermo_from_pT,
    the sound speed, velocity; [rho, p, T, u, a, vx, vy, vz, S, massf]."""
    _, p, T, vx, vy = analytic_state(x, y, case)
    rho = p / (T * R_MODEL)
    return np.array([rho, p, T, CV_MODEL * T, np.sqrt(GAMMA * R_MODEL * T), vx, vy, 0.0, 0.0, 1.0])


# block layout of the case: FluidBlockArray nib = njb = 2 over the unit square
FACE_NORTH, FACE_EAST, FACE_SOUTH, FACE_WEST = 0, 1, 2, 3


def block_layout(ncells=16, nib=2, njb=2):
    """[(ib, jb, ni, nj, origin)] of the block array, and dx = dy."""
    h = L / ncells
    ni, nj = ncells // nib, ncells // njb
    return [(ib, jb, ni, nj, (ib * ni * h, jb * nj * h, 0.0)) for jb in range(njb) for ib in range(nib)], h


def ghost_states_for_face(ni, nj, h, face, origin, case=1):
    """[2][face cells][10] ghost states of a south or west face from the
    analytic field at the ghost-cell centres (layer-major, in-face index i
    for south, j for west)."""
    out = []
    for layer in (1, 2):
        row = []
        n = ni if face == FACE_SOUTH else nj
        for a in range(n):
            if face == FACE_SOUTH:
                x, y = origin[0] + (a + 0.5) * h, origin[1] - (layer - 0.5) * h
            else:
                x, y = origin[0] - (layer - 0.5) * h, origin[1] + (a + 0.5) * h
            row.append(ghost_state(x, y, case))
        out.append(row)
    return np.array(out)


def cell_sources(ni, nj, h, origin, case=1):
    """Q [5][cells] at the cell centres of a block (mass, x/y/z momentum, energy)."""
    j, i = np.meshgrid(np.arange(nj), np.arange(ni), indexing="ij")
    x = origin[0] + (i.ravel() + 0.5) * h
    y = origin[1] + (j.ravel() + 0.5) * h
    fm, fx, fy, fe = source_terms(x, y, case)
    return np.stack([fm, fx, fy, np.zeros_like(fm), fe])


def density_norms(blocks_rho, layout, h, case=1):
    """Volume-weighted norms of rho - rho_exact over all cells
    (compute_volume_weighted_norms, flowsolution.d:265-307)."""
    err = []
    for rho, (ib, jb, ni, nj, origin) in zip(blocks_rho, layout):
        j, i = np.meshgrid(np.arange(nj), np.arange(ni), indexing="ij")
        x = origin[0] + (i.ravel() + 0.5) * h
        y = origin[1] + (j.ravel() + 0.5) * h
        err.append(np.asarray(rho).ravel() - analytic_state(x, y, case)[0])
    e = np.concatenate(err)
    vol = h * h
    vsum = vol * e.size
    return {"L1": float(np.sum(vol * np.abs(e)) / vsum), "L2": float(np.sqrt(np.sum(vol * e * e) / vsum)),
            "Linf": float(np.max(np.abs(e)))}


def run_oracle(po, case=1, ncells=16, max_time=60.0e-3, nib=2, njb=2, decode_initial=False):
    """The case on the oracle: steps taken and the density error norms.
    decode_initial: start from decode(encode(initial flow state)), as a device
    block does after its upload, instead of the flow state itself (as the
    reference's solver does after reading its initial flow file); the two
    differ by rounding."""
    gas = po.Gas.ideal(mMass=MMASS, gamma=GAMMA)
    layout, h = block_layout(ncells, nib, njb)
    blks = {}
    for (ib, jb, ni, nj, origin) in layout:
        b = po.Block(gas, 2, ni, nj, 1, h, h)
        b.set_config(flux_calc=po.AUSMDV, update_scheme=po.PC, interp_order=2, apply_limiter=0,
                     extrema_clipping=0)
        b.set_stringent_cfl(True)
        b.set_from_pT(1.0e5, 1.0e5 / RGAS, 800.0, 800.0, 0.0)
        if decode_initial:
            b.set_U(b.get_U())
        b.set_source_terms(cell_sources(ni, nj, h, origin, case))
        if ib == nib - 1:
            b.set_bc(FACE_EAST, po.BC_EXTRAPOLATE_1)
        if jb == njb - 1:
            b.set_bc(FACE_NORTH, po.BC_EXTRAPOLATE_1)
        if ib == 0:
            b.set_bc(FACE_WEST, po.BC_INFLOW_PROFILE, ghost_states_for_face(ni, nj, h, FACE_WEST, origin, case))
        if jb == 0:
            b.set_bc(FACE_SOUTH, po.BC_INFLOW_PROFILE, ghost_states_for_face(ni, nj, h, FACE_SOUTH, origin, case))
        blks[(ib, jb)] = b
    for (ib, jb), b in blks.items():
        if ib + 1 < nib:
            po.connect_blocks(b, FACE_EAST, blks[(ib + 1, jb)], FACE_WEST)
        if jb + 1 < njb:
            po.connect_blocks(b, FACE_NORTH, blks[(ib, jb + 1)], FACE_SOUTH)
    order = [blks[(ib, jb)] for (ib, jb, *_rest) in layout]
    # simcore.d:1347-1452 over all blocks: cfl 0.5, cfl_count 10, dt_max 1e-3
    dt_global, time, step = 1.0e-6, 0.0, 0
    while True:
        if step % 10 == 0:
            res = [b.determine_dt(dt_global, 0.5, step > 0) for b in order]
            dt_allow = min(r[0] for r in res)
            if step == 0:
                dt_allow = min(1.0e-6, dt_allow)
            if dt_allow <= dt_global:
                dt_global = dt_allow
            else:
                dt_global = min(min(dt_global * 1.5, dt_allow), 1.0e-3)
        rc, nbad = po.multi_step(order, dt_global)
        if rc != 0:
            raise RuntimeError("oracle step failed rc=%d" % rc)
        time += dt_global
        step += 1
        if time >= max_time:
            break
    rho = [b.get_fs()[0] for b in order]
    return step, density_norms(rho, layout, h, case), order
