"""Synthetic initial fields of the benchmark workloads (primitive p, T, vel per
cell) and their boundary conditions.  Seeds are fixed so that every run,
CPU or GPU, starts from identical inputs.

planar_bump_2d   2D 128x64 cells on 1.0 m x 0.5 m, ideal air, uniform
                 p=100 kPa, T=300 K, Mach 2 in +x, plus a Gaussian bump
                 (+10 % density and pressure, radius 0.05 m) centred in the
                 left third; supersonic inflow west, extrapolated outflow
                 east, slip walls north/south.
shocked_box_3d   3D unit cube, n^3 cells, ideal air; smooth field from random
                 Fourier modes with 1 <= |k| <= 8 (amplitudes ~ |k|^-2) mapped
                 into density factor [0.5, 2], velocity perturbation
                 [-300, 300] m/s and pressure factor [0.5, 2], multiplied onto
                 a planar Mach-3 normal shock at x = 0.3 (Rankine-Hugoniot
                 jump from rho=1 kg/m^3, p=100 kPa); periodic y/z, supersonic
                 inflow west, extrapolated outflow east.
"""
import numpy as np

from .fvcore import BCType, Face

SEED_BUMP_2D = 20261201
SEED_BOX_3D = 20261202
SEED_TP_BOX_3D = 20261204


def planar_bump_2d(gas, ni=128, nj=64, Lx=1.0, Ly=0.5, seed=SEED_BUMP_2D):
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0.0, Lx / 3.0)
    cy = rng.uniform(0.0, Ly)
    dx, dy = Lx / ni, Ly / nj
    x = (np.arange(ni) + 0.5) * dx
    y = (np.arange(nj) + 0.5) * dy
    X, Y = np.meshgrid(x, y, indexing="xy")          # [nj][ni]
    p0, T0 = 100.0e3, 300.0
    a0 = float(np.sqrt(gas.gamma * gas.Rgas * T0))
    u0 = 2.0 * a0
    r2 = (X - cx) ** 2 + (Y - cy) ** 2
    bump = 0.1 * np.exp(-r2 / 0.05 ** 2)
    rho0 = p0 / (T0 * gas.Rgas)
    rho = rho0 * (1.0 + bump)
    p = p0 * (1.0 + bump)
    T = p / (rho * gas.Rgas)
    shape = (1, nj, ni)
    fields = dict(p=p.reshape(shape), T=T.reshape(shape), vx=np.full(shape, u0), vy=np.zeros(shape),
                  vz=np.zeros(shape))
    inflow = gas.flowstate(p0, T0, (u0, 0.0, 0.0))
    bcs = {Face.west: (BCType.inflow, inflow), Face.east: (BCType.extrapolate, None),
           Face.north: (BCType.slip_wall, None), Face.south: (BCType.slip_wall, None)}
    geom = dict(dims=2, ni=ni, nj=nj, nk=1, dx=dx, dy=dy, dz=1.0)
    return geom, fields, bcs, dict(centre=(cx, cy), a0=a0, u0=u0)


def _smooth_field(n, rng, kmax=8):
    """Real field on an n^3 periodic grid: sum of modes 1 <= |k| <= kmax with
    amplitude |k|^-2 and uniform random phase, scaled to [-1, 1]."""
    k = np.fft.fftfreq(n, d=1.0 / n)
    kz = np.fft.rfftfreq(n, d=1.0 / n)
    KX, KY, KZ = np.meshgrid(k, k, kz, indexing="ij")
    kk = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    mask = (kk >= 1.0) & (kk <= kmax)
    spec = np.zeros(KX.shape, dtype=np.complex128)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=int(mask.sum()))
    spec[mask] = (kk[mask] ** -2.0) * np.exp(1j * phase)
    f = np.fft.irfftn(spec, s=(n, n, n), axes=(0, 1, 2))   # indexed [x][y][z]
    f = f - f.min()
    f = 2.0 * f / f.max() - 1.0
    return np.ascontiguousarray(np.transpose(f, (2, 1, 0)))   # -> [k][j][i]


def shocked_box_3d(gas, n=256, seed=SEED_BOX_3D, x_shock=0.3, mach=3.0):
    rng = np.random.default_rng(seed)
    s_rho, s_p, s_u, s_v, s_w = (_smooth_field(n, rng) for _ in range(5))
    g = gas.gamma
    rho1, p1 = 1.0, 100.0e3
    a1 = float(np.sqrt(g * p1 / rho1))
    u1 = mach * a1
    # Rankine-Hugoniot ratios across a normal shock (the relations of gasdyn/idealgasflow.d:115-186)
    M2 = mach * mach
    r_rho = (g + 1.0) * M2 / ((g - 1.0) * M2 + 2.0)
    r_p = 1.0 + 2.0 * g / (g + 1.0) * (M2 - 1.0)
    rho2, p2, u2 = rho1 * r_rho, p1 * r_p, u1 / r_rho
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    up = (x < x_shock)[None, None, :]
    f_rho = 1.25 + 0.75 * s_rho          # [0.5, 2.0]
    f_p = 1.25 + 0.75 * s_p              # [0.5, 2.0]
    rho = np.where(up, rho1, rho2) * f_rho
    p = np.where(up, p1, p2) * f_p
    T = p / (rho * gas.Rgas)
    vx = np.where(up, u1, u2) + 300.0 * s_u
    vy = 300.0 * s_v
    vz = 300.0 * s_w
    fields = dict(p=p, T=T, vx=vx, vy=vy, vz=vz)
    T1 = p1 / (rho1 * gas.Rgas)
    inflow = gas.flowstate(p1, T1, (u1, 0.0, 0.0))
    bcs = {Face.west: (BCType.inflow, inflow), Face.east: (BCType.extrapolate, None),
           Face.north: (BCType.periodic, None), Face.south: (BCType.periodic, None),
           Face.top: (BCType.periodic, None), Face.bottom: (BCType.periodic, None)}
    geom = dict(dims=3, ni=n, nj=n, nk=n, dx=dx, dy=dx, dz=dx)
    return geom, fields, bcs, dict(u1=u1, u2=u2, rho2=rho2, p2=p2)


def initial_U(gas, fields):
    """Conserved state [5][ncells] of the primitive fields (FlowState from p, T;
    encode_conserved)."""
    p, T = fields["p"].ravel(), fields["T"].ravel()
    rho, u, _ = gas.rho_u_a_from_pT(p, T)
    return gas.encode(rho, u, fields["vx"], fields["vy"], fields["vz"])


def _smooth_unit(n, rng, kmax=8):
    return 0.5 * (_smooth_field(n, rng, kmax) + 1.0)          # [0, 1]


def thermally_perfect_box_3d(gas, n=128, seed=SEED_TP_BOX_3D):
    """3D n^3 unit cube of 5-species air (N2, O2, NO, N, O): temperature
    300-8000 K and density 0.01-1.0 kg/m^3 from smooth random modes, velocity
    +-300 m/s; mass fractions drawn from a Dirichlet distribution of
    concentration 200 about a composition that moves from N2 0.767 / O2 0.233
    toward dissociated air as (T - 300)/7700 rises.  Periodic in every
    direction.  Returns geometry, primitive fields (incl. Y [nsp][...]), BCs."""
    rng = np.random.default_rng(seed)
    T = 300.0 + 7700.0 * _smooth_unit(n, rng)
    rho = 0.01 + 0.99 * _smooth_unit(n, rng)
    vx, vy, vz = (300.0 * _smooth_field(n, rng) for _ in range(3))
    cold = np.array([0.767, 0.233, 0.0, 0.0, 0.0])
    hot = np.array([0.60, 0.0, 0.02, 0.148, 0.232])
    w = (T - 300.0) / 7700.0
    base = (1.0 - w)[None] * cold[:, None, None, None] + w[None] * hot[:, None, None, None]
    alpha = 200.0 * (base + 1.0e-3)
    g = rng.standard_gamma(alpha)
    Y = g / g.sum(axis=0, keepdims=True)
    Rm = gas.Rmix(Y)
    p = rho * Rm * T
    fields = dict(p=p, T=T, rho=rho, vx=vx, vy=vy, vz=vz, Y=Y)
    bcs = {f: (BCType.periodic, None) for f in (Face.west, Face.east, Face.north, Face.south, Face.top, Face.bottom)}
    geom = dict(dims=3, ni=n, nj=n, nk=n, dx=1.0 / n, dy=1.0 / n, dz=1.0 / n)
    return geom, fields, bcs, {}


def initial_U_tp(gas, fields):
    """Conserved state [5+nsp][ncells] and temperature guess of a thermally
    perfect field (FlowState from rho, T, Y; encode_conserved)."""
    T = fields["T"].ravel()
    Y = fields["Y"].reshape(gas.n_species, -1)
    u = gas.energy(T, Y)
    U = gas.encode(fields["rho"].ravel(), u, fields["vx"], fields["vy"], fields["vz"], Y)
    return U, T.copy()


SEED_MULTIBLOCK = 20261205
_IN_FACE = {0: (0, 2), 1: (1, 2), 2: (0, 2), 3: (1, 2), 4: (0, 1), 5: (0, 1)}   # face -> in-face axes (i,j,k)
_NORMAL = {0: 1, 1: 0, 2: 1, 3: 0, 4: 2, 5: 2}


def _smooth_block(shape, rng, nmodes=3, kmax=3):
    """Smooth field on a (nk, nj, ni) block scaled to [-1, 1]: a sum of
    separable products of short random Fourier series in x, y, z."""
    nk, nj, ni = shape
    f = np.zeros(shape)
    for _ in range(nmodes):
        parts = []
        for n in (nk, nj, ni):
            x = (np.arange(n) + 0.5) / n
            g = np.zeros(n)
            for k in range(1, kmax + 1):
                g += rng.normal() / k * np.cos(2.0 * np.pi * k * x + rng.uniform(0, 2.0 * np.pi))
            parts.append(g)
        f += parts[0][:, None, None] * parts[1][None, :, None] * parts[2][None, None, :]
    f -= f.min()
    return 2.0 * f / max(f.max(), 1e-300) - 1.0


def _block_tree(rng, nblocks, base_lo=(1, 1, 1), base_hi=(10, 6, 4)):
    """Random tree of face-matched connections: block sizes in units of 8
    cells, each new block joined by a random free face and orientation to a
    random free face of an earlier block, its in-face sizes taken from that
    face (swapped for odd orientations), its normal size drawn log-uniform."""
    def logu(lo, hi):
        return int(round(np.exp(rng.uniform(np.log(lo), np.log(hi + 0.999)) - 0.0)))
    dims = [[max(base_lo[a], min(base_hi[a], logu(base_lo[a], base_hi[a]))) for a in range(3)]]
    free = [set(range(6))]
    conns = []
    while len(dims) < nblocks:
        A = int(rng.choice([b for b in range(len(dims)) if free[b]]))
        fa = int(rng.choice(sorted(free[A])))
        fb = int(rng.integers(6))
        o = int(rng.integers(4))
        na = [dims[A][_IN_FACE[fa][0]], dims[A][_IN_FACE[fa][1]]]
        if o % 2 == 1:
            na = na[::-1]
        nb = [0, 0, 0]
        nb[_IN_FACE[fb][0]], nb[_IN_FACE[fb][1]] = na
        ax = _NORMAL[fb]
        nb[ax] = max(base_lo[ax], min(base_hi[ax], logu(base_lo[ax], base_hi[ax])))
        dims.append(nb)
        free.append(set(range(6)) - {fb})
        free[A].discard(fa)
        conns.append((A, fa, len(dims) - 1, fb, o))
    return dims, conns, free


def multiblock_c5(ideal_gas, tp_gas, seed=SEED_MULTIBLOCK, n_ideal=32, n_tp=16, target_cells=40.0e6):
    """48 uniform Cartesian blocks (32 ideal air, 16 five-species air), each
    class a random tree of face-matched full-face connections in all
    orientations (no connection crosses gas classes).  Base sizes are
    log-uniform in 8..80 x 8..48 x 8..32 cells (units of 8), scaled by one
    integer factor so that the total is near target_cells (factor 2 gives the
    16^3 .. 160x96x64 range).  Free faces: one supersonic inflow in the ideal
    class, otherwise extrapolated outflow or slip walls.  Fields are smooth per
    block (ideal: p 70-130 kPa, T 210-390 K, |v| <= 100 m/s; mixture: T
    300-8000 K, rho 0.01-1 kg/m^3, Dirichlet composition as in
    thermally_perfect_box_3d).  Returns a list of block descriptions
    (gas class, sizes, cell size, BCs, U0, Tguess) and the connections
    (block a, face a, block b, face b, orientation) over the global list."""
    rng = np.random.default_rng(seed)
    di, ci, fi = _block_tree(rng, n_ideal)
    dt_, ct, ft = _block_tree(rng, n_tp)
    base = sum(int(np.prod(d)) for d in di + dt_) * 512
    m = max(1, int(round((target_cells / base) ** (1.0 / 3.0))))
    blocks, conns = [], []
    for cls, dims, cn, free in (("ideal", di, ci, fi), ("tp", dt_, ct, ft)):
        off = len(blocks)
        for b, (ui, uj, uk) in enumerate(dims):
            ni, nj, nk = 8 * m * ui, 8 * m * uj, 8 * m * uk
            shape = (nk, nj, ni)
            bcs = {}
            for f in sorted(free[b]):
                bcs[Face(f)] = (BCType.extrapolate if rng.random() < 0.5 else BCType.slip_wall, None)
            if cls == "ideal":
                p = 1.0e5 * (1.0 + 0.3 * _smooth_block(shape, rng))
                T = 300.0 * (1.0 + 0.3 * _smooth_block(shape, rng))
                v = [100.0 * _smooth_block(shape, rng) for _ in range(3)]
                rho, u, _ = ideal_gas.rho_u_a_from_pT(p.ravel(), T.ravel())
                U0 = ideal_gas.encode(rho, u, *[x.ravel() for x in v])
                Tg = None
            else:
                T = 300.0 + 7700.0 * 0.5 * (_smooth_block(shape, rng) + 1.0)
                rho = 0.01 + 0.99 * 0.5 * (_smooth_block(shape, rng) + 1.0)
                v = [300.0 * _smooth_block(shape, rng) for _ in range(3)]
                cold = np.array([0.767, 0.233, 0.0, 0.0, 0.0])
                hot = np.array([0.60, 0.0, 0.02, 0.148, 0.232])
                w = (T - 300.0) / 7700.0
                basec = (1.0 - w)[None] * cold[:, None, None, None] + w[None] * hot[:, None, None, None]
                g = rng.standard_gamma(200.0 * (basec + 1.0e-3))
                Y = (g / g.sum(axis=0, keepdims=True)).reshape(tp_gas.n_species, -1)
                u = tp_gas.energy(T.ravel(), Y)
                U0 = tp_gas.encode(rho.ravel(), u, *[x.ravel() for x in v], Y)
                Tg = T.ravel().copy()
            blocks.append(dict(cls=cls, ni=ni, nj=nj, nk=nk, dx=1.0 / ni, dy=1.0 / nj, dz=1.0 / nk, bcs=bcs,
                               U0=U0, Tguess=Tg))
        for (a, fa, b, fb, o) in cn:
            conns.append((off + a, fa, off + b, fb, o))
    # one supersonic inflow face in the ideal class (Mach 2 air at 100 kPa, 300 K)
    for blk in blocks[:n_ideal]:
        free_w = [f for f, (k, _) in blk["bcs"].items() if f == Face.west]
        if free_w:
            a0 = float(np.sqrt(ideal_gas.gamma * ideal_gas.Rgas * 300.0))
            blk["bcs"][Face.west] = (BCType.inflow, ideal_gas.flowstate(1.0e5, 300.0, (2.0 * a0, 0.0, 0.0)))
            break
    return blocks, conns
