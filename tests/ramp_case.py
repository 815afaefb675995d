"""The reference's 3D inviscid ramp case (examples/eilmer/3D/simple-ramp/sg/
ramp.lua, ramp-test.rb:21-53): Mach 1.5 air over a 10-degree ramp, two
blocks of hexahedral cells clustered toward the ramp surface, euler update,
default flux calculator (adaptive_hanel_ausmdv), CFL 0.5, dt_init 1e-6, run to
t = 5 ms.  The reference takes 862 +- 3 steps and estimates a shock angle of
57 +- 1 degrees (estimate_shock_angle.lua).

Grid construction restates the reference's geometry package for this case:
RobertsFunction (geom/misc/univariatefunctions.d:66-137), the parameter
distribution along edges (univariatefunctions.d:24-35), the straight-edged
Coons patches of a TFIVolume built from eight corners (volume/tfivolume.d:90-
143) and StructuredGrid.make_grid_from_volume (grid/sgrid.d:481-523).
"""
import numpy as np

import eilmer4_amd as e4
from eilmer4_amd.fvcore import BCType, Face


def roberts(t, end0, end1, beta):
    """RobertsFunction(end0, end1, beta)(t)."""
    cluster = (end0 or end1) and beta > 1.0
    alpha, reverse = 0.5, False
    if end0 and not end1:
        reverse, alpha = True, 0.0
    if end1 and not end0:
        reverse, alpha = False, 0.0
    t = np.asarray(t, dtype=np.float64)
    if reverse:
        t = 1.0 - t
    if cluster:
        lam = ((beta + 1.0) / (beta - 1.0)) ** ((t - alpha) / (1.0 - alpha))
        tbar = ((beta + 2.0 * alpha) * lam - beta + 2.0 * alpha) / ((2.0 * alpha + 1.0) * (1.0 + lam))
    else:
        tbar = t
    if reverse:
        tbar = 1.0 - tbar
    return tbar


def distribute(n, f=None):
    t = np.arange(n) * (1.0 / (n - 1))
    tv = t if f is None else f(t)
    return (1.0 - tv) * 0.0 + tv * 1.0


def box_corners(xPos=0.0, yPos=0.0, zPos=0.0, xSize=1.0, ySize=1.0, zSize=1.0):
    return np.array([[xPos, yPos, zPos], [xPos + xSize, yPos, zPos], [xPos + xSize, yPos + ySize, zPos],
                     [xPos, yPos + ySize, zPos], [xPos, yPos, zPos + zSize], [xPos + xSize, yPos, zPos + zSize],
                     [xPos + xSize, yPos + ySize, zPos + zSize], [xPos, yPos + ySize, zPos + zSize]])


def _coons(p00, p10, p11, p01, r, s):
    """Coons patch on straight edges (surface/coonspatch.d): the boundary
    lines blended less the bilinear corners."""
    south = (1 - r) * p00 + r * p10
    north = (1 - r) * p01 + r * p11
    west = (1 - s) * p00 + s * p01
    east = (1 - s) * p10 + s * p11
    return ((1 - s) * south + s * north + (1 - r) * west + r * east -
            ((1 - r) * (1 - s) * p00 + (1 - r) * s * p01 + r * (1 - s) * p10 + r * s * p11))


def tfi_grid(p, niv, njv, nkv, cluster_k=None):
    """make_grid_from_volume of TFIVolume(vertices=p) with the clustering
    function cluster_k on the four k-edges (edge04, 15, 26, 37) and linear
    distributions elsewhere: vertices [nkv][njv][niv][3]."""
    r01 = r32 = r45 = r76 = distribute(niv)
    s12 = s03 = s56 = s47 = distribute(njv)
    t04 = t15 = t26 = t37 = distribute(nkv, cluster_k)
    k = np.arange(nkv)[:, None, None] / (nkv - 1)
    j = np.arange(njv)[None, :, None] / (njv - 1)
    i = np.arange(niv)[None, None, :] / (niv - 1)
    t, s, r = k, j, i
    tK, sJ, rI = (t04[:, None, None], s03[None, :, None], r01[None, None, :])
    tdash = (1 - r) * (1 - s) * tK + r * s * t26[:, None, None] + (1 - s) * r * t15[:, None, None] + \
        s * (1 - r) * t37[:, None, None]
    sdash = (1 - t) * (1 - r) * sJ + t * r * s56[None, :, None] + (1 - t) * r * s12[None, :, None] + \
        t * (1 - r) * s47[None, :, None]
    rdash = (1 - s) * (1 - t) * rI + s * t * r76[None, None, :] + (1 - s) * t * r45[None, None, :] + \
        s * (1 - t) * r32[None, None, :]
    R, S, T = (rdash[..., None], sdash[..., None], tdash[..., None])
    P = [p[q][None, None, None, :] for q in range(8)]
    pW = _coons(P[0], P[3], P[7], P[4], S, T)
    pE = _coons(P[1], P[2], P[6], P[5], S, T)
    pS = _coons(P[0], P[1], P[5], P[4], R, T)
    pN = _coons(P[3], P[2], P[6], P[7], R, T)
    pB = _coons(P[0], P[1], P[2], P[3], R, S)
    pT = _coons(P[4], P[5], P[6], P[7], R, S)
    omr, oms, omt = 1.0 - R, 1.0 - S, 1.0 - T
    big = (omr * oms * omt * P[0] + omr * oms * T * P[4] + omr * S * omt * P[3] + omr * S * T * P[7] +
           R * oms * omt * P[1] + R * oms * T * P[5] + R * S * omt * P[2] + R * S * T * P[6])
    return 0.5 * (omr * pW + R * pE + oms * pS + S * pN + omt * pB + T * pT) - 0.5 * big


def grids():
    ck = lambda t: roberts(t, True, False, 1.2)
    g0 = tfi_grid(box_corners(xSize=0.2, ySize=0.1), 11, 5, 41, ck)
    c1 = box_corners(xPos=0.2, xSize=0.8, ySize=0.1)
    c1[1, 2] = 0.8 * np.tan(np.pi * 10.0 / 180.0)
    c1[2, 2] = c1[1, 2]
    g1 = tfi_grid(c1, 31, 5, 41, ck)
    return g0, g1


GAS = dict(mMass=0.02896, gamma=1.4)
INITIAL = dict(p=5955.0, T=304.0, vel=(0.0, 0.0, 0.0))
INFLOW = dict(p=95.84e3, T=1103.0, vel=(1000.0, 0.0, 0.0))
MAX_TIME = 5.0e-3
DT_INIT = 1.0e-6


def cell_centres(g):
    """Mean of each hexahedron's eight vertices [nk][nj][ni][3]."""
    return 0.125 * (g[:-1, :-1, :-1] + g[:-1, :-1, 1:] + g[:-1, 1:, :-1] + g[:-1, 1:, 1:] +
                    g[1:, :-1, :-1] + g[1:, :-1, 1:] + g[1:, 1:, :-1] + g[1:, 1:, 1:])


def shock_angle(pos, p):
    """estimate_shock_angle.lua: along every (j, k) strip of cells through
    both blocks (cells in i order), the position where p first passes
    p_first + 0.3 (p_max - p_first), interpolated; strips with x < 0.65 kept;
    a least-squares line z(x); its angle in degrees and the mean deviation."""
    xs_shock, zs_shock = [], []
    nk, nj = p.shape[0], p.shape[1]
    for k in range(nk):
        for j in range(nj):
            xs, ys, ps = pos[k, j, :, 0], pos[k, j, :, 2], p[k, j, :]
            p_trigger = ps[0] + 0.3 * (ps.max() - ps[0])
            x_old, y_old, p_old = xs[0], ys[0], ps[0]
            x_new, y_new, p_new = x_old, y_old, p_old
            for q in range(1, len(ps)):
                x_new, y_new, p_new = xs[q], ys[q], ps[q]
                if p_new > p_trigger:
                    break
                x_old, y_old, p_old = x_new, y_new, p_new
            frac = (p_trigger - p_old) / (p_new - p_old)
            x_loc = x_old * (1.0 - frac) + x_new * frac
            y_loc = y_old * (1.0 - frac) + y_new * frac
            if x_loc < 0.65:
                xs_shock.append(x_loc)
                zs_shock.append(y_loc)
    x, y = np.array(xs_shock), np.array(zs_shock)
    N = len(x)
    a1 = (np.sum(x * y) / N - np.sum(x) / N * np.sum(y) / N) / (np.sum(x * x) / N - (np.sum(x) / N) ** 2)
    a0 = np.sum(y) / N - a1 * np.sum(x) / N
    return float(np.degrees(np.arctan(a1))), float(np.mean(np.abs(a0 + a1 * x - y)))


def boundary_conditions(gas_flowstate):
    """Per block: {face: (BCType, flow state)}; the faces left out are slip
    walls (prep.lua default); blocks 0 east and 1 west are connected."""
    inflow = gas_flowstate(INFLOW["p"], INFLOW["T"], INFLOW["vel"])
    return [{Face.west: (BCType.inflow, inflow)}, {Face.east: (BCType.outflow_simple_flux, None)}]
