"""e4hip_determine_dt_noalloc (the CFL reduction without device allocation)
against e4hip_determine_dt, through the C ABI: bit-identical dt_allow and
cfl_max on two calls in a row, the second with fewer and smaller blocks and
with the cfl_allow rescale taken; more blocks than the fixed scratch holds are
refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import eilmer4_amd as e4
from eilmer4_amd import workloads
from eilmer4_amd.fvcore import BCType, Face

pytestmark = pytest.mark.gpu


def _dt_pair(L, ids, dt_current, cfl, check):
    arr = (C.c_int * len(ids))(*ids)
    a1, c1 = C.c_double(), C.c_double()
    a2, c2 = C.c_double(), C.c_double()
    assert L.e4hip_determine_dt(arr, len(ids), dt_current, cfl, check, C.byref(a1), C.byref(c1)) == 0
    assert L.e4hip_determine_dt_noalloc(arr, len(ids), dt_current, cfl, check, C.byref(a2), C.byref(c2)) == 0
    return (a1.value, c1.value), (a2.value, c2.value)


def test_determine_dt_noalloc_matches_determine_dt(gpu):
    L = e4.load()
    gas = e4.IdealGas()
    blocks = []
    # uniform 3D ideal block (shocked box)
    geom, fields, bcs, _ = workloads.shocked_box_3d(gas, n=24, seed=5)
    b3 = e4.SFluidBlock(gas, 3, geom["ni"], geom["nj"], geom["nk"], geom["dx"], geom["dy"], geom["dz"])
    for face, (kind, fs) in bcs.items():
        b3.set_bc(face, kind, fs)
    b3.upload(workloads.initial_U(gas, fields))
    blocks.append(b3)
    # 2D ideal block
    geom2, fields2, bcs2, _ = workloads.planar_bump_2d(gas)
    b2 = e4.SFluidBlock(gas, geom2["dims"], geom2["ni"], geom2["nj"], geom2["nk"], geom2["dx"], geom2["dy"], geom2["dz"])
    for face, (kind, fs) in bcs2.items():
        b2.set_bc(face, kind, fs)
    b2.upload(workloads.initial_U(gas, fields2))
    blocks.append(b2)
    # five-species block (primitive-array CFL kernel)
    tg = e4.ThermallyPerfectGas.from_json()
    geomt, fieldst, bcst, _ = workloads.thermally_perfect_box_3d(tg, n=12)
    bt = e4.SFluidBlock(tg, 3, 12, 12, 12, geomt["dx"], geomt["dy"], geomt["dz"])
    for face, (kind, fs) in bcst.items():
        bt.set_bc(face, kind, fs)
    U0, Tg = workloads.initial_U_tp(tg, fieldst)
    bt.upload(U0, Tg)
    blocks.append(bt)
    # block built from its grid (face-normal CFL kernel)
    from test_oracle_geometry import curvy_grid, _flow
    g = curvy_grid(10, 8, 6, seed=3)
    bg = e4.SFluidBlock.from_grid(gas, g)
    p, T, v = _flow((6, 8, 10), 4)
    rho, u, _ = gas.rho_u_a_from_pT(p.ravel(), T.ravel())
    bg.upload(gas.encode(rho, u, *[x.ravel() for x in v]))
    blocks.append(bg)
    sim = e4.Simulation(blocks, e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv))
    ids = [b.id for b in blocks]
    print("case noalloc dt: 4 blocks (3D, 2D, mixture, grid)", flush=True)
    ref, got = _dt_pair(L, ids, 1.0e-6, 0.5, 0)
    print("  determine_dt %r noalloc %r" % (ref, got), flush=True)
    assert ref == got and np.isfinite(ref[0]) and ref[0] > 0.0
    # second call: fewer, smaller blocks, dt_current far above the limit so
    # that the cfl_allow rescale with the last cell's signal is taken
    print("case noalloc dt: 2 smaller blocks, rescale", flush=True)
    ref2, got2 = _dt_pair(L, [bt.id, bg.id], 1.0e-2, 0.5, 1)
    print("  determine_dt %r noalloc %r" % (ref2, got2), flush=True)
    assert ref2 == got2 and ref2[1] == 0.5 * 1.2
    # more blocks than the scratch holds: refused on the host
    too_many = (C.c_int * 4097)(*([bg.id] * 4097))
    a, c = C.c_double(), C.c_double()
    assert L.e4hip_determine_dt_noalloc(too_many, 4097, 1.0e-6, 0.5, 0, C.byref(a), C.byref(c)) == -1
    assert "at most 4096" in e4.lib.last_error()
    for b in blocks:
        b.destroy()
