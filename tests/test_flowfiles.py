"""Eilmer4's native grid and flow files (gziptext and rawbinary) on the CPU.

The fixtures tests/golden/MHD-Blunt-Nose.{grid,flow}.b0000.t0000.gz are the
reference's own example files (examples/eilmer/2D/mhd-blunt-nose/WithB): a
75 x 75 ideal-air block written by the reference (with MHD variables and the
older names e[0], T[0], k[0]).  The checks tie every column to its meaning
through the gas relations, and the grid to the flow through the cell
positions, so a shifted column or a transposed index fails."""
import gzip
import os
import struct

import numpy as np
import pytest

import eilmer4_amd as e4
from eilmer4_amd import flowfiles as ff

HERE = os.path.dirname(os.path.abspath(__file__))
GRID = os.path.join(HERE, "golden", "MHD-Blunt-Nose.grid.b0000.t0000.gz")
FLOW = os.path.join(HERE, "golden", "MHD-Blunt-Nose.flow.b0000.t0000.gz")


def test_reference_grid_file():
    g = ff.read_grid(GRID)
    assert g.dimensions == 2 and g.vertices.shape == (1, 76, 76, 3) and g.label == ""
    # the first two vertex lines of the file
    assert np.array_equal(g.vertices[0, 0, 0], [-1.1430000000000001e-01, 0.0, 0.0])
    assert np.array_equal(g.vertices[0, 0, 1], [-1.1090014188888889e-01, 0.0, 0.0])
    with pytest.raises(ValueError):
        g.cartesian_spacing()          # a body-fitted grid


def test_reference_flow_file_columns():
    f = ff.read_flow(FLOW)
    assert (f.nic, f.njc, f.nkc, f.dimensions) == (75, 75, 1, 2) and len(f.variables) == 25
    rho, p, a = f.get("rho"), f.get("p"), f.get("a")
    T, e = f.get("T", "T[0]"), f.get("u", "e[0]")
    gas = e4.IdealGas()                # ideal air: R = 8.31451/0.02896, gamma 1.4
    assert np.max(np.abs(p / (rho * T) / gas.Rgas - 1.0)) < 1e-14
    assert np.max(np.abs(a * a * rho / p / gas.gamma - 1.0)) < 1e-14
    assert np.max(np.abs(e / T / gas.Cv - 1.0)) < 1e-14
    assert np.all(f.get("massf[0]-air") == 1.0)
    assert set(np.unique(f.get("S"))) <= {0.0, 1.0}


def test_reference_flow_positions_lie_in_their_grid_cells():
    """Cell (i, j) of the flow file (i fastest) sits inside the quad of
    vertices (i..i+1, j..j+1) of the grid file."""
    g = ff.read_grid(GRID).vertices[0]
    f = ff.read_flow(FLOW)
    x = f.get("pos.x").reshape(75, 75)
    y = f.get("pos.y").reshape(75, 75)
    quad = np.stack([g[:-1, :-1], g[:-1, 1:], g[1:, 1:], g[1:, :-1]])     # [4][j][i][3]
    lo, hi = quad.min(axis=0), quad.max(axis=0)
    assert np.all((x >= lo[..., 0]) & (x <= hi[..., 0]) & (y >= lo[..., 1]) & (y <= hi[..., 1]))


def test_variable_list_matches_the_reference_file():
    """build_flow_variable_list for ideal air is the reference file's list
    without its MHD variables, under the current names of u, T and k."""
    f = ff.read_flow(FLOW)
    rename = {"k[0]": "k", "e[0]": "u", "T[0]": "T"}
    ref = [rename.get(v, v) for v in f.variables if v not in ("B.x", "B.y", "B.z", "divB", "psi")]
    assert ff.flow_variable_list(e4.IdealGas()) == ref
    tp = ff.flow_variable_list(e4.ThermallyPerfectGas.from_json())
    assert tp[17:] == ["massf[0]-N2", "massf[1]-O2", "massf[2]-NO", "massf[3]-N", "massf[4]-O", "dt_chem", "u", "T"]


@pytest.mark.parametrize("fmt", ["gziptext", "rawbinary"])
def test_round_trips(tmp_path, fmt):
    g = ff.read_grid(GRID)
    ff.write_grid(tmp_path / "g", g, fmt)
    g2 = ff.read_grid(tmp_path / "g")
    assert np.array_equal(g2.vertices, g.vertices) and g2.dimensions == g.dimensions
    f = ff.read_flow(FLOW)
    f.sim_time = 1.2345678901234567e-3
    f.label = "blk-0"
    ff.write_flow(tmp_path / "f", f, fmt)
    f2 = ff.read_flow(tmp_path / "f")
    assert f2.variables == f.variables and f2.sim_time == f.sim_time and f2.label == "blk-0"
    assert (f2.nic, f2.njc, f2.nkc, f2.dimensions) == (f.nic, f.njc, f.nkc, f.dimensions)
    for n in f.variables:
        assert np.array_equal(f2.values[n], f.values[n]), n


def test_rawbinary_layout():
    """The raw-binary layouts of sgrid.d:717-740 and sfluidblock.d:1908-1930,
    built here byte by byte: header text, int32 label length and label, then
    int32 dimensions and counts (flow: double sim_time and the variable
    names first), then float64 records with i fastest."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        g = ff.cartesian_grid(2, 2, 1, 0, 0.5, 0.25, label="ab")
        ff.write_grid(os.path.join(d, "g"), g, "rawbinary")
        want = b"structured_grid 1.0" + struct.pack("<i", 2) + b"ab" + struct.pack("<4i", 2, 3, 2, 1)
        for j in range(2):
            for i in range(3):
                want += struct.pack("<3d", 0.5 * i, 0.25 * j, 0.0)
        assert open(os.path.join(d, "g"), "rb").read() == want
        fl = ff.FlowSolution(["rho", "S"], {"rho": np.array([1.5, 2.5]), "S": np.array([0.0, 1.0])}, 2, 2, 1, 1,
                             0.75, "")
        ff.write_flow(os.path.join(d, "f"), fl, "rawbinary")
        want = (b"structured_grid_flow 1.0" + struct.pack("<i", 0) + struct.pack("<d", 0.75) + struct.pack("<i", 2)
                + struct.pack("<i", 3) + b"rho" + struct.pack("<i", 1) + b"S" + struct.pack("<4i", 2, 2, 1, 1)
                + struct.pack("<4d", 1.5, 0.0, 2.5, 1.0))
        assert open(os.path.join(d, "f"), "rb").read() == want
        ff.write_flow(os.path.join(d, "t"), fl, "gziptext")
        text = gzip.open(os.path.join(d, "t"), "rt").read().splitlines()
        assert text[:9] == ["structured_grid_flow 1.0", "label: ", "sim_time: 7.500000000000000000e-01",
                            "variables: 2", ' "rho" "S"', "dimensions: 2", "nicell: 2", "njcell: 1", "nkcell: 1"]
        assert text[9] == " 1.500000000000000000e+00 0"


def test_cartesian_grid_spacing():
    g = ff.cartesian_grid(3, 4, 3, 2, 0.1, 0.2, 0.3, origin=(1.0, -2.0, 0.5))
    dx, dy, dz, o = g.cartesian_spacing()
    assert abs(dx - 0.1) < 1e-15 and abs(dy - 0.2) < 1e-15 and abs(dz - 0.3) < 1e-15
    assert np.array_equal(o, [1.0, -2.0, 0.5])


def test_conserved_from_reference_flow():
    """encode_conserved of the file's states decodes back to them."""
    f = ff.read_flow(FLOW)
    gas = e4.IdealGas()
    U, Tg = ff.conserved_from_flow(f, gas)
    assert Tg is None and U.shape == (5, 75 * 75)
    rho = U[0]
    assert np.array_equal(rho, f.get("rho"))
    assert np.max(np.abs(U[1] / rho - f.get("vel.x"))) <= 1e-12 * np.abs(f.get("vel.x")).max()
    ke = 0.5 * (f.get("vel.x") ** 2 + f.get("vel.y") ** 2 + f.get("vel.z") ** 2)
    assert np.max(np.abs(U[4] / rho - ke - f.get("e[0]")) / f.get("e[0]")) < 1e-12


def _profile_flow(gas, ny):
    """A column of ny cells at x = 0.5, y = (j + 0.5)/ny varying in y."""
    y = (np.arange(ny) + 0.5) / ny
    p = 1.0e5 * (1.0 + 0.2 * y)
    T = 300.0 + 100.0 * y
    rho, u, a = gas.rho_u_a_from_pT(p, T)
    zero = np.zeros(ny)
    v = {"pos.x": np.full(ny, 0.5), "pos.y": y, "pos.z": zero, "volume": np.full(ny, 1.0 / ny), "rho": rho,
         "vel.x": 600.0 + 50.0 * y, "vel.y": 10.0 * y, "vel.z": zero, "p": p, "a": a, "mu": zero, "k": zero,
         "mu_t": zero, "k_t": zero, "S": (np.arange(ny) % 3 == 0).astype(float), "tke": zero, "omega": np.ones(ny),
         "massf[0]-air": np.ones(ny), "u": u, "T": T}
    return ff.FlowSolution(ff.flow_variable_list(gas), v, 2, 1, ny, 1)


def test_static_profile_matching(tmp_path):
    """flowstate.d:589-736: each ghost cell takes the nearest profile point
    (first on ties).  Profile points at y = (m + 0.5)/8; a block of 16 cells
    in y has its west ghosts at y = (j + 0.5)/16, whose nearest points are
    m = j // 2 (the lower one on the exact tie is impossible here: no ghost
    is equidistant)."""
    gas = e4.IdealGas()
    fl = _profile_flow(gas, 8)
    ff.write_profile(tmp_path / "prof.dat", fl, range(8))
    prof = ff.FlowProfile(tmp_path / "prof.dat", gas)
    assert prof.data.shape == (8, len(ff.flow_variable_list(gas)))
    g = ff.ghost_centres(2, 4, 16, 1, 0.25, 1.0 / 16, 1.0, 3)       # west face
    assert g.shape == (2, 16, 3)
    assert np.allclose(g[0, :, 0], -0.125) and np.allclose(g[1, :, 0], -0.375)
    for j in range(16):
        assert prof.nearest(g[0, j]) == j // 2
        st = prof.state(g[0, j])
        m = j // 2
        assert st[0] == fl.values["rho"][m] and st[5] == fl.values["vel.x"][m] and st[8] == fl.values["S"][m]
    # AyA-to-AyA ignores x and z
    prof_y = ff.FlowProfile(tmp_path / "prof.dat", gas, "AyA-to-AyA")
    assert prof_y.nearest([100.0, 0.3, -7.0]) == 2
    with pytest.raises(ValueError):
        ff.FlowProfile(tmp_path / "prof.dat", gas, "xyz-to-abc")


def test_static_profile_uniform_is_supersonic_inflow():
    """A profile of one state everywhere is InFlowBC_Supersonic with that
    state (the oracle, bit for bit)."""
    import pyoracle as O
    gas = e4.IdealGas()
    og = O.Gas.ideal()
    rng = np.random.default_rng(3)
    n = 8
    shape = (n, n, n)
    U0 = gas.encode(*[x.ravel() for x in (np.full(shape, 1.2) + 0.1 * rng.random(shape),
                                          2.2e5 + 1e4 * rng.random(shape), 100 * rng.random(shape),
                                          np.zeros(shape), np.zeros(shape))])
    fs = O.flowstate_pT(og, 1.0e5, 300.0, (700.0, 5.0, -3.0))
    states = np.zeros((2, n * n, 10))
    states[..., :8] = fs[:8]
    states[..., 9] = 1.0
    blocks = []
    for kind, arg in ((O.BC_INFLOW, fs), (O.BC_INFLOW_PROFILE, states)):
        b = O.Block(og, 3, n, n, n, 0.1, 0.1, 0.1)
        b.set_config(flux_calc=O.AUSMDV)
        b.set_bc(O.WEST, kind, arg)
        b.set_bc(O.EAST, O.BC_EXTRAPOLATE)
        b.set_U(U0)
        for _ in range(3):
            assert b.step(1e-5)[0] == 0
        blocks.append(b.get_U())
    assert np.array_equal(blocks[0], blocks[1])


def test_job_layout_and_times(tmp_path):
    """fileutil.d:15-33 names and the times file of simcore.d:678-683."""
    assert ff.make_file_name("flow", "cone", 3, 12) == "flow/t0012/cone.flow.b0003.t0012.gz"
    assert ff.make_file_name("grid", "cone", 0, 0, "rawbinary") == "grid/t0000/cone.grid.b0000.t0000.bin"
    ff.append_times(tmp_path, "cone", 0, 0.0, 1.0e-6)
    ff.append_times(tmp_path, "cone", 1, 2.5e-3, 3.0e-6)
    lines = open(tmp_path / "config" / "cone.times").read().splitlines()
    assert lines[0] == "# tindx sim_time dt_global"
    assert lines[2] == "0001 2.500000000000000052e-03 3.000000000000000076e-06"
    assert ff.read_times(tmp_path, "cone") == [(0, 0.0, 1.0e-6), (1, 2.5e-3, 3.0e-6)]


def test_reference_named_boundary_conditions():
    from eilmer4_amd import bc
    from eilmer4_amd.fvcore import BCType
    assert bc.WallBC_WithSlip().kind == BCType.slip_wall and bc.WallBC_WithSlip1().kind == BCType.slip_wall_1
    assert bc.OutFlowBC_Simple().kind == BCType.outflow_simple_flux
    assert bc.OutFlowBC_SimpleExtrapolate(1).kind == BCType.extrapolate_1
    fs = e4.IdealGas().flowstate(1e5, 300.0, (500.0, 0.0, 0.0))
    b = bc.InFlowBC_Supersonic(fs)
    assert b.kind == BCType.inflow and np.array_equal(b.fs, fs)
    with pytest.raises(ValueError):
        bc.OutFlowBC_SimpleExtrapolate(2)
