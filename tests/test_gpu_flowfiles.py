"""A run that starts from Eilmer4 grid and flow files and ends in one: the
device block built from the files (Cartesian spacing from the grid, conserved
state by encode_conserved of the flow file's states) matches the oracle run
from the same state bit for bit in exact FP64, and the flow file written from
the device holds exactly the downloaded states."""
import numpy as np
import pytest

import pyoracle as O
import eilmer4_amd as e4
from eilmer4_amd import flowfiles as ff
from eilmer4_amd import workloads
from parity_helpers import oracle_block, device_block, device_sim

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", ["gziptext", "rawbinary"])
def test_flow_file_pipeline(gpu, tmp_path, fmt):
    gas = e4.IdealGas()
    n = 16
    geom, fields, bcs, _ = workloads.shocked_box_3d(gas, n=n)
    U0 = workloads.initial_U(gas, fields)
    print("case flow files shocked box n=%d format=%s" % (n, fmt), flush=True)
    d0 = device_block(gas, geom, bcs, U0)
    ff.write_flow(tmp_path / "init.flow", ff.flow_from_block(d0, label="box"), fmt)
    ff.write_grid(tmp_path / "box.grid", ff.cartesian_grid(3, n, n, n, geom["dx"], geom["dy"], geom["dz"]), fmt)
    d0.destroy()
    g = ff.read_grid(tmp_path / "box.grid")
    dx, dy, dz, _ = g.cartesian_spacing()
    fl = ff.read_flow(tmp_path / "init.flow")
    U, _ = ff.conserved_from_flow(fl, gas)
    geom2 = dict(dims=3, ni=fl.nic, nj=fl.njc, nk=fl.nkc, dx=dx, dy=dy, dz=dz)
    ob = oracle_block(geom2, bcs, U, flux=O.AUSMDV)
    db = device_block(gas, geom2, bcs, U)
    sim = device_sim([db], flux=e4.FluxCalculator.ausmdv, precision="fp64_exact")
    osim = O.Simulation(ob)
    for s in range(3):
        osim.run(max_step=s + 1)
        sim.state.dt_global = osim.dt_global
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    Ud, fs = db.download(with_fs=True)
    assert np.array_equal(Ud, ob.get_U())
    ff.write_flow(tmp_path / "t0003.flow", ff.flow_from_block(db, sim_time=3.0e-6), fmt)
    out = ff.read_flow(tmp_path / "t0003.flow")
    for q, name in enumerate(["rho", "p", "T", "u", "a", "vel.x", "vel.y", "vel.z"]):
        assert np.array_equal(out.values[name], fs[q]), name
    db.destroy()
