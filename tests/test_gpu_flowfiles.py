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


def test_job_from_reference_named_conditions(gpu, tmp_path):
    """Boundary conditions given by their reference names (a static-profile
    inflow from a profile file, a simple outflow, walls with and without ghost
    data) and the solution written as a job's flow files and times file: the
    device run equals the oracle run with the same boundary data (to the last
    bits of the wall flux's pow())."""
    from eilmer4_amd import bc
    gas = e4.IdealGas()
    ni, nj = 24, 12
    dx, dy = 1.0 / ni, 0.5 / nj
    y = (np.arange(6) + 0.5) / 12.0
    p = 1.0e5 * (1.0 + 0.2 * y)
    T = 300.0 + 50.0 * y
    rho, u, a = gas.rho_u_a_from_pT(p, T)
    zero = np.zeros(6)
    prof = ff.FlowSolution(ff.flow_variable_list(gas),
                           {"pos.x": np.full(6, -0.01), "pos.y": y, "pos.z": zero, "volume": zero + 1e-3, "rho": rho,
                            "vel.x": np.full(6, 700.0), "vel.y": zero, "vel.z": zero, "p": p, "a": a, "mu": zero,
                            "k": zero, "mu_t": zero, "k_t": zero, "S": zero, "tke": zero, "omega": zero + 1.0,
                            "massf[0]-air": zero + 1.0, "u": u, "T": T}, 2, 6, 1, 1)
    ff.write_profile(tmp_path / "profile.dat", prof, range(6))
    conds = {e4.Face.west: bc.InFlowBC_StaticProfile(tmp_path / "profile.dat", gas),
             e4.Face.east: bc.OutFlowBC_Simple(), e4.Face.north: bc.WallBC_WithSlip1(),
             e4.Face.south: bc.WallBC_WithSlip()}
    rng = np.random.default_rng(71)
    shape = (1, nj, ni)
    rho0, u0, _ = gas.rho_u_a_from_pT(1.0e5 * (1 + 0.1 * rng.random(shape)).ravel(), np.full(ni * nj, 300.0))
    U0 = gas.encode(rho0, u0, np.full(ni * nj, 650.0), np.zeros(ni * nj), np.zeros(ni * nj))
    db = e4.SFluidBlock(gas, 2, ni, nj, 1, dx, dy, 1.0)
    ob = O.Block(O.Gas.ideal(), 2, ni, nj, 1, dx, dy)
    ob.set_config(flux_calc=O.AUSMDV)
    for face, cond in conds.items():
        db.set_bc(face, cond)
        data = cond.fs(db, face) if callable(cond.fs) else cond.fs
        ob.set_bc(int(face), int(cond.kind), data)
    db.upload(U0)
    ob.set_U(U0)
    sim = device_sim([db], flux=e4.FluxCalculator.ausmdv, precision="fp64_exact")
    osim = O.Simulation(ob)
    print("case job from reference-named conditions", flush=True)
    t = 0.0
    for s in range(5):
        osim.run(max_step=s + 1)
        sim.state.dt_global = osim.dt_global
        sim.gasdynamic_explicit_increment_with_fixed_grid()
        t += osim.dt_global
    # the north wall's flux uses pow(), whose last bit may differ between the
    # device and the host library
    from parity_helpers import rel_err
    assert rel_err(db.download(), ob.get_U()).max() <= 1.0e-13
    ff.write_solution_files(tmp_path / "job", "demo", [db], 1, t, osim.dt_global, with_grid=True)
    back = ff.read_flow(tmp_path / "job" / ff.make_file_name("flow", "demo", 0, 1))
    _, fs = db.download(with_fs=True)
    assert np.array_equal(back.values["rho"], fs[0]) and back.sim_time == t
    assert ff.read_times(tmp_path / "job", "demo") == [(1, t, osim.dt_global)]
    db.destroy()
