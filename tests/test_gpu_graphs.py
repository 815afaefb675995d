"""Graph replay of repeated steps (use_graphs): every step of a run that
captures and replays HIP graphs equals, bit for bit, the same step run
launch by launch.  Each case runs the eager sequence first (storing U after
every step), then the graph sequence on fresh blocks, downloading and
comparing after every step, so each replay is followed by eager work and a
synchronisation.  Cases: four batched 2D blocks with stored ghost states and
source terms under the CFL policy (dt changes every 10 steps: capture, replay,
re-capture); a mixture block beside an ideal block with the shock detector
(buffer swaps inside the step); a change of source terms between replays."""
import numpy as np
import pytest

import eilmer4_amd as e4
from eilmer4_amd.fvcore import BCType
import mms_case as mms

pytestmark = pytest.mark.gpu


def _mms_blocks():
    gas = e4.IdealGas(mMass=mms.MMASS, gamma=mms.GAMMA)
    layout, h = mms.block_layout(16, 2, 2)
    blks = {}
    for (ib, jb, ni, nj, origin) in layout:
        b = e4.SFluidBlock(gas, 2, ni, nj, 1, h, h)
        b.set_from_pT(1.0e5, 1.0e5 / mms.RGAS, 800.0, 800.0, 0.0)
        b.set_source_terms(mms.cell_sources(ni, nj, h, origin))
        if ib == 1:
            b.set_bc(e4.Face.east, BCType.extrapolate_1)
        if jb == 1:
            b.set_bc(e4.Face.north, BCType.extrapolate_1)
        if ib == 0:
            b.set_bc(e4.Face.west, BCType.inflow_profile, mms.ghost_states_for_face(ni, nj, h, 3, origin))
        if jb == 0:
            b.set_bc(e4.Face.south, BCType.inflow_profile, mms.ghost_states_for_face(ni, nj, h, 2, origin))
        blks[(ib, jb)] = b
    for (ib, jb), b in blks.items():
        if ib == 0:
            e4.connect_blocks(b, e4.Face.east, blks[(1, jb)], e4.Face.west)
        if jb == 0:
            e4.connect_blocks(b, e4.Face.north, blks[(ib, 1)], e4.Face.south)
    return [blks[(ib, jb)] for (ib, jb, *_r) in layout], layout, h


def _run(blocks, cfg, nsteps, fixed_dt=None, hook=None):
    sim = e4.Simulation(blocks, cfg)
    out = []
    for s in range(nsteps):
        if hook:
            hook(s, blocks)
        if fixed_dt is None:
            sim.determine_time_step_size()
        else:
            sim.state.dt_global = fixed_dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
        sim.state.step += 1
        out.append((sim.state.dt_global, [b.download() for b in blocks]))
    return out


def _compare(eager, graph, label):
    for s, ((dte, Ue), (dtg, Ug)) in enumerate(zip(eager, graph)):
        assert dte == dtg, (label, s)
        for k, (ue, ug) in enumerate(zip(Ue, Ug)):
            assert np.array_equal(ue, ug), (label, s, k, np.max(np.abs(ue - ug)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_graph_replay_batched_blocks(gpu, precision):
    print("case graphs: 2x2 batched 2D blocks, sources, profile and x_order-1 ghosts, %s, 25 steps" % precision,
          flush=True)
    runs = []
    for graphs in (False, True):
        blocks, _, _ = _mms_blocks()
        cfg = e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv, apply_limiter=False, extrema_clipping=False,
                              stringent_cfl=True, dt_init=1.0e-6, precision_mode=e4.PrecisionMode[precision],
                              use_graphs=graphs)
        runs.append(_run(blocks, cfg, 25))
        for b in blocks:
            b.destroy()
    assert len({dt for dt, _ in runs[0]}) >= 3       # dt changed, so graphs were re-captured
    _compare(runs[0], runs[1], "batched")


def test_graph_replay_mixture_and_detector(gpu):
    print("case graphs: mixture block + ideal block, adaptive_hanel_ausmdv, 12 steps", flush=True)
    from eilmer4_amd import workloads
    gas_tp = e4.ThermallyPerfectGas.from_json()
    gas_id = e4.IdealGas()
    n = 8
    rng = np.random.default_rng(21)
    shape = (n, n, n)
    Y = np.array([0.7, 0.2, 0.05, 0.03, 0.02])
    Yf = np.broadcast_to(Y[:, None], (5, n ** 3)).copy()
    T = 1500.0 * (1.0 + 0.2 * rng.random(shape))
    p = 1.0e5 * (1.0 + 0.3 * rng.random(shape))
    rho = (p.ravel() / (gas_tp.Rmix(Yf) * T.ravel())).reshape(shape)
    fields = dict(p=p, T=T, rho=rho, vx=400.0 * (rng.random(shape) - 0.5), vy=np.zeros(shape), vz=np.zeros(shape),
                  Y=Yf)
    U_tp, Tg = workloads.initial_U_tp(gas_tp, fields)
    pi = 1.0e5 * (1.0 + 0.5 * rng.random(shape))
    Ti = 300.0 * (1.0 + 0.5 * rng.random(shape))
    r, u, _ = gas_id.rho_u_a_from_pT(pi.ravel(), Ti.ravel())
    U_id = gas_id.encode(r, u, 500.0 * (rng.random(n ** 3) - 0.5), np.zeros(n ** 3), np.zeros(n ** 3))
    runs = []
    for graphs in (False, True):
        a = e4.SFluidBlock(gas_tp, 3, n, n, n, 0.1, 0.1, 0.1)
        b = e4.SFluidBlock(gas_id, 3, n, n, n, 0.01, 0.01, 0.01)
        for face in e4.Face:
            a.set_bc(face, BCType.slip_wall)
            b.set_bc(face, BCType.extrapolate_1 if face in (e4.Face.east, e4.Face.west) else BCType.slip_wall)
        a.upload(U_tp, Tg)
        b.upload(U_id)
        cfg = e4.GlobalConfig(flux_calculator=e4.FluxCalculator.adaptive_hanel_ausmdv,
                              precision_mode=e4.PrecisionMode.fp64, use_graphs=graphs)
        runs.append(_run([a, b], cfg, 12, fixed_dt=2.0e-7))
        a.destroy()
        b.destroy()
    _compare(runs[0], runs[1], "mixture")


def test_graph_dropped_on_change(gpu):
    """New source terms between steps: the graph run follows the eager run."""
    print("case graphs: source terms changed at step 6", flush=True)

    def hook(s, blocks):
        if s == 6:
            for k, b in enumerate(blocks):
                b.set_source_terms(np.full((b.nq, b.ncells), 10.0 * (k + 1)))

    runs = []
    for graphs in (False, True):
        blocks, _, _ = _mms_blocks()
        cfg = e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv, apply_limiter=False, extrema_clipping=False,
                              precision_mode=e4.PrecisionMode.fp64_exact, use_graphs=graphs)
        runs.append(_run(blocks, cfg, 12, fixed_dt=1.0e-6, hook=hook))
        for b in blocks:
            b.destroy()
    _compare(runs[0], runs[1], "change")


@pytest.mark.parametrize("graphs", [False, True])
def test_batched_2d_and_3d_blocks_in_one_call(gpu, graphs):
    """Two 2D and two 3D ideal-gas blocks in one call: each dimensionality is
    one batched launch per stage with its own argument table; the result
    equals one launch per block, bit for bit, eager and under graph replay."""
    print("case batch: 2D + 3D blocks in one call, graphs=%s" % graphs, flush=True)
    from eilmer4_amd import workloads
    gas = e4.IdealGas()
    g3, f3, bc3, _ = workloads.shocked_box_3d(gas, n=12, seed=4)
    g2, f2, bc2, _ = workloads.planar_bump_2d(gas, ni=40, nj=24)
    U3 = workloads.initial_U(gas, f3)
    U2 = workloads.initial_U(gas, f2)
    runs = []
    for batch in (False, True):
        blocks = []
        for geom, bcs, U in ((g2, bc2, U2), (g3, bc3, U3), (g2, bc2, U2 * 1.01), (g3, bc3, U3 * 0.99)):
            b = e4.SFluidBlock(gas, geom["dims"], geom["ni"], geom["nj"], geom["nk"], geom["dx"], geom["dy"],
                               geom["dz"])
            for face, (kind, fs) in bcs.items():
                b.set_bc(face, kind, fs)
            b.upload(U)
            blocks.append(b)
        cfg = e4.GlobalConfig(flux_calculator=e4.FluxCalculator.ausmdv, precision_mode=e4.PrecisionMode.fp64,
                              batch_blocks=batch, use_graphs=graphs and batch)
        runs.append(_run(blocks, cfg, 8, fixed_dt=2.0e-6))
        for b in blocks:
            b.destroy()
    _compare(runs[0], runs[1], "2d+3d")
