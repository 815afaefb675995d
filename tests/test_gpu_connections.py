"""Full-face block connections on the device against the oracle: every
face/orientation pairing (2D and 3D), decomposition invariance, a connected
pair of thermally perfect blocks, and the argument checks of
e4hip_block_connect."""
import numpy as np
import pytest

import pyoracle as O
import eilmer4_amd as e4
from eilmer4_amd import workloads
from parity_helpers import rel_err

pytestmark = pytest.mark.gpu


def _state(gas, shape, seed):
    rng = np.random.default_rng(seed)
    p = 1.0e5 * (1.0 + 0.4 * rng.random(shape))
    T = 300.0 * (1.0 + 0.4 * rng.random(shape))
    v = [150.0 * (rng.random(shape) - 0.5) for _ in range(3)]
    rho, u, _ = gas.rho_u_a_from_pT(p.ravel(), T.ravel())
    return gas.encode(rho, u, *[x.ravel() for x in v])


def _pair(dims, f, g, o, n, precision, steps=2, seed=11):
    gas = e4.IdealGas()
    shape = (n, n, n) if dims == 3 else (1, n, n)
    nk = n if dims == 3 else 1
    UA, UB = _state(gas, shape, seed), _state(gas, shape, seed + 1)
    oa, ob = (O.Block(O.Gas.ideal(), dims, n, n, nk, 0.1, 0.1, 0.1) for _ in range(2))
    for b, U in ((oa, UA), (ob, UB)):
        b.set_config()
        b.set_U(U)
    O.connect_blocks(oa, f, ob, g, o)
    da, db = (e4.SFluidBlock(gas, dims, n, n, nk, 0.1, 0.1, 0.1) for _ in range(2))
    e4.connect_blocks(da, f, db, g, o)
    sim = e4.Simulation([da, db], e4.GlobalConfig(precision_mode=e4.PrecisionMode[precision]))
    da.upload(UA)
    db.upload(UB)
    dt = 1.0e-5
    for _ in range(steps):
        assert O.multi_step([oa, ob], dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    out = (oa.get_U(), ob.get_U(), da.download(), db.download())
    da.destroy()
    db.destroy()
    return out


def test_every_3d_pairing_matches_oracle(gpu):
    """All 6 x 6 x 4 face/orientation pairings of two 5^3 blocks, two RK2
    steps, exact FP64: bit-identical to the oracle."""
    bad = []
    for f in range(6):
        for g in range(6):
            for o in range(4):
                print("case 3d connect f=%d g=%d o=%d" % (f, g, o), flush=True)
                Uoa, Uob, Uda, Udb = _pair(3, f, g, o, 5, "fp64_exact")
                if not (np.array_equal(Uoa, Uda) and np.array_equal(Uob, Udb)):
                    bad.append((f, g, o, rel_err(Uda, Uoa).max(), rel_err(Udb, Uob).max()))
    assert not bad, bad


def test_every_2d_pairing_matches_oracle(gpu):
    bad = []
    for f in range(4):
        for g in range(4):
            print("case 2d connect f=%d g=%d" % (f, g), flush=True)
            Uoa, Uob, Uda, Udb = _pair(2, f, g, 0, 7, "fp64_exact", steps=3)
            if not (np.array_equal(Uoa, Uda) and np.array_equal(Uob, Udb)):
                bad.append((f, g, rel_err(Uda, Uoa).max(), rel_err(Udb, Uob).max()))
    assert not bad, bad


@pytest.mark.parametrize("precision", ["fp64_exact", "fp64"])
def test_decomposition_invariance(gpu, precision):
    """The 64^3 shocked box against the same cells in four blocks (split in i
    and k, joined east-west and top-bottom, periodic in j through
    self-connections): identical after 5 steps (the connection-test pattern)."""
    print("case decomposition precision=%s" % precision, flush=True)
    gas = e4.IdealGas()
    n = 64
    geom, fields, bcs, _ = workloads.shocked_box_3d(gas, n=n)
    U0 = workloads.initial_U(gas, fields).reshape(5, n, n, n)
    whole = e4.SFluidBlock(gas, 3, n, n, n, geom["dx"], geom["dy"], geom["dz"])
    for face, (kind, fs) in bcs.items():
        whole.set_bc(face, kind, fs)
    ci, ck = 24, 40
    parts = {}
    for a, isl in enumerate((slice(0, ci), slice(ci, n))):
        for c, ksl in enumerate((slice(0, ck), slice(ck, n))):
            ni, nk = isl.stop - isl.start, ksl.stop - ksl.start
            b = e4.SFluidBlock(gas, 3, ni, n, nk, geom["dx"], geom["dy"], geom["dz"])
            parts[(a, c)] = (b, isl, ksl)
    F = e4.Face
    for (a, c), (b, isl, ksl) in parts.items():
        if a == 0:
            b.set_bc(F.west, *bcs[F.west])
        else:
            b.set_bc(F.east, *bcs[F.east])
        e4.connect_blocks(b, F.north, b, F.south, 0)       # periodic in j
    for c in (0, 1):
        e4.connect_blocks(parts[(0, c)][0], F.east, parts[(1, c)][0], F.west, 0)
    for a in (0, 1):                                        # periodic in k across the two parts
        e4.connect_blocks(parts[(a, 0)][0], F.top, parts[(a, 1)][0], F.bottom, 0)
        e4.connect_blocks(parts[(a, 1)][0], F.top, parts[(a, 0)][0], F.bottom, 0)
    blocks = [whole] + [p[0] for p in parts.values()]
    # (AUSMDV: with the shock indicator the ghost cells carry the previous
    # detection's flags, so a decomposition is not invariant)
    cfg = dict(flux_calculator=e4.FluxCalculator.ausmdv, precision_mode=e4.PrecisionMode[precision])
    sim_w = e4.Simulation([whole], e4.GlobalConfig(**cfg))
    whole.upload(U0)
    for b, isl, ksl in parts.values():
        b.upload(np.ascontiguousarray(U0[:, ksl, :, isl]))
    sim_p = e4.Simulation([p[0] for p in parts.values()], e4.GlobalConfig(**cfg))
    dt = 2.0e-7
    for _ in range(5):
        for sim in (sim_w, sim_p):
            sim.state.dt_global = dt
            sim.gasdynamic_explicit_increment_with_fixed_grid()
    Uw = whole.download().reshape(5, n, n, n)
    for b, isl, ksl in parts.values():
        Ub = b.download().reshape(5, ksl.stop - ksl.start, n, isl.stop - isl.start)
        assert np.array_equal(Uw[:, ksl, :, isl], Ub), (isl, ksl, rel_err(Ub, Uw[:, ksl, :, isl]))
    for b in blocks:
        b.destroy()


def test_connected_thermally_perfect_blocks_match_oracle(gpu):
    """Two 5-species blocks joined east-to-north with orientation 1, three
    steps, against the oracle (bar 1e-10 as for the single-block TP tests)."""
    print("case tp connected", flush=True)
    gas = e4.ThermallyPerfectGas.from_json()
    n = 8
    geom, fields, _, _ = workloads.thermally_perfect_box_3d(gas, n=n)
    U0, Tg = workloads.initial_U_tp(gas, fields)
    og = O.Gas.from_json("therm-perf-5-species-air.json")
    oa, ob = (O.Block(og, 3, n, n, n, geom["dx"], geom["dy"], geom["dz"]) for _ in range(2))
    Ushift = U0.reshape(U0.shape[0], n, n, n)[:, ::-1].reshape(U0.shape[0], -1).copy()
    Tshift = Tg.reshape(n, n, n)[::-1].ravel().copy()
    for b, U, T in ((oa, U0, Tg), (ob, Ushift, Tshift)):
        b.set_config()
        b.set_U(U, T)
    O.connect_blocks(oa, O.EAST, ob, O.NORTH, 1)
    da, db = (e4.SFluidBlock(gas, 3, n, n, n, geom["dx"], geom["dy"], geom["dz"]) for _ in range(2))
    e4.connect_blocks(da, e4.Face.east, db, e4.Face.north, 1)
    sim = e4.Simulation([da, db], e4.GlobalConfig(precision_mode=e4.PrecisionMode.fp64_exact))
    da.upload(U0, Tg)
    db.upload(Ushift, Tshift)
    dt = 2.0e-7
    for _ in range(3):
        assert O.multi_step([oa, ob], dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    ea, eb = rel_err(da.download(), oa.get_U()), rel_err(db.download(), ob.get_U())
    print("  rel err", ea.max(), eb.max(), flush=True)
    assert ea.max() <= 1e-10 and eb.max() <= 1e-10
    da.destroy()
    db.destroy()


def test_connect_rejects_mismatched_faces(gpu):
    gas = e4.IdealGas()
    a = e4.SFluidBlock(gas, 3, 8, 6, 4, 0.1, 0.1, 0.1)
    b = e4.SFluidBlock(gas, 3, 8, 6, 5, 0.1, 0.1, 0.1)
    with pytest.raises(e4.E4hipError):
        a.connect(e4.Face.east, b, e4.Face.west, 0)     # 6x4 face against 6x5
    with pytest.raises(e4.E4hipError):
        a.connect(e4.Face.north, a, e4.Face.south, 1)   # 8x4 face rotated onto 8x4 needs 4x8
    a.connect(e4.Face.north, a, e4.Face.south, 0)
    c = e4.SFluidBlock(e4.ThermallyPerfectGas.from_json(), 3, 8, 6, 4, 0.1, 0.1, 0.1)
    with pytest.raises(e4.E4hipError):
        a.connect(e4.Face.east, c, e4.Face.west, 0)     # different gas model
    for blk in (a, b, c):
        blk.destroy()


def test_batched_launch_matches_per_block_launches(gpu):
    """A small C5-style set (random face-matched tree of ideal and mixture
    blocks, every orientation) stepped with the ideal blocks batched into one
    launch per stage and with one launch per block: bit-identical in exact
    FP64, and the batched run matches the oracle's multi-block step."""
    ig, tg = e4.IdealGas(), e4.ThermallyPerfectGas.from_json()
    descs, conns = workloads.multiblock_c5(ig, tg, n_ideal=6, n_tp=3, target_cells=2.0e5)
    print("case batched %d blocks %d cells" % (len(descs), sum(d["ni"] * d["nj"] * d["nk"] for d in descs)),
          flush=True)
    res = {}
    for batch in (True, False):
        blks = []
        for d in descs:
            b = e4.SFluidBlock(ig if d["cls"] == "ideal" else tg, 3, d["ni"], d["nj"], d["nk"], d["dx"], d["dy"],
                               d["dz"])
            for face, (kind, fs) in d["bcs"].items():
                b.set_bc(face, kind, fs)
            blks.append(b)
        for (a, fa, b, fb, o) in conns:
            e4.connect_blocks(blks[a], fa, blks[b], fb, o)
        sim = e4.Simulation(blks, e4.GlobalConfig(precision_mode=e4.PrecisionMode.fp64_exact, batch_blocks=batch))
        for b, d in zip(blks, descs):
            b.upload(d["U0"], d["Tguess"])
        for _ in range(2):
            sim.state.dt_global = 1.0e-7
            sim.gasdynamic_explicit_increment_with_fixed_grid()
        res[batch] = [b.download() for b in blks]
        for b in blks:
            b.destroy()
    for Ua, Ub in zip(res[True], res[False]):
        assert np.array_equal(Ua, Ub)
    # the ideal blocks against the oracle (the mixture blocks are their own class)
    ideal = [i for i, d in enumerate(descs) if d["cls"] == "ideal"]
    ob = []
    for i in ideal:
        d = descs[i]
        b = O.Block(O.Gas.ideal(), 3, d["ni"], d["nj"], d["nk"], d["dx"], d["dy"], d["dz"])
        b.set_config()
        for face, (kind, fs) in d["bcs"].items():
            b.set_bc(int(face), int(kind), fs)
        b.set_U(d["U0"])
        ob.append(b)
    for (a, fa, b, fb, o) in conns:
        if a in ideal and b in ideal:
            O.connect_blocks(ob[ideal.index(a)], fa, ob[ideal.index(b)], fb, o)
    for _ in range(2):
        assert O.multi_step(ob, 1.0e-7)[0] == 0
    for k, i in enumerate(ideal):
        assert np.array_equal(res[True][i], ob[k].get_U()), (i, rel_err(res[True][i], ob[k].get_U()))
