"""The learned temperature guess: training kernels against the FP64 CPU
reference, the BF16 MFMA inference against a host emulation, and the decode
with the guess against the reference solve on identical U."""
import numpy as np
import pytest

import mlp_ref
import pyoracle as O
import eilmer4_amd as e4
from eilmer4_amd import learned_eos as L, workloads
from parity_helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gas(gpu):
    return e4.ThermallyPerfectGas.from_json()


@pytest.fixture(scope="module")
def small_data(gas):
    X, T = L.newton_labels(gas, n=24, seeds=L.DATA_SEEDS[:3])
    return X, T


def test_training_gradients_match_cpu_reference(gas, small_data):
    X, T = small_data
    model = L.GuessModel.init(X, T, seed=3)
    rng = np.random.default_rng(0)
    idx = rng.choice(X.shape[0], 512, replace=False)
    Z, t = model.features(X[idx]), model.targets(T[idx])
    loss_d, g_d = L.mlp_grad(model.params, Z, t)
    loss_r, g_r = mlp_ref.loss_and_grad(model.params, Z, t)
    rel = np.linalg.norm(g_d - g_r) / np.linalg.norm(g_r)
    print("loss device %.6e reference %.6e; gradient rel err %.3e" % (loss_d, loss_r, rel))
    assert abs(loss_d - loss_r) / loss_r < 1e-4
    assert rel < 1e-4
    # every parameter block, so a wrong index in one layer's backward cannot hide
    o = 0
    for n in (512, 64, 4096, 64, 4096, 64, 64, 1):
        b = np.linalg.norm(g_d[o:o + n] - g_r[o:o + n]) / max(np.linalg.norm(g_r[o:o + n]), 1e-30)
        assert b < 1e-3, (o, n, b)
        o += n


def test_mfma_inference_matches_bf16_emulation(gas, small_data):
    """Device guesses (BF16 MFMA, 16x16x32 operand layout) against a host
    emulation of the same BF16 arithmetic: a wrong lane layout gives O(1)
    errors, rounding differences give ~1e-3 of the output scale."""
    X, T = small_data
    model = L.GuessModel.init(X, T, seed=5)
    model.install(gas)
    Xs = X[:1000]
    Td = L.device_guess(gas, Xs)
    y = mlp_ref.forward_bf16(model.params, model.features(Xs))
    Th = model.T_mean + model.T_std * y
    err = np.max(np.abs(Td - Th)) / model.T_std
    print("max |T_device - T_emulated| / T_std = %.3e" % err)
    assert err < 5e-3


@pytest.fixture(scope="module")
def trained(gas, small_data):
    X, T = small_data
    model = L.GuessModel.init(X, T, seed=11)
    losses = L.train(model, X, T, epochs=25, batch=2048, lr=3e-3)
    print("training loss per epoch", losses)
    model.install(gas)
    return model, losses


def test_training_reduces_loss(trained, gas, small_data):
    model, losses = trained
    X, T = small_data
    assert losses[-1] < 0.05 * losses[0]
    Tg = L.device_guess(gas, X[:20000])
    rms = float(np.sqrt(np.mean((Tg - T[:20000]) ** 2)))
    print("guess rms error %.2f K over T in [%.0f, %.0f]" % (rms, T.min(), T.max()))
    assert rms < 100.0


def _upload_decode(gas, U, Tg, n, learned):
    blk = e4.SFluidBlock(gas, 3, n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
    e4.Simulation([blk], e4.GlobalConfig(precision_mode=e4.PrecisionMode.fp64_exact, learned_eos_guess=learned))
    L.decode_stats(reset=True)
    blk.upload(U, Tg)
    ev, fb = L.decode_stats(reset=True)
    _, fs = blk.download(with_fs=True)
    blk.destroy()
    return fs, ev, fb


def test_learned_decode_matches_reference_solve(trained, gas):
    """Per decode call on identical U (the thermally perfect workload field,
    seeded apart from the training data): |T_learned - T_reference| / T <= 1e-10,
    and fewer energy evaluations than the reference's bracketed solve from a
    stored temperature 5 % off."""
    n = 24
    geom, fields, bcs, _ = workloads.thermally_perfect_box_3d(gas, n=n)
    U, Tgen = workloads.initial_U_tp(gas, fields)
    Tstored = Tgen * (1.0 + 0.05 * np.random.default_rng(1).uniform(-1, 1, Tgen.shape))
    fs_ref, ev_ref, _ = _upload_decode(gas, U, Tstored, n, False)
    fs_ml, ev_ml, fb_ml = _upload_decode(gas, U, Tstored, n, True)
    errT = float(np.max(np.abs(fs_ml[2] - fs_ref[2]) / fs_ref[2]))
    print("decode: T rel diff %.3e; energy evaluations reference %d learned %d (saved %.1f %%); fallbacks %d of %d"
          % (errT, ev_ref, ev_ml, 100.0 * (1 - ev_ml / ev_ref), fb_ml, n ** 3))
    assert errT <= 1e-10
    assert ev_ml < ev_ref
    assert fb_ml < 0.1 * n ** 3


def test_learned_steps_against_oracle(trained, gas):
    """Five RK2 steps with the learned guess in the decode against the oracle
    (reference solve): the guess changes only the Newton iterate path, so
    density stays within 1e-5 and temperature within 1e-8 (the u side effect of
    the last iterate differs at ~1e-9)."""
    n = 12
    geom, fields, bcs, _ = workloads.thermally_perfect_box_3d(gas, n=n)
    U0, Tg = workloads.initial_U_tp(gas, fields)
    og = O.Gas.from_json("therm-perf-5-species-air.json")
    ob = O.Block(og, 3, n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
    ob.set_config()
    for face, (kind, fs) in bcs.items():
        ob.set_bc(int(face), int(kind), fs)
    ob.set_U(U0, Tg)
    db = e4.SFluidBlock(gas, 3, n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
    for face, (kind, fs) in bcs.items():
        db.set_bc(face, kind, fs)
    sim = e4.Simulation([db], e4.GlobalConfig(learned_eos_guess=True))
    db.upload(U0, Tg)
    dt, _, _ = ob.determine_dt(1e-3, 0.5, False)
    L.decode_stats(reset=True)
    for _ in range(5):
        assert ob.step(dt)[0] == 0
        sim.state.dt_global = dt
        sim.gasdynamic_explicit_increment_with_fixed_grid()
    ev, fb = L.decode_stats(reset=True)
    U, fs = db.download(with_fs=True)
    err = rel_err(U, ob.get_U())
    errT = float(np.max(np.abs(fs[2] - ob.get_fs()[2]) / ob.get_fs()[2]))
    print("learned, 5 steps: U rel err %s, T rel err %.3e, decode energy evals %d, fallbacks %d" % (err, errT, ev, fb))
    assert err[0] <= 1e-5
    assert errT <= 1e-8
    db.destroy()
