"""Learned initial guess of the thermally perfect temperature solve.

Data: states (rho, u, Y0..Y4) of separately seeded 5-species fields, labelled
with the temperature the device's own decode Newton solve
(therm_perf_gas_mix_eos.d:70-153 on the device) returns for them.  Training:
the library's own FP32 forward/backward/Adam kernels (e4hip_mlp_train).
Inference: BF16 MFMA inside the decode (e4hip_gas_tp_set_guess_model).
"""
import ctypes as C
import os

import numpy as np

from . import lib as _lib
from . import workloads
from .fluidblock import SFluidBlock
from .simcore import GlobalConfig, Simulation

NPARAM = 8961
NH = 64
DATA_SEEDS = tuple(range(20261210, 20261220))
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tp5_guess_model.npz")

_fp = C.POINTER(C.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def decode_inputs(gas, U):
    """(rho, u, Y0..Y4) exactly as decode_conserved forms them from U
    (fvcell.d:590-675): u = E/rho - ke, Y = rhoY/rho then scale_mass_fractions."""
    rho = U[0]
    dinv = 1.0 / rho
    vx, vy, vz = U[1] * dinv, U[2] * dinv, U[3] * dinv
    u = U[4] * dinv
    u = u - 0.5 * (vx * vx + vy * vy + vz * vz)
    Y = U[5:] * dinv
    Y = np.where(Y >= 0.0, Y, 0.0)
    s = Y.sum(axis=0)
    Y = np.where(np.abs(s - 1.0) > 0.0, Y / s, Y)
    return np.vstack([rho[None], u[None], Y]).T       # [n][7]


def newton_labels(gas, n=60, seeds=DATA_SEEDS, guess_noise=0.05):
    """States of the seeded fields and the temperatures the device decode's
    Newton solve finds for them, starting from a perturbed guess."""
    X, T = [], []
    rng = np.random.default_rng(12345)
    for seed in seeds:
        geom, fields, bcs, _ = workloads.thermally_perfect_box_3d(gas, n=n, seed=seed)
        U, Tgen = workloads.initial_U_tp(gas, fields)
        Tguess = Tgen * (1.0 + guess_noise * rng.uniform(-1.0, 1.0, Tgen.shape))
        blk = SFluidBlock(gas, 3, n, n, n, 1.0 / n, 1.0 / n, 1.0 / n)
        for face, (kind, fs) in bcs.items():
            blk.set_bc(face, kind, fs)
        blk.upload(U, Tguess)
        _, fs = blk.download(with_fs=True)
        blk.destroy()
        X.append(decode_inputs(gas, U))
        T.append(fs[2].copy())
    return np.concatenate(X), np.concatenate(T)


class GuessModel:
    """Parameters (training layout) and normalisation of the guess network."""

    def __init__(self, params, in_mean, in_std, T_mean, T_std):
        self.params = _f32(params)
        self.in_mean = _f32(in_mean)
        self.in_std = _f32(in_std)      # multiplicative: z = (x - mean) * in_std
        self.T_mean = float(T_mean)
        self.T_std = float(T_std)

    @staticmethod
    def init(X, T, seed=7):
        rng = np.random.default_rng(seed)
        mean = X.mean(axis=0)
        std = X.std(axis=0)
        inv = np.where(std > 0, 1.0 / np.maximum(std, 1e-30), 0.0)
        p = np.zeros(NPARAM, dtype=np.float32)
        o = 0
        for fan_in, fan_out in ((8, NH), (NH, NH), (NH, NH)):
            lim = np.sqrt(6.0 / (fan_in + fan_out))
            p[o:o + fan_in * fan_out] = rng.uniform(-lim, lim, fan_in * fan_out)
            o += fan_in * fan_out + fan_out
        lim = np.sqrt(6.0 / (NH + 1))
        p[o:o + NH] = rng.uniform(-lim, lim, NH)
        return GuessModel(p, mean, inv, float(T.mean()), float(T.std()))

    def features(self, X):
        Z = np.zeros((X.shape[0], 8), dtype=np.float32)
        Z[:, :7] = (X - self.in_mean.astype(np.float64)) * self.in_std.astype(np.float64)
        return Z

    def targets(self, T):
        return _f32((T - self.T_mean) / self.T_std)

    def save(self, path=MODEL_PATH, **extra):
        np.savez(path, params=self.params, in_mean=self.in_mean, in_std=self.in_std, T_mean=self.T_mean,
                 T_std=self.T_std, **extra)

    @staticmethod
    def load(path=MODEL_PATH):
        d = np.load(path, allow_pickle=False)
        return GuessModel(d["params"], d["in_mean"], d["in_std"], float(d["T_mean"]), float(d["T_std"]))

    def install(self, gas):
        """Upload to the device as the decode's guess model of this gas."""
        _lib.check(_lib.load().e4hip_gas_tp_set_guess_model(
            gas.handle(), self.params.ctypes.data_as(_fp), self.in_mean.ctypes.data_as(_fp),
            self.in_std.ctypes.data_as(_fp), self.T_mean, self.T_std))


def mlp_grad(params, Z, t):
    """Loss and gradient of one batch with the device kernels."""
    params, Z, t = _f32(params), _f32(Z), _f32(t)
    g = np.zeros(NPARAM, dtype=np.float32)
    loss = C.c_float(0)
    _lib.init()
    _lib.check(_lib.load().e4hip_mlp_grad(params.ctypes.data_as(_fp), Z.ctypes.data_as(_fp), t.ctypes.data_as(_fp),
                                          Z.shape[0], C.byref(loss), g.ctypes.data_as(_fp)))
    return loss.value, g


def train(model, X, T, epochs=30, batch=4096, lr=3e-3, seed=1):
    Z = model.features(X)
    t = model.targets(T)
    losses = np.zeros(epochs, dtype=np.float32)
    _lib.init()
    _lib.check(_lib.load().e4hip_mlp_train(model.params.ctypes.data_as(_fp), Z.ctypes.data_as(_fp),
                                           t.ctypes.data_as(_fp), Z.shape[0], epochs, batch, lr, seed,
                                           losses.ctypes.data_as(_fp)))
    return losses


def device_guess(gas, X):
    X = np.ascontiguousarray(X, dtype=np.float64)
    T = np.zeros(X.shape[0])
    dp = C.POINTER(C.c_double)
    _lib.check(_lib.load().e4hip_tp_guess(gas.handle(), X.shape[0], X.ctypes.data_as(dp), T.ctypes.data_as(dp)))
    return T


def decode_stats(reset=True):
    """(energy evaluations, polish fall-backs) of the decode temperature solves."""
    ev, fb = C.c_longlong(0), C.c_longlong(0)
    _lib.check(_lib.load().e4hip_tp_decode_stats(C.byref(ev), C.byref(fb), int(reset)))
    return ev.value, fb.value


def face_stats(reset=True):
    """(energy evaluations, polish fall-backs) of the face-state temperature solves."""
    ev, fb = C.c_longlong(0), C.c_longlong(0)
    _lib.check(_lib.load().e4hip_tp_face_stats(C.byref(ev), C.byref(fb), int(reset)))
    return ev.value, fb.value


def main(argv=None):
    """Generate the training states (ten seeded 60^3 fields, 2.16 M states),
    train, report the held-out guess error, and save the model."""
    import argparse
    import time
    from .gas import ThermallyPerfectGas
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--n", type=int, default=60)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=3e-3)
    args = ap.parse_args(argv)
    gas = ThermallyPerfectGas.from_json()
    t0 = time.time()
    X, T = newton_labels(gas, n=args.n)
    print("states: %d (%.1f s)" % (X.shape[0], time.time() - t0), flush=True)
    model = GuessModel.init(X, T)
    t0 = time.time()
    losses = train(model, X, T, epochs=args.epochs, batch=args.batch, lr=args.lr)
    print("trained %d epochs in %.1f s; loss %s" % (args.epochs, time.time() - t0, losses), flush=True)
    model.install(gas)
    Xh, Th = newton_labels(gas, n=32, seeds=(workloads.SEED_TP_BOX_3D,))
    Tg = device_guess(gas, Xh)
    rms = float(np.sqrt(np.mean((Tg - Th) ** 2)))
    mx = float(np.max(np.abs(Tg - Th)))
    print("held-out field (seed %d): guess rms %.2f K, max %.2f K" % (workloads.SEED_TP_BOX_3D, rms, mx), flush=True)
    model.save(args.out, losses=losses, heldout_rms=rms, heldout_max=mx, n_states=X.shape[0])
    print("saved", args.out)


if __name__ == "__main__":
    main()
