"""Pins of oracle functions that no reference vector covers (CPU only):

* the temperature-guess network restatement (oracle/mlp_ref.py): the
  gradient of loss_and_grad against central finite differences of its own
  loss on every parameter block, the layer structure of forward against
  hand-computed single-path networks, and forward_bf16 against forward within
  a bound derived from BF16 rounding (8 significant bits);
* HLLC's star state (Toro 2009, section 10.4): an isolated contact
  discontinuity is resolved exactly -- a stationary one carries no mass or
  energy, a moving one gives the upwind side's convective flux -- which HLL,
  or a wrong star-state energy, does not.
"""
import math

import numpy as np
import pytest

import mlp_ref
import pyoracle as O

NPAR = 8961
BLOCKS = [("W1", 0, 512), ("b1", 512, 64), ("W2", 576, 4096), ("b2", 4672, 64), ("W3", 4736, 4096),
          ("b3", 8832, 64), ("w4", 8896, 64), ("b4", 8960, 1)]


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, scale, NPAR)
    return p


def _batch(seed, n=32):
    rng = np.random.default_rng(seed)
    Z = rng.normal(0.0, 1.0, (n, 8))
    Z[:, 7] = 0.0          # the eighth feature is zero in the trained model's layout
    t = rng.normal(0.0, 1.0, n)
    return Z, t


def test_block_layout_covers_all_parameters():
    assert sum(n for _, _, n in BLOCKS) == NPAR
    assert all(BLOCKS[i][1] + BLOCKS[i][2] == BLOCKS[i + 1][1] for i in range(len(BLOCKS) - 1))


@pytest.mark.parametrize("name,off,n", BLOCKS)
def test_mlp_gradient_matches_central_differences(name, off, n):
    """d loss / d p_i by central differences (h = 1e-6 |p_i|-scaled) on eight
    entries of each parameter block: a wrong index, transpose or missing
    activation derivative in one layer's backward fails its block."""
    p = _params(1)
    Z, t = _batch(2)
    _, g = mlp_ref.loss_and_grad(p, Z, t)
    rng = np.random.default_rng(off + 7)
    idx = off + rng.choice(n, size=min(8, n), replace=False)
    for i in idx:
        h = 1e-6 * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = mlp_ref.loss_and_grad(pp, Z, t)
        lm, _ = mlp_ref.loss_and_grad(pm, Z, t)
        fd = (lp - lm) / (2.0 * h)
        scale = max(abs(fd), 1e-6 * np.abs(g).max())
        assert abs(g[i] - fd) / scale < 1e-5, (name, int(i), g[i], fd)


def test_mlp_loss_is_mean_square_error_of_forward():
    p = _params(3)
    Z, t = _batch(4)
    y, _ = mlp_ref.forward(p, Z)
    loss, _ = mlp_ref.loss_and_grad(p, Z, t)
    assert loss == pytest.approx(float(np.mean((y - t) ** 2)), rel=1e-14)


def _silu(x):
    return x / (1.0 + math.exp(-x))


def test_mlp_forward_single_paths():
    """Networks whose output is computable by hand: W[out][in] row-major,
    biases per output, silu between layers, a linear output layer."""
    z = np.array([[0.7, -1.3, 0.2, 0.0, 0.5, -0.4, 1.1, 0.0]])
    # one path: input feature 2 -> hidden unit 5 -> unit 9 -> unit 40 -> output
    p = np.zeros(NPAR)
    p[0 + 5 * 8 + 2] = 1.5            # W1[5][2]
    p[512 + 5] = -0.25                # b1[5]
    p[576 + 9 * 64 + 5] = 2.0         # W2[9][5]
    p[4736 + 40 * 64 + 9] = -0.75     # W3[40][9]
    p[8832 + 40] = 0.1                # b3[40]
    p[8896 + 40] = 3.0                # w4[40]
    p[8960] = 0.05                    # b4
    h1 = _silu(1.5 * 0.2 - 0.25)
    h2 = _silu(2.0 * h1)
    h3 = _silu(-0.75 * h2 + 0.1)
    expect = 3.0 * h3 + 0.05
    y, _ = mlp_ref.forward(p, z)
    assert y[0] == pytest.approx(expect, rel=1e-14)
    # biases alone: every hidden unit of layer 3 at silu(b3), output w4 . silu(b3) + b4
    p = np.zeros(NPAR)
    rng = np.random.default_rng(5)
    p[8832:8896] = rng.normal(size=64)
    p[8896:8960] = rng.normal(size=64)
    p[8960] = -0.3
    expect = sum(p[8896 + k] * _silu(p[8832 + k]) for k in range(64)) - 0.3
    y, _ = mlp_ref.forward(p, z)
    assert y[0] == pytest.approx(expect, rel=1e-13)


def test_bf16_rounding_is_round_to_nearest_even():
    u = 2.0 ** -7                     # spacing of BF16 values in [1, 2): 8 significant bits
    x = np.array([1.0, 1.0 + u, 1.0 + 0.5 * u, 1.0 + 1.5 * u, 1.0 + 0.49 * u, -(1.0 + 0.51 * u), 3.0e38])
    r = mlp_ref.bf16(x)
    assert list(r[:6]) == [1.0, 1.0 + u, 1.0, 1.0 + 2 * u, 1.0, -(1.0 + u)]
    assert np.isfinite(r[6])


def test_bf16_matches_torch_bfloat16_cast():
    """mlp_ref.bf16 against torch's own float32 -> bfloat16 conversion (round
    to nearest even) on random values, ties and signed zeros."""
    import torch
    rng = np.random.default_rng(21)
    x = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e4, 2000), rng.normal(0, 1e-4, 2000),
                        [0.0, -0.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(2.0 + 2.0 ** -7)]]).astype(np.float32)
    ref = torch.from_numpy(x.copy()).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mlp_ref.bf16(x), ref)


def test_forward_bf16_structure_without_rounding(monkeypatch):
    """With the BF16 rounding replaced by the identity, forward_bf16 is the
    FP32 evaluation of forward (only float32 casts remain): a transposed or
    swapped layer would differ by O(1)."""
    monkeypatch.setattr(mlp_ref, "bf16", lambda x: np.asarray(x, dtype=np.float64))
    p = _params(13, scale=0.2)
    Z, _ = _batch(14, n=256)
    y, _ = mlp_ref.forward(p, Z)
    yb = mlp_ref.forward_bf16(p, Z)
    assert np.max(np.abs(yb - y)) <= 1e-5 * np.abs(y).max()


def test_forward_bf16_within_derived_rounding_bound():
    """|forward_bf16 - forward| <= a bound propagated layer by layer from BF16
    rounding (relative 2^-8 per rounding of a weight or an activation, all
    errors of one sign: a rigorous worst case) with silu's Lipschitz constant
    1.0998; the difference is not zero (the rounding is applied) and stays
    within 2 % of the output scale."""
    u = 2.0 ** -8
    lip = 1.0998
    p = _params(11, scale=0.2)
    Z, _ = _batch(12, n=256)
    W1, b1, W2, b2, W3, b3, w4, b4 = mlp_ref.unpack(p)
    y, (Z64, a1, h1, a2, h2, a3, h3) = mlp_ref.forward(p, Z)
    yb = mlp_ref.forward_bf16(p, Z)
    # e: bound on |h_bf16 - h| before the next layer rounds its input
    h0 = np.abs(Z64[:, :7])
    e = np.zeros_like(h0)
    hs = [(h0, np.abs(W1[:, :7])), (np.abs(h1), np.abs(W2)), (np.abs(h2), np.abs(W3))]
    for h, W in hs:
        e_in = e + u * (h + e)                          # rounding the layer input
        ea = e_in @ ((1.0 + u) * W).T + h @ (u * W).T   # rounded weights
        e = lip * ea                                    # silu
    e_out = (e + u * (np.abs(h3) + e)) @ np.abs(w4)     # the output layer reads rounded h3
    diff = np.abs(yb - y)
    assert np.all(diff <= e_out * (1.0 + 1e-9) + 1e-15), (diff.max(), e_out.min())
    assert diff.max() > 1e-5 * np.abs(y).max()
    assert diff.max() < 0.02 * np.abs(y).max()


G = 1.4
RGAS = 8.31451 / 0.02896


def _fs(gas, rho, p, vel):
    return O.flowstate_pT(gas, p, p / (rho * RGAS), vel)


def _hll(L, R):
    """HLL flux with HLLC's Davis wave speeds (Toro 10.21, 10.48)."""
    def U_F(s):
        rho, p, u, a, vn, vt1, vt2 = s[0], s[1], s[3], s[4], s[5], s[6], s[7]
        E = rho * (u + 0.5 * (vn * vn + vt1 * vt1 + vt2 * vt2))
        return (np.array([rho, rho * vn, rho * vt1, rho * vt2, E]),
                np.array([rho * vn, rho * vn * vn + p, rho * vn * vt1, rho * vn * vt2, vn * (E + p)]))
    (UL, FL), (UR, FR) = U_F(L), U_F(R)
    SL = min(L[5] - L[4], R[5] - R[4])
    SR = max(L[5] + L[4], R[5] + R[4])
    return (SR * FL - SL * FR + SL * SR * (UR - UL)) / (SR - SL)


@pytest.mark.parametrize("vt", [(0.0, 0.0), (35.0, -12.0)])
def test_hllc_resolves_a_stationary_contact(vt):
    """u_L = u_R = 0, p_L = p_R, rho_L != rho_R: mass and energy fluxes vanish,
    the normal momentum flux is p, the tangential ones vanish (Toro 10.4: the
    contact wave of speed S* = 0 separates the two star states).  HLL smears the
    same contact: its mass flux is S_L S_R (rho_R - rho_L) / (S_R - S_L)."""
    gas = O.Gas.ideal()
    for rhoL, rhoR, p in ((1.0, 0.125, 1.0e5), (0.3, 2.5, 4.0e4), (1.2, 1.1, 2.0e5)):
        L = _fs(gas, rhoL, p, (0.0, vt[0], vt[1]))
        R = _fs(gas, rhoR, p, (0.0, -vt[1], vt[0]))
        F = O.flux(gas, O.HLLC, L, R)
        a = max(L[4], R[4])
        assert abs(F[0]) <= 1e-13 * rhoL * a
        assert F[1] == pytest.approx(p, rel=1e-13)
        assert abs(F[2]) <= 1e-13 * rhoL * a * 100.0 and abs(F[3]) <= 1e-13 * rhoL * a * 100.0
        assert abs(F[4]) <= 1e-13 * p * a
        Fh = _hll(L, R)
        assert abs(Fh[0]) > 1e-3 * rhoL * a            # the property separates HLLC from HLL


def test_hllc_moving_contact_is_upwind_convection():
    """u_L = u_R = u0 (subsonic), p_L = p_R: the exact flux is the upwind
    side's convective flux plus the pressure terms, which pins the star-state
    energy (Toro eq. 10.39)."""
    gas = O.Gas.ideal()
    for u0 in (120.0, -90.0):
        L = _fs(gas, 1.0, 1.0e5, (u0, 10.0, -5.0))
        R = _fs(gas, 0.2, 1.0e5, (u0, -30.0, 8.0))
        F = O.flux(gas, O.HLLC, L, R)
        K = L if u0 > 0 else R
        rho, p, e = K[0], K[1], K[3]
        vel = np.array([u0, K[6], K[7]])
        E = rho * (e + 0.5 * float(vel @ vel))
        expect = np.array([rho * u0, rho * u0 * u0 + p, rho * u0 * vel[1], rho * u0 * vel[2], u0 * (E + p)])
        assert np.max(np.abs(F - expect) / np.maximum(np.abs(expect), 1.0)) < 1e-12
