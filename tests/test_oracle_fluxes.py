"""Pins of the oracle's efm, ausm_plus_up, roe and adaptive_efm_ausmdv
(fluxcalc.d:626-1037, 1202-1371) by what the mathematics fixes, beyond the
consistency, upwinding, mirror-symmetry, conservation and Riemann-problem
checks they share with the other calculators (test_oracle_invariants.py):

  exxef        the polynomial erf of Abramowitz and Stegun 7.1.26 (max error
               1.5e-7) and exp(-x^2) exactly; both clamped beyond |x| = 5;
  efm          the mass flux of a gas at rest is rho sqrt(RT / (2 pi)) per
               side (the one-sided Maxwellian flux, Pullin 1980);
  roe          a stationary contact (u = 0, equal p, rho jump) is resolved
               exactly; the entropy fix is inactive for a compression;
  ausm_plus_up the low-Mach pressure-diffusion term of Liou 2006 eq. 73 on a
               pressure jump at rest, and p_half = (pL + pR) / 2 (eq. 75);
  adaptive_efm_ausmdv  efm at flagged faces with small shear, else ausmdv.
"""
import math

import numpy as np
import pytest

import pyoracle as O

RGAS = 8.31451 / 0.02896
G = 1.4


def fs(rho, p, vel, gas=None):
    gas = gas or O.Gas.ideal()
    return O.flowstate_pT(gas, p, p / (rho * RGAS), vel)


def test_exxef_is_the_abramowitz_stegun_erf():
    xs = np.concatenate([np.linspace(-6, 6, 241), [0.0, 1e-9, -1e-9, 4.999, 5.0, 5.001]])
    for x in xs:
        ex, ef = O.exxef(float(x))
        if abs(x) > 5.0:
            assert ex == 0.138879e-10 and ef == math.copysign(1.0, x)
        else:
            assert ex == math.exp(-x * x)
            assert abs(ef - math.erf(x)) <= 1.5e-7
            assert math.copysign(1.0, ef) == math.copysign(1.0, x) or ef == 0.0


def test_efm_mass_flux_of_a_gas_at_rest():
    """Both sides at rest: each side sends rho sqrt(RT/(2 pi)) through the
    face, with dtwspi = 1/(2 sqrt(pi)) to the reference's nine digits."""
    gas = O.Gas.ideal()
    rng = np.random.default_rng(11)
    for _ in range(50):
        rL, rR = rng.uniform(0.2, 3.0, 2)
        pL, pR = rng.uniform(2e4, 3e5, 2)
        L, R = fs(rL, pL, (0.0, 0.0, 0.0), gas), fs(rR, pR, (0.0, 0.0, 0.0), gas)
        F = O.flux(gas, O.EFM, L, R)
        # the mixture's R is the sqrt(rho)-weighted mean of p/(rho T), here RGAS
        cL, cR = math.sqrt(2 * RGAS * L[2]), math.sqrt(2 * R[4] ** 2 / G)
        expect = (rL * cL - rR * cR) / (2 * math.sqrt(math.pi))
        assert abs(F[0] - expect) <= 1e-8 * (rL * cL + rR * cR)
        # pressure: half of each side's, up to the polynomial's error at 0
        # (erf(0) = 1 - (A1 + ... + A5) = 1e-9, within its 1.5e-7 bound)
        assert abs(F[1] - 0.5 * (pL + pR)) <= 0.5 * 1.5e-7 * abs(pL - pR) + 1e-12 * (pL + pR)
        assert F[2] == 0.0 and F[3] == 0.0


@pytest.mark.parametrize("vt", [(0.0, 0.0), (35.0, -12.0)])
def test_roe_resolves_a_stationary_contact(vt):
    gas = O.Gas.ideal()
    rng = np.random.default_rng(12)
    for _ in range(50):
        p = rng.uniform(2e4, 3e5)
        rL, rR = rng.uniform(0.2, 3.0, 2)
        L = fs(rL, p, (0.0, vt[0], vt[1]), gas)
        R = fs(rR, p, (0.0, vt[0], vt[1]), gas)
        F = O.flux(gas, O.ROE, L, R)
        assert abs(F[0]) <= 1e-12 * rL * 300
        assert abs(F[1] - p) <= 1e-12 * p
        assert abs(F[2]) <= 1e-9 and abs(F[3]) <= 1e-9
        assert abs(F[4]) <= 1e-12 * p * 300


def test_roe_compression_is_the_plain_roe_flux():
    """Where u - a falls across the face the entropy fix is off: the flux is
    (FL + FR)/2 - sum alpha |lambda| K / 2, which for two supersonic states
    moving right equals the left flux; for a subsonic compression it differs
    from both sides' fluxes but is conservative: F(L,R) - F(L,L) and
    F(R,R) - F(L,R) add up to F(R,R) - F(L,L)."""
    gas = O.Gas.ideal()
    L = fs(1.0, 1.0e5, (200.0, 0.0, 0.0), gas)
    R = fs(1.5, 1.6e5, (100.0, 0.0, 0.0), gas)
    assert (R[5] - R[4]) < (L[5] - L[4])
    F = O.flux(gas, O.ROE, L, R)
    # the Roe average is exact for this jump: A_hat (UR - UL) = FR - FL, so the
    # flux is also FL + A_hat^- (UR - UL) = FR - A_hat^+ (UR - UL); check the
    # two one-sided forms through the flux of the swapped, mirrored problem
    Lm, Rm = R.copy(), L.copy()
    Lm[5], Rm[5] = -Lm[5], -Rm[5]
    Fm = O.flux(gas, O.ROE, Lm, Rm)
    s = np.abs(F).max()
    assert abs(F[0] + Fm[0]) / s < 1e-12 and abs(F[1] - Fm[1]) / s < 1e-12 and abs(F[4] + Fm[4]) / s < 1e-12


def test_ausm_plus_up_pressure_diffusion_at_rest():
    """uL = uR = 0, equal densities, a pressure jump: MbarSq = 0, so M0^2 =
    M_inf and fa = sqrt(M_inf) (2 - sqrt(M_inf)); the mass flux is
    a_half Mp rho with Mp = -Kp/fa (pR - pL)/(rho a_half^2) (Liou 2006 eq. 73),
    and p_half = (pL + pR)/2 since P5+(0) = P5-(0) = 1/2 (eq. 75)."""
    gas = O.Gas.ideal()
    for M_inf in (0.01, 0.09, 0.25):
        for pL, pR in ((1.0e5, 1.2e5), (1.5e5, 0.9e5)):
            rho = 1.1
            L, R = fs(rho, pL, (0.0, 0.0, 0.0), gas), fs(rho, pR, (0.0, 0.0, 0.0), gas)
            F = O.flux_opt(gas, O.AUSM_PLUS_UP, L, R, M_inf=M_inf)
            a_half = 0.5 * (L[4] + R[4])
            fa = math.sqrt(M_inf) * (2.0 - math.sqrt(M_inf))
            # a_half Mp rho = -Kp/fa (pR - pL)/a_half
            mass = -0.25 / fa * (pR - pL) / a_half
            assert abs(F[0] - mass) <= 1e-12 * abs(mass)
            assert abs(F[1] - 0.5 * (pL + pR)) <= 1e-12 * pL
            # the energy follows the upwind total enthalpy
            H = (L if F[0] >= 0 else R)[3] + (L if F[0] >= 0 else R)[1] / rho
            assert abs(F[4] - F[0] * H) <= 1e-12 * abs(F[0] * H)


def test_ausm_plus_up_default_m_inf():
    """flux() uses the reference's default M_inf = 0.01 (globalconfig.d:591)."""
    gas = O.Gas.ideal()
    L, R = fs(1.0, 1.0e5, (30.0, 5.0, 0.0), gas), fs(1.2, 1.1e5, (10.0, 0.0, 2.0), gas)
    assert np.array_equal(O.flux(gas, O.AUSM_PLUS_UP, L, R), O.flux_opt(gas, O.AUSM_PLUS_UP, L, R, M_inf=0.01))
    assert not np.array_equal(O.flux(gas, O.AUSM_PLUS_UP, L, R),
                              O.flux_opt(gas, O.AUSM_PLUS_UP, L, R, M_inf=0.2))


def test_adaptive_efm_ausmdv_switch():
    gas = O.Gas.ideal()
    L, R = fs(1.0, 1.0e5, (300.0, 20.0, 0.0), gas), fs(2.0, 3.0e5, (100.0, 25.0, 0.0), gas)
    efm, ausmdv = O.flux(gas, O.EFM, L, R), O.flux(gas, O.AUSMDV, L, R)
    assert not np.array_equal(efm, ausmdv)
    assert np.array_equal(O.flux_opt(gas, O.ADAPTIVE_EFM_AUSMDV, L, R, SL=0, SR=0), ausmdv)
    assert np.array_equal(O.flux_opt(gas, O.ADAPTIVE_EFM_AUSMDV, L, R, SL=1, SR=0), efm)
    assert np.array_equal(O.flux_opt(gas, O.ADAPTIVE_EFM_AUSMDV, L, R, SL=0, SR=1), efm)
    # shear |vt1 jump| / mean a above the tolerance keeps AUSMDV at a shock point
    Rs = fs(2.0, 3.0e5, (100.0, 400.0, 0.0), gas)
    assert np.array_equal(O.flux_opt(gas, O.ADAPTIVE_EFM_AUSMDV, L, Rs, SL=1, SR=1),
                          O.flux(gas, O.AUSMDV, L, Rs))


@pytest.mark.parametrize("calc", [O.EFM, O.AUSM_PLUS_UP, O.ROE])
def test_mixture_species_follow_the_mass_flux(calc):
    """Five-species air: species fluxes are the mass flux times the upwind
    side's mass fractions, and the gas-dynamic part is that of the same
    states' flux (consistency for equal states)."""
    tp = O.Gas.from_json("therm-perf-5-species-air.json")
    rng = np.random.default_rng(13)
    for _ in range(20):
        YL, YR = rng.dirichlet(np.ones(5) * 3), rng.dirichlet(np.ones(5) * 3)
        TL, TR = rng.uniform(400, 5000, 2)
        L = O.flowstate_pT(tp, rng.uniform(2e4, 2e5), TL, rng.uniform(-300, 300, 3), YL)
        R = O.flowstate_pT(tp, rng.uniform(2e4, 2e5), TR, rng.uniform(-300, 300, 3), YR)
        F = O.flux(tp, calc, L, R)
        up = YL if (F[0] > 0 or (F[0] == 0 and calc != O.EFM)) else YR
        assert np.allclose(F[5:], F[0] * up, rtol=1e-14, atol=0)
        Fs = O.flux(tp, calc, L, L)
        E = L[0] * (L[3] + 0.5 * (L[5] ** 2 + L[6] ** 2 + L[7] ** 2))
        phys = np.array([L[0] * L[5], L[0] * L[5] ** 2 + L[1], L[0] * L[5] * L[6], L[0] * L[5] * L[7],
                         L[5] * (E + L[1])])
        assert np.max(np.abs(Fs[:5] - phys)) <= 1e-11 * np.abs(phys).max()


def _roe_dissipation_by_eigendecomposition(L, R, gamma=G):
    """|A_hat| (UR - UL) with A_hat the x-direction Euler flux Jacobian at the
    Roe average (Toro 2009, eqs. 3.79 and 11.60), |A| = V |Lambda| V^-1 from
    numpy's eigendecomposition."""
    def cons(s):
        rho, p, u, v, w = s[0], s[1], s[5], s[6], s[7]
        return np.array([rho, rho * u, rho * v, rho * w, p / (gamma - 1) + 0.5 * rho * (u * u + v * v + w * w)])
    sl, sr = math.sqrt(L[0]), math.sqrt(R[0])
    H = [s[3] + s[1] / s[0] + 0.5 * (s[5] ** 2 + s[6] ** 2 + s[7] ** 2) for s in (L, R)]
    u, v, w = [(sl * L[5 + q] + sr * R[5 + q]) / (sl + sr) for q in range(3)]
    Hh = (sl * H[0] + sr * H[1]) / (sl + sr)
    g1 = gamma - 1
    q2 = u * u + v * v + w * w
    A = np.array([[0, 1, 0, 0, 0],
                  [g1 * q2 / 2 - u * u, (3 - gamma) * u, -g1 * v, -g1 * w, g1],
                  [-u * v, v, u, 0, 0],
                  [-u * w, w, 0, u, 0],
                  [u * (g1 * q2 / 2 - Hh), Hh - g1 * u * u, -g1 * u * v, -g1 * u * w, gamma * u]])
    lam, V = np.linalg.eig(A)
    absA = (V @ np.diag(np.abs(lam)) @ np.linalg.inv(V)).real
    return absA @ (cons(R) - cons(L))


def test_roe_flux_against_the_eigendecomposition_of_the_roe_matrix():
    """Where the entropy fix is off (u - a does not grow across the face),
    roe = (FL + FR)/2 - |A_hat| (UR - UL)/2 with |A_hat| from an independent
    eigendecomposition of the Roe-averaged Jacobian."""
    gas = O.Gas.ideal()
    rng = np.random.default_rng(14)
    done = 0
    for _ in range(400):
        L = fs(rng.uniform(0.3, 3), rng.uniform(2e4, 3e5), rng.uniform(-500, 500, 3), gas)
        R = fs(rng.uniform(0.3, 3), rng.uniform(2e4, 3e5), rng.uniform(-500, 500, 3), gas)
        if (R[5] - R[4]) > (L[5] - L[4]):
            continue
        F = O.flux(gas, O.ROE, L, R)
        phys = []
        for s in (L, R):
            E = s[0] * (s[3] + 0.5 * (s[5] ** 2 + s[6] ** 2 + s[7] ** 2))
            phys.append(np.array([s[0] * s[5], s[0] * s[5] ** 2 + s[1], s[0] * s[5] * s[6], s[0] * s[5] * s[7],
                                  s[5] * (E + s[1])]))
        expect = 0.5 * (phys[0] + phys[1]) - 0.5 * _roe_dissipation_by_eigendecomposition(L, R)
        scale = np.abs(phys[0]).max() + np.abs(phys[1]).max()
        assert np.max(np.abs(F - expect)) <= 1e-10 * scale
        done += 1
    assert done > 100
