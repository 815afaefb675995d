"""Gas models on the host side: parameters for the device library and the
state constructors the input layer needs (FlowState from p and T).

IdealGas follows gas/ideal_gas.d:31-144 (R = R_universal/mMass, Cv = R/(gamma-1)).
"""
import json
import os

import numpy as np

from . import lib as _lib

R_universal = 8.31451                 # gas/physical_constants.d:13
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


class IdealGas:
    def __init__(self, mMass=0.02896, gamma=1.4, name="air"):
        self.name = name
        self.mMass = float(mMass)
        self.gamma = float(gamma)
        self.Rgas = R_universal / self.mMass
        self.Cv = self.Rgas / (self.gamma - 1.0)
        self.Cvinv = 1.0 / self.Cv
        self.Cp = self.Rgas * self.gamma / (self.gamma - 1.0)
        self.n_species = 1
        self._handle = None

    @staticmethod
    def from_json(name="ideal-air-gas-model.json"):
        path = name if os.path.isabs(name) else os.path.join(DATA, name)
        d = json.load(open(path))
        return IdealGas(d["mMass"], d["gamma"], d.get("speciesName", "air"))

    def handle(self):
        """Device gas-model handle (created on first use)."""
        if self._handle is None:
            import ctypes as C
            h = C.c_int(-1)
            _lib.init()
            _lib.check(_lib.load().e4hip_gas_ideal(self.mMass, self.gamma, C.byref(h)))
            self._handle = h.value
        return self._handle

    # update_thermo_from_pT (ideal_gas.d:84-92) and update_sound_speed (137-144)
    def rho_u_a_from_pT(self, p, T):
        p = np.asarray(p, dtype=np.float64)
        T = np.asarray(T, dtype=np.float64)
        rho = p / (T * self.Rgas)
        u = self.Cv * T
        a = np.sqrt(self.gamma * self.Rgas * T)
        return rho, u, a

    def flowstate(self, p, T, vel=(0.0, 0.0, 0.0)):
        """FlowState vector [rho,p,T,u,a,vx,vy,vz,massf] built from p, T and velocity."""
        rho, u, a = self.rho_u_a_from_pT(p, T)
        return np.array([rho, p, T, u, a, vel[0], vel[1], vel[2], 1.0], dtype=np.float64)

    def encode(self, rho, u, vx, vy, vz):
        """encode_conserved (eilmer/fvcell.d:494-560) on arrays -> U [5][n]."""
        rho = np.asarray(rho, dtype=np.float64).ravel()
        vx, vy, vz, u = (np.asarray(a, dtype=np.float64).ravel() for a in (vx, vy, vz, u))
        U = np.empty((5, rho.size))
        U[0] = rho
        U[1] = rho * vx
        U[2] = rho * vy
        U[3] = rho * vz
        ke = 0.5 * (vx * vx + vy * vy + vz * vz)
        U[4] = rho * (u + ke)
        return U

    def sound_speed(self, T):
        return np.sqrt(self.gamma * self.Rgas * np.asarray(T, dtype=np.float64))


class ThermallyPerfectGas:
    """Thermally perfect mixture with NASA-9 (CEA) curves
    (gas/therm_perf_gas.d:41-110, gas/thermo/cea_thermo_curves.d).  The host
    side evaluates energies for building initial states; the temperature solve
    of the update runs on the device."""

    def __init__(self, desc):
        self.desc = desc
        self.species = list(desc["species"])
        self.n_species = len(self.species)
        self.M = np.array(desc["M"], dtype=np.float64)
        self.R = R_universal / self.M
        self.nseg = int(desc["thermo"][0]["nsegments"])
        self.Tb = np.array([t["T_break_points"] for t in desc["thermo"]], dtype=np.float64)
        self.Tbl = np.array([t["T_blend_ranges"] for t in desc["thermo"]], dtype=np.float64)
        self.coeffs = np.array([t["segments"] for t in desc["thermo"]], dtype=np.float64)   # [nsp][nseg][9]
        self._handle = None

    @staticmethod
    def from_json(name="therm-perf-5-species-air.json"):
        path = name if os.path.isabs(name) else os.path.join(DATA, name)
        return ThermallyPerfectGas(json.load(open(path)))

    def handle(self):
        if self._handle is None:
            import ctypes as C
            dp = C.POINTER(C.c_double)
            h = C.c_int(-1)
            _lib.init()
            args = [np.ascontiguousarray(a.ravel()) for a in (self.M, self.Tb, self.Tbl, self.coeffs)]
            self._args = args
            _lib.check(_lib.load().e4hip_gas_tp(self.n_species, args[0].ctypes.data_as(dp), self.nseg,
                                               args[1].ctypes.data_as(dp), args[2].ctypes.data_as(dp),
                                               args[3].ctypes.data_as(dp), C.byref(h)))
            self._handle = h.value
        return self._handle

    def _coeffs_at(self, s, T):
        """Blended coefficient sets (cea_thermo_curves.d:136-169) for an array of T."""
        Tb, Tbl, c = self.Tb[s], self.Tbl[s], self.coeffs[s]
        nb = self.nseg + 1
        a = np.empty(T.shape + (9,))
        seg = np.zeros(T.shape, dtype=int)
        blend = np.zeros(T.shape, dtype=bool)
        wB = np.zeros(T.shape)
        seg[T > Tb[nb - 2] + 0.5 * Tbl[nb - 3]] = self.nseg - 1
        for i in range(1, nb - 1):
            lo, hi = Tb[i] - 0.5 * Tbl[i - 1], Tb[i] + 0.5 * Tbl[i - 1]
            inb = (T >= lo) & (T <= hi)
            blend |= inb
            wB = np.where(inb, (1.0 / Tbl[i - 1]) * (T - lo), wB)
            seg = np.where(inb, i, seg)
            if i < nb - 2:      # the last interior break's upper test is never reached
                above = (T > hi) & (T < Tb[i + 1] - 0.5 * Tbl[i])
                seg = np.where(above, i, seg)
        a[:] = c[seg]
        bl = np.where(blend)
        a[bl] = ((1.0 - wB[bl])[:, None] * c[seg[bl] - 1] + wB[bl][:, None] * c[seg[bl]])
        return a

    def eval_h(self, s, T):
        T = np.asarray(T, dtype=np.float64)
        lo, hi = self.Tb[s][0], self.Tb[s][-1]
        Tc = np.clip(T, lo, hi)
        a = self._coeffs_at(s, Tc)
        h = self.R[s] * Tc * (-a[..., 0] / (Tc * Tc) + a[..., 1] / Tc * np.log(Tc) + a[..., 2] + a[..., 3] * Tc / 2.0 +
                              a[..., 4] * Tc * Tc / 3.0 + a[..., 5] * Tc ** 3 / 4.0 + a[..., 6] * Tc ** 4 / 5.0 +
                              a[..., 7] / Tc)
        Cp = self.eval_Cp(s, Tc)
        return h + Cp * (T - Tc)

    def eval_Cp(self, s, T):
        T = np.clip(np.asarray(T, dtype=np.float64), self.Tb[s][0], self.Tb[s][-1])
        a = self._coeffs_at(s, T)
        return self.R[s] * (a[..., 0] / (T * T) + a[..., 1] / T + a[..., 2] + a[..., 3] * T + a[..., 4] * T * T +
                            a[..., 5] * T ** 3 + a[..., 6] * T ** 4)

    def energy(self, T, Y):
        """u = sum_s Y_s (h_s(T) - R_s T), Y as [nsp][n]."""
        return sum(Y[s] * (self.eval_h(s, T) - self.R[s] * T) for s in range(self.n_species))

    def Rmix(self, Y):
        return sum(Y[s] * self.R[s] for s in range(self.n_species))

    def sound_speed(self, T, Y):
        Cp = sum(Y[s] * self.eval_Cp(s, T) for s in range(self.n_species))
        Rm = self.Rmix(Y)
        return np.sqrt(Cp / (Cp - Rm) * Rm * T)

    def flowstate(self, p, T, vel=(0.0, 0.0, 0.0), Y=None):
        Y = np.asarray(Y, dtype=np.float64).reshape(self.n_species, 1)
        Rm = self.Rmix(Y)[0]
        rho = p / (Rm * T)
        u = float(self.energy(np.array([T]), Y)[0])
        a = float(self.sound_speed(np.array([T]), Y)[0])
        return np.concatenate([[rho, p, T, u, a, vel[0], vel[1], vel[2]], Y[:, 0]])

    def encode(self, rho, u, vx, vy, vz, Y):
        """encode_conserved with species (fvcell.d:494-560) -> U [5+nsp][n]."""
        rho = np.asarray(rho, dtype=np.float64).ravel()
        vx, vy, vz, u = (np.asarray(a, dtype=np.float64).ravel() for a in (vx, vy, vz, u))
        Y = np.asarray(Y, dtype=np.float64).reshape(self.n_species, -1)
        U = np.empty((5 + self.n_species, rho.size))
        U[0] = rho
        U[1] = rho * vx
        U[2] = rho * vy
        U[3] = rho * vz
        U[4] = rho * (u + 0.5 * (vx * vx + vy * vy + vz * vz))
        U[5:] = rho * Y
        return U
