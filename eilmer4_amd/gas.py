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
