"""Time-stepping coordinator over the device library, mirroring the calls of
eilmer/simcore.d that touch this path:

  determine_time_step_size()                     simcore.d:1347-1452
  gasdynamic_explicit_increment_with_fixed_grid() simcore.d:2303-2952
  integrate_in_time(target_time)                  simcore.d:791-1144 (loop 856-1131)

The per-block work is one C-ABI call per stage sequence; the host keeps the
reference's dt policy and bad-cell policy.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import lib as _lib
from .fvcore import (FlowSolverException, FluxCalculator, GasdynamicUpdate, InterpolateOption, PrecisionMode,
                     DEVICE_FLUX_CALCULATORS, number_of_stages_for_update_scheme)


@dataclass
class GlobalConfig:
    """The GlobalConfig/LocalConfig fields the path reads, with the reference
    defaults (eilmer/globalconfig.d:440-569, 585, 600, 707-734)."""
    flux_calculator: int = FluxCalculator.adaptive_hanel_ausmdv
    gasdynamic_update_scheme: int = GasdynamicUpdate.pc
    interpolation_order: int = 2
    apply_limiter: bool = True
    extrema_clipping: bool = True
    interpolate_in_local_frame: bool = True
    thermo_interpolator: int = InterpolateOption.rhou
    precision_mode: int = PrecisionMode.fp64
    max_invalid_cells: int = 0
    learned_eos_guess: bool = False
    batch_blocks: bool = True
    shear_tolerance: float = 0.20           # globalconfig.d:585
    compression_tolerance: float = -0.30    # globalconfig.d:600
    cfl_value: float = 0.5
    cfl_count: int = 10
    dt_init: float = 1.0e-3
    dt_max: float = 1.0e-3
    max_time: float = 1.0e-3
    max_step: int = 100
    fixed_time_step: bool = False
    stringent_cfl: bool = False             # globalconfig.d:729
    M_inf: float = 0.01                     # ausm_plus_up, globalconfig.d:591
    autotune: bool = True                   # k-chunk heights of the stage kernel timed per block shape
    use_graphs: bool = False                # replay repeated steps from a captured HIP graph

    def step_config(self):
        if FluxCalculator(self.flux_calculator) not in DEVICE_FLUX_CALCULATORS:
            raise FlowSolverException("flux_calculator %s is not available on the device path"
                                      % FluxCalculator(self.flux_calculator).name)
        return _lib.StepConfig(int(self.flux_calculator), int(self.gasdynamic_update_scheme),
                               int(self.interpolation_order), int(bool(self.apply_limiter)),
                               int(bool(self.extrema_clipping)), int(bool(self.interpolate_in_local_frame)),
                               int(self.thermo_interpolator), int(self.precision_mode), int(self.max_invalid_cells),
                               int(bool(self.learned_eos_guess)), int(bool(self.batch_blocks)),
                               float(self.shear_tolerance), float(self.compression_tolerance),
                               int(bool(self.stringent_cfl)), int(bool(self.use_graphs)), float(self.M_inf),
                               int(bool(self.autotune)))


class SimState:
    def __init__(self, dt_init):
        self.time = 0.0
        self.step = 0
        self.dt_global = dt_init
        self.dt_allow = 0.0
        self.cfl_max = 0.0


class Simulation:
    """One simulation on one GPU: blocks, config and the time loop."""

    def __init__(self, blocks, config=None, device=0):
        _lib.check(_lib.load().e4hip_init(device))
        self.blocks = list(blocks)
        self.cfg = config or GlobalConfig()
        self.apply_config()
        self.state = SimState(self.cfg.dt_init)
        self.last_bad_cells = [0, 0, 0]
        self.last_bad_cells_per_block = []

    def apply_config(self):
        sc = self.cfg.step_config()
        _lib.check(_lib.load().e4hip_config(C.byref(sc)))

    def _ids(self):
        arr = (C.c_int * len(self.blocks))(*[b.id for b in self.blocks])
        return arr, len(self.blocks)

    def block_dt(self, dt_current, check_cfl):
        ids, n = self._ids()
        dt_allow, cfl_max = C.c_double(0), C.c_double(0)
        _lib.check(_lib.load().e4hip_determine_dt(ids, n, dt_current, self.cfg.cfl_value, int(check_cfl),
                                                  C.byref(dt_allow), C.byref(cfl_max)))
        return dt_allow.value, cfl_max.value

    def determine_time_step_size(self):
        """simcore.d:1347-1452 without super/local time-stepping or dt_override."""
        S, cfg = self.state, self.cfg
        if S.step % cfg.cfl_count != 0:
            return
        S.dt_allow, S.cfl_max = self.block_dt(S.dt_global, S.step > 0)
        if S.step == 0:
            S.dt_allow = min(cfg.dt_init, S.dt_allow)
        if S.dt_allow <= S.dt_global:
            S.dt_global = S.dt_allow
        else:
            S.dt_global = min(S.dt_global * 1.5, S.dt_allow)
            S.dt_global = min(S.dt_global, cfg.dt_max)

    def gasdynamic_explicit_increment_with_fixed_grid(self):
        ids, n = self._ids()
        nbad = (C.c_int * 3)()
        L = _lib.load()
        rc = L.e4hip_explicit_increment(ids, n, self.state.dt_global, nbad)
        self.last_bad_cells = list(nbad)
        per = (C.c_int * (3 * n))()
        if L.e4hip_last_step_bad_cells(n, per) == 0:
            self.last_bad_cells_per_block = [list(per[3 * b:3 * b + 3]) for b in range(n)]
        if rc == -4:
            for b, counts in enumerate(self.last_bad_cells_per_block):
                for s, c in enumerate(counts):
                    if c > self.cfg.max_invalid_cells:
                        print("Following stage-%d gasdynamic update: %d bad cells in block[%d]." % (s + 1, c, b))
            raise FlowSolverException("Too many bad cells; go home.")
        if rc == -5:
            raise FlowSolverException("Bad cell with negative mass.")
        _lib.check(rc)
        self.state.time = self.state.time + self.state.dt_global

    def integrate_in_time(self, target_time=None, max_step=None):
        cfg, S = self.cfg, self.state
        target_time = cfg.max_time if target_time is None else target_time
        max_step = cfg.max_step if max_step is None else max_step
        finished = False
        while not finished:
            if not cfg.fixed_time_step:
                self.determine_time_step_size()
            self.gasdynamic_explicit_increment_with_fixed_grid()
            S.step += 1
            if S.time >= target_time:
                finished = True
            if S.step >= max_step:
                finished = True
        return S.step

    def synchronize(self):
        _lib.check(_lib.load().e4hip_synchronize())

    def stages(self):
        return number_of_stages_for_update_scheme(self.cfg.gasdynamic_update_scheme)
