"""eilmer4_amd: MI355X-native explicit gas-dynamic update of structured fluid
blocks (Eilmer4's gasdynamic_explicit_increment_with_fixed_grid) behind the C
ABI of include/e4hip.h.  The compute path is libe4hip.so (hand-written HIP for
gfx950); this package is the host side that drives it.
"""
from .fvcore import (BCType, Face, FlowSolverException, FluxCalculator, GasdynamicUpdate, InterpolateOption,
                     PrecisionMode, flux_calculator_from_name, flux_calculator_name, update_scheme_from_name)
from .gas import IdealGas, ThermallyPerfectGas
from .lib import E4hipError, load
from .fluidblock import SFluidBlock, connect_blocks
from .simcore import GlobalConfig, Simulation
from . import bc, flowfiles, udf

__all__ = ["BCType", "Face", "FlowSolverException", "FluxCalculator", "GasdynamicUpdate", "InterpolateOption",
           "PrecisionMode", "flux_calculator_from_name", "flux_calculator_name", "update_scheme_from_name",
           "IdealGas", "ThermallyPerfectGas", "E4hipError", "load", "SFluidBlock", "connect_blocks", "GlobalConfig",
           "Simulation", "bc", "flowfiles", "udf"]
