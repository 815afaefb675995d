"""Symbolic names of the selection API (eilmer/fvcore.d).

Enum values are the reference's (so a D caller passes ``cast(int)`` of its
own enum), with ``hllc`` appended to FluxCalculator.
"""
from enum import IntEnum


class FlowSolverException(RuntimeError):
    """Mirror of the reference's FlowSolverException (eilmer/fvcore.d)."""


class FluxCalculator(IntEnum):          # fvcore.d:184-194
    ausmdv = 0
    hanel = 1
    efm = 2
    ausm_plus_up = 3
    adaptive_efm_ausmdv = 4
    adaptive_hanel_ausmdv = 5
    adaptive_hlle_roe = 6
    hlle = 7
    roe = 8
    hllc = 9                            # new member (Toro 2009, section 10.4)


# calculators the device path implements
# (hlle is MHD only, fluxcalc.d:1197, and adaptive_hlle_roe takes it near shocks)
DEVICE_FLUX_CALCULATORS = (FluxCalculator.ausmdv, FluxCalculator.hanel, FluxCalculator.efm,
                           FluxCalculator.ausm_plus_up, FluxCalculator.adaptive_efm_ausmdv,
                           FluxCalculator.adaptive_hanel_ausmdv, FluxCalculator.roe, FluxCalculator.hllc)


def flux_calculator_name(fcalc):
    return FluxCalculator(fcalc).name


def flux_calculator_from_name(name):    # fvcore.d:213-229
    if name == "adaptive":
        return FluxCalculator.adaptive_efm_ausmdv
    try:
        return FluxCalculator[name]
    except KeyError:
        raise FlowSolverException("Invalid flux calculator name")


class GasdynamicUpdate(IntEnum):        # fvcore.d:30-41
    euler = 0
    pc = 1
    midpoint = 2
    classic_rk3 = 3
    tvd_rk3 = 4
    denman_rk3 = 5
    rkl1 = 6
    rkl2 = 7
    moving_grid_1_stage = 8
    moving_grid_2_stage = 9


_UPDATE_ALIASES = {"predictor_corrector": "pc", "predictor-corrector": "pc", "classic-rk3": "classic_rk3",
                   "tvd-rk3": "tvd_rk3", "denman-rk3": "denman_rk3",
                   "moving-grid-1-stage": "moving_grid_1_stage", "moving-grid-2-stage": "moving_grid_2_stage"}


def update_scheme_from_name(name):      # fvcore.d:95-117
    name = _UPDATE_ALIASES.get(name, name)
    try:
        return GasdynamicUpdate[name]
    except KeyError:
        raise FlowSolverException("Invalid gasdynamic update scheme name")


def number_of_stages_for_update_scheme(gdut):   # fvcore.d:61-75
    return {0: 1, 1: 2, 2: 2, 3: 3, 4: 3, 5: 3, 6: 3, 7: 3, 8: 1, 9: 2}[int(gdut)]


class InterpolateOption(IntEnum):       # fvcore.d:155
    pt = 0
    rhou = 1
    rhop = 2
    rhot = 3


class Face(IntEnum):                    # fvcore.d Face
    north = 0
    east = 1
    south = 2
    west = 3
    top = 4
    bottom = 5


class BCType(IntEnum):
    """Ghost-cell boundary effects available on the device path."""
    slip_wall = 0          # WallBC_WithSlip (internal_copy_then_reflect)
    inflow = 1             # InFlowBC_Supersonic (flow_state_copy)
    extrapolate = 2        # OutFlowBC_SimpleExtrapolate, x_order 0
    periodic = 3           # ExchangeBC_FullFace of a block with itself
    connected = 4          # ExchangeBC_FullFace to another block (SFluidBlock.connect)
    outflow_simple_flux = 5  # OutFlowBC_SimpleFlux / OutFlowBC_Simple (bc.lua:1124-1148)
    slip_wall_1 = 6        # WallBC_WithSlip1: no ghost data, one-sided reconstruction, wall flux
    extrapolate_1 = 7      # OutFlowBC_SimpleExtrapolate{xOrder=1}
    inflow_profile = 8     # InFlowBC_StaticProfile (states from flowfiles.profile_ghost_states)


class PrecisionMode(IntEnum):
    """Arithmetic of the reconstruction and flux path (E4HIP_PRECISION_*)."""
    fp64 = 0           # plain FP64: FMA contraction, reciprocal-based division
    fp64_exact = 1     # FP64 with the reference's roundings
    mixed = 2          # FP32 reconstruction and flux, FP64 accumulate/update/decode
