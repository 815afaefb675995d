"""The boundary conditions of the reference's input scripts (bc.lua), under
their names, for SFluidBlock.set_bc: each returns the device boundary kind and
its data.  Connections (ExchangeBC_FullFace) are made with
SFluidBlock.connect / connect_blocks.

    blk.set_bc(Face.west, InFlowBC_Supersonic(gas.flowstate(p, T, vel)))
    blk.set_bc(Face.east, OutFlowBC_Simple())
    blk.set_bc(Face.north, WallBC_WithSlip1())
    blk.set_bc(Face.west, InFlowBC_StaticProfile("profile.dat", gas, match="xyz-to-xyz"))
"""
from dataclasses import dataclass

import numpy as np

from .fvcore import BCType


@dataclass
class BoundaryCondition:
    kind: BCType
    fs: object = None          # a flow state vector, or a function of (block, face) giving one


def WallBC_WithSlip0():
    """bc.lua:528-552: mirror ghost cells with the normal velocity reversed."""
    return BoundaryCondition(BCType.slip_wall)


WallBC_WithSlip = WallBC_WithSlip0          # bc.lua:607


def WallBC_WithSlip1():
    """bc.lua:554-577: no ghost-cell data; one-sided reconstruction and the
    wall flux (needs at least 3 cells across the wall)."""
    return BoundaryCondition(BCType.slip_wall_1)


def InFlowBC_Supersonic(flowState):
    """flow_state_copy.d: every ghost cell holds flowState ([rho, p, T, u, a,
    vx, vy, vz, massf...], e.g. gas.flowstate(p, T, vel))."""
    return BoundaryCondition(BCType.inflow, np.ascontiguousarray(flowState, dtype=np.float64))


def OutFlowBC_SimpleExtrapolate(xOrder=0):
    """bc.lua:1103-1122: ghost cells extrapolated from the cells inside
    (copied for xOrder 0, linearly for xOrder 1)."""
    if xOrder not in (0, 1):
        raise ValueError("xOrder must be 0 or 1")
    return BoundaryCondition(BCType.extrapolate if xOrder == 0 else BCType.extrapolate_1)


def OutFlowBC_SimpleFlux():
    """bc.lua:1124-1146: extrapolated ghosts and the boundary face's flux from
    the interior cell (closed to inflow)."""
    return BoundaryCondition(BCType.outflow_simple_flux)


OutFlowBC_Simple = OutFlowBC_SimpleFlux      # bc.lua:1148


def InFlowBC_StaticProfile(filename, gas, match="xyz-to-xyz", origin=(0.0, 0.0, 0.0)):
    """bc.lua:996-1019: each ghost cell takes the flow state of the nearest
    point of a profile file; the matching needs the block's geometry, so the
    states are made when the condition is set on a block."""
    from .flowfiles import FlowProfile, profile_ghost_states
    prof = FlowProfile(filename, gas, match)
    return BoundaryCondition(BCType.inflow_profile,
                             lambda blk, face: profile_ghost_states(prof, blk, int(face), origin))


__all__ = ["BoundaryCondition", "WallBC_WithSlip0", "WallBC_WithSlip", "WallBC_WithSlip1", "InFlowBC_Supersonic",
           "OutFlowBC_SimpleExtrapolate", "OutFlowBC_SimpleFlux", "OutFlowBC_Simple", "InFlowBC_StaticProfile"]
