"""The boundary conditions of the reference's input scripts (bc.lua), under
their names, for SFluidBlock.set_bc: each returns the device boundary kind and
its data.  Connections (ExchangeBC_FullFace) are made with
SFluidBlock.connect / connect_blocks.

    blk.set_bc(Face.west, InFlowBC_Supersonic(gas.flowstate(p, T, vel)))
    blk.set_bc(Face.east, OutFlowBC_Simple())
    blk.set_bc(Face.north, WallBC_WithSlip1())
    blk.set_bc(Face.west, InFlowBC_StaticProfile("profile.dat", gas, match="xyz-to-xyz"))
    blk.set_bc(Face.south, UserDefinedBC(lambda t, x, y, z: {"p": p(x, y), "T": T(x, y)}))
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


def UserDefinedGhostCellBC(ghost_cell, t=0.0, origin=(0.0, 0.0, 0.0)):
    """bc.lua:1286-1311 (UserDefinedBC): the ghost cells take the flow state a
    user function gives at their centres (user_defined_effects.d:179-226:
    p, T and massf, then update_thermo_from_pT and update_sound_speed, then
    velx, vely, velz, each 0 when absent).  ghost_cell(t, x, y, z) returns a
    dict of those names; it is evaluated once, at time t, when the condition
    is set on a block (the reference calls its Lua function every stage)."""
    def states(blk, face):
        from .flowfiles import ghost_centres
        g = ghost_centres(blk.dims, blk.ni, blk.nj, blk.nk, blk.dx, blk.dy, blk.dz, int(face), origin)
        nsp = getattr(blk.gas, "n_species", 1)
        out = np.empty(g.shape[:2] + (9 + nsp,))
        for layer in range(g.shape[0]):
            for n in range(g.shape[1]):
                x, y, z = g[layer, n]
                d = ghost_cell(t, x, y, z)
                vel = (d.get("velx", 0.0), d.get("vely", 0.0), d.get("velz", 0.0))
                if nsp > 1:
                    Y = d.get("massf")
                    Y = np.zeros(nsp) if Y is None else np.asarray(Y, dtype=np.float64)
                    fs = blk.gas.flowstate(d["p"], d["T"], vel, Y)
                else:
                    fs = blk.gas.flowstate(d["p"], d["T"], vel)[:8]
                    fs = np.concatenate([fs, [1.0]])
                out[layer, n] = np.concatenate([fs[:8], [0.0], fs[8:]])
        return out
    return BoundaryCondition(BCType.inflow_profile, states)


UserDefinedBC = UserDefinedGhostCellBC      # bc.lua:1311


__all__ = ["BoundaryCondition", "WallBC_WithSlip0", "WallBC_WithSlip", "WallBC_WithSlip1", "InFlowBC_Supersonic",
           "OutFlowBC_SimpleExtrapolate", "OutFlowBC_SimpleFlux", "OutFlowBC_Simple", "InFlowBC_StaticProfile",
           "UserDefinedGhostCellBC", "UserDefinedBC"]
