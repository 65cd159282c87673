"""flowpanel_amd -- MI355X-native (gfx950) hot path of the FLOWPanel.jl panel method.

Host-side mirror of the reference's API over the C ABI of
include/flowpanel_hip.h (libflowpanel_hip.so, built by ``build.py``):

    grid, shedding = geometry.loft_wing(...)
    body = RigidWakeBody(grid, shedding)
    solve(body, Uinfs, Das, Dbs, solver=solve_ludiv)
    Us = calcfield_U(body, body)
    Cps = calcfield_Cp(body, magUinf)
"""
from . import geometry, meshio, vtk
from .vtk import save
from .geometry import Grid, calc_shedding, direction
from .meshio import read_msh, read_msh_nodes, read_vtu
from .bodies import (ConstantDoublet, ConstantSource, MultiBody, NonLiftingBody, RigidWakeBody, VortexRing,
                     calc_areas, calc_controlpoints, calc_normals, generate_revolution_liftbody)
from .solvers import (assemble, calc_bc_noflowthrough, getrf, solve, solve_backslash, solve_gmres,
                      solve_gmres_ir, solve_lu32_ir, solve_ludiv, solve_freestreams, BodyBatch, calc_elprescribe,
                      last_stats)
from .postprocess import (Uind, phi, calcfield_U, calcfield_Ugradmu_cell, calcfield_Ugradmu_node, calcfield_Ugradmu,
                          addfields, calcfield_Cp, calcfield_F, calcfield_LDS, calcfield_Ftot, calcfield_Mtot,
                          calcfield_lmn, calcfield_sectionalforce, loads)
from .adjoint import lift_drag, lift_drag_gradient, set_nodes
from ._native import FPError, context

__all__ = [
    "geometry", "meshio", "vtk", "save", "Grid", "calc_shedding", "read_msh", "read_msh_nodes", "read_vtu", "direction", "ConstantSource", "ConstantDoublet", "VortexRing", "NonLiftingBody",
    "RigidWakeBody", "MultiBody", "generate_revolution_liftbody", "calc_normals", "calc_controlpoints", "calc_areas", "assemble",
    "calc_bc_noflowthrough", "getrf", "solve", "solve_ludiv", "solve_backslash", "solve_gmres", "solve_lu32_ir",
    "solve_gmres_ir", "last_stats", "Uind", "phi", "calcfield_U", "calcfield_Ugradmu_cell", "calcfield_Ugradmu_node", "calcfield_Ugradmu", "addfields", "calcfield_Cp", "calcfield_F", "calcfield_LDS",
    "calcfield_Ftot", "calcfield_Mtot", "calcfield_lmn", "calcfield_sectionalforce", "loads",
    "FPError", "context", "solve_freestreams", "BodyBatch", "calc_elprescribe", "lift_drag", "lift_drag_gradient", "set_nodes",
]
