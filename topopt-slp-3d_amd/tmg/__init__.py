"""tmg -- MI355X-native multigrid equilibrium solver for 3-D SIMP topology
optimisation (the GMG-preconditioned CG of TopOpt-SLP-3D) on hand-written
gfx950 HIP kernels behind the C-ABI of include/tmg.h."""
from . import problem
from ._lib import TMG_FP64, TMG_MIXED, TmgError, lib
from .gmg import (ApplyFilter, ElemNeighborhood, Filter, FMVCycle, GlobalStiffMatrix, Grad, RefreshGMG, SetupGMG,
                  Smooth, SolveMG, Solver, VCycle, WCycle)

__all__ = ["problem", "lib", "TmgError", "TMG_FP64", "TMG_MIXED", "Solver", "Filter", "GlobalStiffMatrix",
           "SetupGMG", "RefreshGMG", "SolveMG", "VCycle", "WCycle", "FMVCycle", "Smooth", "ElemNeighborhood",
           "ApplyFilter", "Grad"]
