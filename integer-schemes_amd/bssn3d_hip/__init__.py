"""bssn3d_hip: the BSSN right-hand side and RK4/MSRK step of bssn3d on MI355X (gfx950).

Host-side mirror of the reference's bssn3d interface over the C ABI of libbssn_hip.so
(include/bssn_hip.h). Device buffers are torch tensors (or any __dlpack__ exporter).
"""
from .grid import Grid, ghost_width
from .state import (BSSNState, NAME_TO_IDX, NUM_VARS, PhysicsParams, SYM_IDX, VAR_NAMES,
                    pack_S, scalars)
from .rhs import BSSNSolver, cuda_fused_algebra, device_rhs_fused
from .evolve import BSSNEvolution
from .constraints import ConstraintSolver
from . import initial_data

__all__ = ["Grid", "ghost_width", "BSSNState", "NAME_TO_IDX", "NUM_VARS", "PhysicsParams",
           "SYM_IDX", "VAR_NAMES", "pack_S", "scalars", "BSSNSolver", "cuda_fused_algebra",
           "device_rhs_fused", "BSSNEvolution", "ConstraintSolver", "initial_data"]
