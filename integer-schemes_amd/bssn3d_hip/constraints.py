"""Hamiltonian and momentum constraints on the device (3D/src/bssn3d/constraints.py:38-62)."""
from __future__ import annotations

from typing import Dict, Tuple

import torch

from . import _lib
from .grid import Grid, ghost_width
from .rhs import as_tensor, stream_handle


class ConstraintSolver:
    def __init__(self, grid: Grid, order: int = 6):
        if order not in (4, 6, 8):
            raise ValueError(f"order {order} not supported on the device (4, 6 or 8)")
        ng = ghost_width(order)
        if ng > grid.ng:
            raise ValueError(f"order {order} needs ng>={ng}, grid ng={grid.ng}")
        self.grid = grid
        self.order = order

    def _h3(self):
        g = self.grid
        return _lib.dbl([g.dx, g.dy, g.dz])

    def constraints(self, state) -> Dict[str, torch.Tensor]:
        """{'H','M0','M1','M2'} over the full padded grid."""
        F = as_tensor(state)
        _, Sx, Sy, Sz = F.shape
        HM = torch.empty((4, Sx, Sy, Sz), dtype=torch.float64, device=F.device)
        _lib.check(_lib.load().bssn_constraints_f64(F.data_ptr(), HM.data_ptr(), Sx, Sy, Sz,
                                                    self.order, self._h3(),
                                                    stream_handle(F.device)))
        return {"H": HM[0], "M0": HM[1], "M1": HM[2], "M2": HM[3]}

    def l2(self, state) -> Tuple[float, float]:
        """Interior RMS (H, |M|), |M| = sqrt(mean(M0^2 + M1^2 + M2^2))."""
        F = as_tensor(state)
        _, Sx, Sy, Sz = F.shape
        lib = _lib.load()
        HM = torch.empty((4, Sx, Sy, Sz), dtype=torch.float64, device=F.device)
        npart = lib.bssn_l2_partial_doubles(Sx, Sy, Sz, self.grid.ng)
        part = torch.empty(max(npart, 1), dtype=torch.float64, device=F.device)
        out = _lib.dbl([0.0, 0.0])
        _lib.check(lib.bssn_constraint_l2_f64(F.data_ptr(), HM.data_ptr(), part.data_ptr(), Sx,
                                              Sy, Sz, self.order, self._h3(), self.grid.ng, out,
                                              stream_handle(F.device)))
        return float(out[0]), float(out[1])
