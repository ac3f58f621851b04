"""Vertex-centred, ghost-padded grid (3D/src/bssn3d/grid.py:14-50).

x_i = xmin + (i - ng) h for i in [0, N + 2 ng). Coordinates are numpy arrays; the padded
extents (Sx, Sy, Sz) = N + 2 ng are what the device kernels take.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def ghost_width(order: int, ko_order: int | None = None) -> int:
    """Widest stencil reach, max(order, ko_order) // 2 (common/src/mcs_common/derivatives.py:72)."""
    if order not in (4, 6, 8):
        raise ValueError(f"Unsupported finite difference order. Expected 4, 6, or 8, but got {order}.")
    ko_order = order if ko_order is None else ko_order
    if ko_order not in (4, 6, 8):
        raise ValueError(f"Unsupported KO order. Expected 4, 6, or 8, but got {ko_order}.")
    return max(order, ko_order) // 2


def default_ko_order(order: int) -> int:
    """The Kreiss-Oliger order the device pairs with FD order ``order``: order + 2 (the
    reference's production rule, common/src/mcs_common/derivatives.py:55-58), at most 8, its
    widest KO stencil (derivatives.py:45-49): 4 -> 6, 6 -> 8, 8 -> 8."""
    return min(order + 2, 8)


@dataclass
class Grid:
    Nx: int
    Ny: int
    Nz: int
    dx: float
    dy: float
    dz: float
    ng: int
    xmin: float = -0.5
    ymin: float = -0.5
    zmin: float = -0.5

    def __post_init__(self):
        ng = self.ng
        self.x = self.xmin + np.arange(-ng, self.Nx + ng) * self.dx
        self.y = self.ymin + np.arange(-ng, self.Ny + ng) * self.dy
        self.z = self.zmin + np.arange(-ng, self.Nz + ng) * self.dz

    @property
    def shape(self):
        return (self.Nx + 2 * self.ng, self.Ny + 2 * self.ng, self.Nz + 2 * self.ng)

    @property
    def interior(self):
        ng = self.ng
        return (slice(ng, -ng or None),) * 3

    def mesh(self):
        return np.meshgrid(self.x, self.y, self.z, indexing="ij")

    @classmethod
    def from_domain(cls, N, order=6, lo=-0.5, hi=0.5, ko_order=8):
        if isinstance(N, int):
            N = (N, N, N)
        ng = ghost_width(order, ko_order)
        d = [(hi - lo) / n for n in N]
        return cls(Nx=N[0], Ny=N[1], Nz=N[2], dx=d[0], dy=d[1], dz=d[2], ng=ng,
                   xmin=lo, ymin=lo, zmin=lo)
