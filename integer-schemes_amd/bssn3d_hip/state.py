"""BSSN state layout and physics parameters (3D/src/bssn3d/state.py:26-153).

A state is a float64 array of shape (24, Sx, Sy, Sz), C order, fields in the order below
(the order the reference's Dendro RHS emits). On the device it is a torch tensor (or anything
exposing __dlpack__); the library reads and writes it in place through its data pointer.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple

ALPHA, CHI, K = 0, 1, 2
GT0, GT1, GT2, GT3, GT4, GT5 = 3, 4, 5, 6, 7, 8
BETA0, BETA1, BETA2 = 9, 10, 11
AT0, AT1, AT2, AT3, AT4, AT5 = 12, 13, 14, 15, 16, 17
GTILDE0, GTILDE1, GTILDE2 = 18, 19, 20
B0, B1, B2 = 21, 22, 23
NUM_VARS = 24

VAR_NAMES = [
    "alpha", "chi", "K",
    "gt0", "gt1", "gt2", "gt3", "gt4", "gt5",
    "beta0", "beta1", "beta2",
    "At0", "At1", "At2", "At3", "At4", "At5",
    "Gt0", "Gt1", "Gt2",
    "B0", "B1", "B2",
]
NAME_TO_IDX = {n: i for i, n in enumerate(VAR_NAMES)}

# upper-triangular, row-major symmetric index (state.py:52-55)
SYM_IDX = {
    (0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5,
    (1, 0): 1, (2, 0): 2, (2, 1): 4,
}


@dataclass(frozen=True)
class PhysicsParams:
    """Gauge and damping parameters, same names and defaults as the reference
    (3D/src/bssn3d/state.py:113-153)."""

    eta: float = 2.0
    lmbda: Tuple[float, float, float, float] = (1.0, 1.0, 1.0, 1.0)
    lambda_f: Tuple[float, float] = (1.0, 0.0)
    cahd_c: float = 0.06
    ssl_h: float = 0.6
    ssl_sigma: float = 20.0
    chi_floor: float = 1.0e-4
    alpha_floor: float = 1.0e-4
    alpha_inf: float = 1.0
    chi_inf: float = 1.0

    def lam(self, i: int) -> float:
        return self.lmbda[i]


class BSSNState:
    """Thin holder of a (24, Sx, Sy, Sz) float64 tensor, with the reference's accessors."""

    def __init__(self, data):
        self.data = data

    alpha = property(lambda s: s.data[ALPHA])
    chi = property(lambda s: s.data[CHI])
    K = property(lambda s: s.data[K])
    beta = property(lambda s: s.data[BETA0:BETA0 + 3])
    Gt = property(lambda s: s.data[GTILDE0:GTILDE0 + 3])
    B = property(lambda s: s.data[B0:B0 + 3])
    gt = property(lambda s: s.data[GT0:GT0 + 6])
    At = property(lambda s: s.data[AT0:AT0 + 6])

    def gt_ij(self, i, j):
        return self.data[GT0 + SYM_IDX[(i, j)]]

    def At_ij(self, i, j):
        return self.data[AT0 + SYM_IDX[(i, j)]]

    @property
    def shape(self):
        return tuple(self.data.shape[1:])


def pack_S(eta, cahd_c, dt, dx, ssl_h, ssl_sigma, t, lmbda, lambda_f, dy, dz):
    """The 16-double scalar pack in the reference's order
    (3D/src/bssn3d/fused_rhs_cuda.py:52-55)."""
    return [float(eta), float(cahd_c), float(dt), float(dx), float(ssl_h), float(ssl_sigma),
            float(t), float(lmbda[0]), float(lmbda[1]), float(lmbda[2]), float(lmbda[3]),
            float(lambda_f[0]), float(lambda_f[1]), float(dx), float(dy), float(dz)]


def scalars(p: PhysicsParams, dx, dy, dz, dt, t):
    return pack_S(p.eta, p.cahd_c, dt, dx, p.ssl_h, p.ssl_sigma, t, p.lmbda, p.lambda_f, dy, dz)
