"""Radiative (Sommerfeld) outer boundary for the BSSN state, restated in numpy.

TEST INFRASTRUCTURE ONLY: the GPU boundary kernel and the Sommerfeld step are checked against
this. Nothing in the product imports it.

The reference has no outer boundary condition for BSSN (its steps are periodic or leave edge-
padded ghosts). This follows its Maxwell-Chern-Simons solver's bc_sommerfeld
(3D/src/mcs3d/main.py:230-250), applied to each of the 24 fields: in the ghost zones the time
derivative is replaced by the outgoing-wave condition

    d_t u = -(x d_x u + y d_y u + z d_z u) / r

with the derivative normal to the face a one-sided first difference (forward on the low face,
backward on the high face) and the two tangential derivatives the centred order-`order`
operator with edge padding (common/src/mcs_common/derivatives.py:78-88). Faces are written in
the order x-low, x-high, y-low, y-high, z-low, z-high, so at edges and corners the later face
wins. Coordinates are vertex-centred: x_i = xmin + (i - ng) h (3D/src/bssn3d/grid.py:27-33).
"""
from __future__ import annotations

import numpy as np

import oracle as O


def _coords(shape, ng, xmin3, h3):
    ax = [xmin3[a] + (np.arange(shape[a]) - ng) * h3[a] for a in range(3)]
    return np.meshgrid(*ax, indexing="ij")


def bc_rhs(u: np.ndarray, ng: int, xmin3, h3, order: int = 4) -> dict:
    """The radiative right-hand side of one (Sx, Sy, Sz) field on each face's ghost slab:
    returns {(axis, side): array over the whole grid} (only the slab entries are meaningful)."""
    X, Y, Z = _coords(u.shape, ng, xmin3, h3)
    R = np.sqrt(X * X + Y * Y + Z * Z)
    R = np.where(R == 0.0, 1.0, R)
    cen = [O.apply_axis(u, a, 0, order, 1.0 / h3[a]) for a in range(3)]
    out = {}
    for a in range(3):
        fwd = (np.roll(u, -1, axis=a) - u) / h3[a]
        bwd = (u - np.roll(u, 1, axis=a)) / h3[a]
        for side, one in ((0, fwd), (1, bwd)):
            d = list(cen)
            d[a] = one
            out[(a, side)] = -(X * d[0] + Y * d[1] + Z * d[2]) / R
    return out


def apply(dtu: np.ndarray, u: np.ndarray, ng: int, xmin3, h3, order: int = 4) -> np.ndarray:
    """dtu with its ghost zones replaced by the radiative condition of u (one field)."""
    dtu = dtu.copy()
    rb = bc_rhs(u, ng, xmin3, h3, order)
    for a in range(3):
        for side in (0, 1):
            sl = [slice(None)] * 3
            sl[a] = slice(0, ng) if side == 0 else slice(u.shape[a] - ng, u.shape[a])
            dtu[tuple(sl)] = rb[(a, side)][tuple(sl)]
    return dtu


def stage_rhs(F: np.ndarray, S16, ng: int, xmin3, order: int = 4, ko_sigma: float = 0.0):
    """f(F): the edge-padded RHS + KO at every padded point, ghosts replaced by the radiative
    condition (all 24 fields)."""
    out = O.rhs(F, S16, order=order, ko_sigma=ko_sigma)   # KO order: the default pairing
    h3 = (S16[13], S16[14], S16[15])
    for f in range(F.shape[0]):
        out[f] = apply(out[f], F[f], ng, xmin3, h3, order)
    return out


def rk4_step(F: np.ndarray, S16, t: float, dt: float, ng: int, xmin3, order: int = 4,
             ko_sigma: float = 0.0, chi_floor: float = 1e-4, alpha_floor: float = 1e-4):
    """Classic RK4 with the radiative ghosts (stage times t, t+dt/2, t+dt/2, t+dt; CAHD on the
    full dt, 3D/src/bssn3d/evolve.py:121-132), then enforce over all points."""
    S = np.array(S16, dtype=np.float64)

    def f(Y, tt):
        s = S.copy()
        s[2] = dt
        s[6] = tt
        return stage_rhs(Y, s, ng, xmin3, order, ko_sigma)
    k1 = f(F, t)
    k2 = f(F + 0.5 * dt * k1, t + 0.5 * dt)
    k3 = f(F + 0.5 * dt * k2, t + 0.5 * dt)
    k4 = f(F + dt * k3, t + dt)
    out = F + (dt / 6.0) * (((k1 + 2.0 * k2) + 2.0 * k3) + k4)
    return O.enforce(out, chi_floor, alpha_floor)
