"""Long-run constraint series on the oracle.

TEST INFRASTRUCTURE ONLY: the device long-run series (bssn3d_hip.series) are checked against
this. Nothing in the product imports it.

Restates the sampling loop of 3D/src/bssn3d/longrun_stability.py:124-188 on the C oracle: the
state is advanced in chunks by oracle.evolve (periodic: every RHS reads a synced copy of its
stage input and the RK combinations run over the whole padded arrays, so the ghosts carry the
reference's values) and after each chunk the interior RMS of H and |M| (oracle.constraint_l2,
which reads those ghosts), max|alpha|, max|U - Minkowski| and finiteness are recorded. A run
stops early when a sample is non-finite or max|alpha| > max_alpha_stop.
"""
from __future__ import annotations

import numpy as np

import oracle as O

ALPHA, CHI = 0, 1
_FLAT = (0, 1, 3, 6, 8)            # alpha, chi, gt_xx, gt_yy, gt_zz: 1 on flat space


def minkowski(shape) -> np.ndarray:
    U = np.zeros((24,) + tuple(shape))
    for f in _FLAT:
        U[f] = 1.0
    return U


def hambump(shape, ng: int, h: float, amp: float, lo: float = -0.5) -> np.ndarray:
    """Flat space with chi + amp cos(2 pi x) cos(2 pi y) cos(2 pi z)
    (3D/src/bssn3d/initial_data.py:33-55); x_i = lo + (i - ng) h."""
    x = lo + (np.arange(shape[0]) - ng) * h
    c = np.cos(2.0 * np.pi * x)
    U = minkowski(shape)
    U[CHI] += amp * c[:, None, None] * c[None, :, None] * c[None, None, :]
    return U


def series(U0: np.ndarray, S16, dt: float, h: float, ng: int, total: int, chunk: int,
           samples: int, ko_sigma: float, order: int = 6, max_alpha_stop: float = 10.0,
           integrator: str = "rk4", chi_floor: float = 1e-4, alpha_floor: float = 1e-4):
    """Rows {ncross, steps, H, M, max_alpha, max_dev, finite} of one periodic long run."""
    mink = minkowski(U0.shape[1:])
    h3 = (h, h, h)

    def row(U, t):
        H, M = O.constraint_l2(U, h3, order, ng)
        return dict(ncross=float(t), steps=round(t / dt), H=H, M=M,
                    max_alpha=float(np.max(np.abs(U[ALPHA]))),
                    max_dev=float(np.max(np.abs(U - mink))),
                    finite=bool(np.all(np.isfinite(U))))

    U = np.array(U0, dtype=np.float64)
    t = 0.0
    rows = [row(U, t)]
    for _ in range(samples):
        U = O.evolve(U, S16, dt, chunk, t0=t, integrator=integrator, order=order,
                     ko_order=0, ko_sigma=ko_sigma, ng=ng,
                     periodic=True, chi_floor=chi_floor, alpha_floor=alpha_floor)
        t += chunk * dt
        rows.append(row(U, t))
        if not rows[-1]["finite"] or rows[-1]["max_alpha"] > max_alpha_stop:
            break
    return rows, U
