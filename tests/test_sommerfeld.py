"""GPU: the radiative (Sommerfeld) outer boundary through the C ABI, against the numpy
restatement (oracle/sommerfeld.py, itself pinned in test_sommerfeld_oracle.py).

Bar: the boundary right-hand side 1e-12 normwise on the ghost points (the tangential stencils
round differently: integer numerators times one scale vs scaled coefficients) with interior
entries untouched; the Sommerfeld RK4 step 1e-12 normwise per field against the restatement's
step (the fused interior RHS against the oracle's, as everywhere else); a 48^3 puncture run of
40 steps stays finite.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
import sommerfeld as SF
from conftest import gpu_available, normwise

from bssn3d_hip import _lib, initial_data as idata
from bssn3d_hip.evolve import BSSNEvolution
from bssn3d_hip.rhs import stream_handle
from bssn3d_hip.state import PhysicsParams, scalars

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DEV = "cuda:0"


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("order", [4, 6, 8])
def test_boundary_rhs_matches_restatement(order):
    say(f"sommerfeld rhs order {order}")
    ng = 3 if order == 4 else 4
    S = (12, 15, 11) if order == 4 else (14, 13, 16)
    h = (0.11, 0.07, 0.13)
    xmin = (-0.5, 0.3, -0.9)
    rng = np.random.default_rng(order)
    ax = [xmin[a] + (np.arange(S[a]) - ng) * h[a] for a in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    F = np.empty((24,) + S)
    for f in range(24):
        k = rng.uniform(0.5, 2.0, size=3)
        F[f] = rng.uniform(-1, 1) + 0.2 * np.sin(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 6))
    Fd = torch.from_numpy(F).to(DEV)
    out = torch.full_like(Fd, 123.0)
    _lib.check(_lib.load().bssn_sommerfeld_rhs_f64(Fd.data_ptr(), out.data_ptr(), *S, order,
                                                   _lib.dbl(h), ng, _lib.dbl(xmin),
                                                   stream_handle(Fd.device)))
    got = out.cpu().numpy()
    I = np.meshgrid(*[np.arange(n) for n in S], indexing="ij")
    ghost = np.zeros(S, bool)
    for a in range(3):
        ghost |= (I[a] < ng) | (I[a] >= S[a] - ng)
    want = np.stack([SF.apply(np.full(S, 123.0), F[f], ng, xmin, h, order) for f in range(24)])
    e = normwise(np.where(ghost, got, 0.0), np.where(ghost, want, 0.0))
    assert e.max() <= 1e-12, e.max()
    assert np.all(got[:, ~ghost] == 123.0)


def test_sommerfeld_step_matches_restatement():
    say("sommerfeld rk4 step vs restatement, puncture 20^3")
    grid, st, _ = idata.config_single_puncture(seed=5, N=20)
    F = st.numpy()
    p = PhysicsParams()
    dt = 0.25 * grid.dx
    evo = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="sommerfeld")
    got = evo.evolve(torch.from_numpy(F).to(DEV), dt, 2, t0=0.2).data.cpu().numpy()
    S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, 0.2)
    xmin = (grid.xmin, grid.ymin, grid.zmin)
    want = F
    for k in range(2):
        want = SF.rk4_step(want, S16, 0.2 + k * dt, dt, 3, xmin, order=4, ko_sigma=0.1)
    e = normwise(got, want)
    assert e.max() <= 1e-12, e.max()


def test_sommerfeld_run_stays_finite():
    say("sommerfeld 40 steps, puncture 48^3")
    grid, st, _ = idata.config_single_puncture(seed=0, N=48, device=DEV)
    evo = BSSNEvolution(grid, PhysicsParams(), order=4, ko_sigma=0.1, bc="sommerfeld")
    out = evo.evolve(st, 0.25 * grid.dx, 40).data
    assert bool(torch.isfinite(out).all())
