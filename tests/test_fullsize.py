"""GPU: the benchmark configurations at their full sizes, against the oracle.

256^3 binary, one RK4 step (the metric's workload): the oracle cannot step 3.45 GB states in a
test's budget, so it steps periodic windows. A window is the cube of interior points within
m = 12 of a check region, taken with periodic wrap from the state before the step; the oracle
advances it with edge-padded stencils (its bc != "periodic" step). The stencil reach of one
stage is 3 (KO6) and an RK4 step has 4 stages, so the window edge cannot reach the check region:
there the oracle's step is the periodic step exactly. Windows: both punctures, random interior
points and the periodic corner. Bar (as for the full-size RHS): per field, the HIP increment's
error against the long-double step, normalised by the field's full-grid maximum increment, is at
most max(1e-12, 2 x) the FP64 oracle's own error there.

INT8-split form at 256^3 (the single-puncture construction of the 48^3 workload at 256^3):
the bundle from 7 int8 digits and the RHS on it at sampled points (random, the puncture's 7^3
neighbourhood, edges and corners) against the oracle's bundle and RHS, per derivative and per
field, normalised by full-grid maxima. Bars: RHS <= 1e-10 per field (the precision form's stated
bound); each derivative within 1e-9 of the long-double bundle (the form's 2^-49 digit truncation
amplified by the stencil's cancellation on O(1) fields).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from conftest import gpu_available, normwise

from bssn3d_hip import initial_data as idata
from bssn3d_hip.evolve import BSSNEvolution
from bssn3d_hip.rhs import rhs_int8_into, rhs_into, stream_handle
from bssn3d_hip import _lib
from bssn3d_hip.state import PhysicsParams, VAR_NAMES, scalars

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DEV = "cuda:0"


def say(msg):
    print(f"[case] {msg}", flush=True)


def test_binary_256_rk4_step_windows():
    say("256^3 binary: one RK4 step (periodic, KO 0.1) vs the oracle's step in 8 windows")
    grid, U, pos = idata.config_binary(seed=0, N=256, device=DEV)
    p = PhysicsParams()
    dt = 0.25 * grid.dx
    t0 = 0.0
    ng, N = grid.ng, 256
    F0 = U.cpu().numpy()
    U0 = U.clone()
    evo = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="periodic", integrator="rk4")
    evo.step_(U, t0, dt)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(U).all())
    fullmax = (U - U0).abs().reshape(24, -1).amax(dim=1).cpu().numpy()
    del U0
    F1 = U.cpu().numpy()
    del U
    torch.cuda.empty_cache()
    S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, t0)
    c, m = 6, 12
    W = c + 2 * m
    rng = np.random.default_rng(5)
    centres = []
    for q in pos:       # puncture positions -> interior indices
        centres.append([int(round((q[a] - lo) / grid.dx)) for a, lo in
                        enumerate((grid.xmin, grid.ymin, grid.zmin))])
    centres += [list(rng.integers(0, N, size=3)) for _ in range(5)]
    centres.append([0, 0, 0])          # the periodic corner: the window wraps on every axis
    e_hip_all, e_ref_all = [], []
    for ci in centres:
        ax = [(ci[a] - c // 2 - m + np.arange(W)) % N + ng for a in range(3)]
        win = F0[np.ix_(np.arange(24), ax[0], ax[1], ax[2])]
        ref = O.rk4_step(win, S16, t0, dt, order=4, ko_sigma=0.1, ng=ng, periodic=False,
                         chi_floor=p.chi_floor, alpha_floor=p.alpha_floor)
        tru = O.rk4_step_ld(win, S16, t0, dt, order=4, ko_sigma=0.1, ng=ng, periodic=False,
                            chi_floor=p.chi_floor, alpha_floor=p.alpha_floor)
        sl = slice(m, m + c)
        core = [ax[a][sl] for a in range(3)]
        got = F1[np.ix_(np.arange(24), core[0], core[1], core[2])]
        base = win[:, sl, sl, sl]
        inc_t = (tru[:, sl, sl, sl] - base.astype(np.longdouble)).astype(np.float64)
        e_hip_all.append(normwise(got - base, inc_t, den=fullmax))
        e_ref_all.append(normwise(ref[:, sl, sl, sl] - base, inc_t, den=fullmax))
    e_hip = np.max(e_hip_all, axis=0)
    e_ref = np.max(e_ref_all, axis=0)
    print("e_ref", np.array2string(e_ref, precision=2), "\ne_hip",
          np.array2string(e_hip, precision=2), flush=True)
    assert np.all(e_hip <= np.maximum(1e-12, 2.0 * e_ref)), (e_hip, e_ref)


def test_int8_split_256_sampled():
    say("256^3 single puncture: INT8-split bundle and RHS vs the oracle, per derivative/field")
    grid, U, pos = idata.config_single_puncture(seed=0, N=256, device=DEV)
    p = PhysicsParams()
    S16 = scalars(p, grid.dx, grid.dy, grid.dz, 0.25 * grid.dx, 0.0)
    h3 = [grid.dx, grid.dy, grid.dz]
    Sx, Sy, Sz = grid.shape
    lib = _lib.load()
    sh = stream_handle(U.device)
    D8 = torch.empty((138, Sx, Sy, Sz), dtype=torch.float64, device=DEV)
    out8 = torch.empty_like(U)
    rhs_int8_into(U, out8, D8, S16, 4)      # leaves the INT8 bundle in D8
    out64 = torch.empty_like(U)
    rhs_into(U, out64, S16, 4)
    torch.cuda.synchronize()
    rhsmax = out64.abs().reshape(24, -1).amax(dim=1).cpu().numpy()
    D64 = torch.empty_like(D8)
    _lib.check(lib.bssn_deriv_bundle_f64(U.data_ptr(), D64.data_ptr(), Sx, Sy, Sz, 4,
                                         _lib.dbl(h3), sh))
    torch.cuda.synchronize()
    dmax = D64.abs().reshape(138, -1).amax(dim=1).cpu().numpy()
    del D64
    rng = np.random.default_rng(2)
    pts = [np.stack([rng.integers(0, n, size=2000) for n in (Sx, Sy, Sz)])]
    pts.append(np.array([[x, y, z] for x in (0, 1, Sx - 2, Sx - 1) for y in (0, 2, Sy - 1)
                         for z in (0, 3, Sz - 1)]).T)
    for q in pos:
        cc = [int(round((q[a] - lo) / grid.dx)) + grid.ng
              for a, lo in enumerate((grid.xmin, grid.ymin, grid.zmin))]
        r = np.arange(-3, 4)
        X, Y, Z = np.meshgrid(cc[0] + r, cc[1] + r, cc[2] + r, indexing="ij")
        pts.append(np.stack([X.ravel(), Y.ravel(), Z.ravel()]))
    P = np.concatenate(pts, axis=1)
    idx = (P[0] * Sy + P[1]) * Sz + P[2]
    it = torch.from_numpy(idx).to(DEV)
    d8 = D8.reshape(138, -1)[:, it].cpu().numpy()
    r8 = out8.reshape(24, -1)[:, it].cpu().numpy()
    del D8, out8, out64
    F = U.cpu().numpy()
    del U
    torch.cuda.empty_cache()
    d_ref = O.deriv_points(F, h3, idx, 4)
    d_tru = O.deriv_points(F, h3, idx, 4, long_double=True)
    r_tru, _ = O.rhs_points_ld(F, S16, idx, 4)
    r_ref = O.rhs_points(F, S16, idx, 4)
    e_d8 = normwise(d8, d_tru, den=dmax)
    e_d64 = normwise(d_ref, d_tru, den=dmax)
    e_r8 = normwise(r8, r_tru, den=rhsmax)
    e_r64 = normwise(r_ref, r_tru, den=rhsmax)
    worst = np.argsort(e_d8)[::-1][:5]
    print("per field RHS error (int8 | fp64 oracle):")
    for f, n in enumerate(VAR_NAMES):
        print(f"  {n:6s} {e_r8[f]:.2e} | {e_r64[f]:.2e}")
    print("derivatives: int8 max", f"{e_d8.max():.2e}", "median", f"{np.median(e_d8):.2e}",
          "| fp64 max", f"{e_d64.max():.2e}", "worst", worst.tolist(), flush=True)
    assert np.all(e_r8 <= 1e-10), e_r8
    assert np.all(e_d8 <= 1e-9), e_d8
