"""GPU: the INT8-split derivative bundle (Ozaki scheme I on INT8 MFMA) through the C ABI.

Bars:
  * bit-identical to the numpy restatement (oracle/int8_split.py), whose contractions are exact
    int64 products: any INT32 MFMA result that differs from the exact integer product, a wrong
    operand lane layout, or a wrong digit would change bits;
  * against the FP64 oracle bundle: per-derivative normwise error <= 5e-11 with 7 digits (the
    FP64 bundle's own cancellation floor on g~ second derivatives is ~1e-11), and the RHS built
    on it (FP64 algebra) <= 1e-12 per field.
Cases: the 16^3 single-puncture grid, ragged extents with partial and multiple 28 x 28 tiles,
fewer digits (error grows ~128x per digit dropped), zero and constant fields.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import int8_split as I8
import oracle as O
from conftest import gpu_available, normwise

from bssn3d_hip import _lib, initial_data as idata
from bssn3d_hip.rhs import BSSNSolver, rhs_int8_into, stream_handle
from bssn3d_hip.state import PhysicsParams, scalars

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DEV = "cuda:0"


def say(msg):
    print(f"[case] {msg}", flush=True)


def hip_bundle(F: np.ndarray, h3, k: int) -> np.ndarray:
    _, Sx, Sy, Sz = F.shape
    Fd = torch.from_numpy(np.ascontiguousarray(F)).to(DEV)
    D = torch.full((138, Sx, Sy, Sz), np.nan, dtype=torch.float64, device=DEV)
    _lib.check(_lib.load().bssn_deriv_bundle_i8_f64(Fd.data_ptr(), D.data_ptr(), Sx, Sy, Sz, 4,
                                                    _lib.dbl(h3), k, stream_handle(Fd.device)))
    torch.cuda.synchronize()
    return D.cpu().numpy()


def ragged_state(shape, seed):
    rng = np.random.default_rng(seed)
    axes = [np.linspace(0, 1, n) for n in shape]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    F = np.empty((24,) + tuple(shape))
    for f in range(24):
        k = rng.uniform(0.5, 3.0, size=3)
        F[f] = rng.uniform(-2, 2) + 0.1 * np.sin(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z)
                                                 + rng.uniform(0, 6))
    return F


@pytest.mark.parametrize("k", [7, 3])
def test_int8_bundle_bitwise_puncture(k):
    say(f"int8 bundle puncture 16^3 k={k}")
    grid, st, _ = idata.config_single_puncture(seed=0, N=16)
    F = st.numpy()
    h = (grid.dx, grid.dy, grid.dz)
    got = hip_bundle(F, h, k)
    want = I8.bundle(F, h, k)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]


@pytest.mark.parametrize("shape", [(31, 9, 37), (5, 58, 6)])
def test_int8_bundle_bitwise_ragged(shape):
    say(f"int8 bundle ragged {shape}")
    F = ragged_state(shape, seed=sum(shape))
    F[5] = 0.0                # an all-zero field (scale 1, digits 0)
    F[9] = 1.25               # a constant field (offset takes all of it)
    h = (0.07, 0.05, 0.11)
    got = hip_bundle(F, h, 7)
    want = I8.bundle(F, h, 7)
    assert np.array_equal(got, want), np.argwhere(got != want)[:5]
    assert np.all(got[3 * 5:3 * 5 + 3] == 0.0) and np.all(got[3 * 9:3 * 9 + 3] == 0.0)


def test_int8_error_vs_fp64_oracle():
    say("int8 vs fp64 oracle (bundle and RHS)")
    grid, st, _ = idata.config_single_puncture(seed=1, N=16)
    F = st.numpy()
    h = (grid.dx, grid.dy, grid.dz)
    D8 = hip_bundle(F, h, 7)
    D = O.deriv_bundle(F, h, order=4)
    e = normwise(D8, D)
    print("bundle normwise: max %.2e median %.2e" % (e.max(), np.median(e)))
    assert e.max() <= 5e-11
    # the RHS on the INT8 bundle (FP64 algebra) against the FP64 oracle RHS
    S16 = scalars(PhysicsParams(), grid.dx, grid.dy, grid.dz, 0.25 * grid.dx, 0.0)
    Fd = torch.from_numpy(F).to(DEV)
    out = torch.empty_like(Fd)
    Dd = torch.empty((138,) + F.shape[1:], dtype=torch.float64, device=DEV)
    rhs_int8_into(Fd, out, Dd, S16, 4)
    want = O.rhs(F, S16, order=4)
    er = normwise(out.cpu().numpy(), want)
    print("rhs normwise per field:", " ".join("%.1e" % v for v in er))
    assert er.max() <= 1e-12
    # fewer digits: the error grows about 2^7 per digit dropped
    e5 = normwise(hip_bundle(F, h, 5), D).max()
    e4 = normwise(hip_bundle(F, h, 4), D).max()
    assert e4 > 30.0 * e5 > 0.0


def test_int8_solver_scheme():
    say("BSSNSolver int8 scheme")
    grid, st, _ = idata.config_single_puncture(seed=2, N=16, device=DEV)
    s8 = BSSNSolver(grid, order=4, scheme="int8")
    s64 = BSSNSolver(grid, order=4, scheme="fused")
    a = s8.rhs_tensor(st, t=0.3).cpu().numpy()
    b = s64.rhs_tensor(st, t=0.3).cpu().numpy()
    assert normwise(a, b).max() <= 1e-12
    with pytest.raises(ValueError):
        BSSNSolver(grid, order=6, scheme="int8")


@pytest.mark.parametrize("scheme,tol", [("twopass", 1e-12), ("int8", 1e-11)])
def test_composed_step_vs_fused(scheme, tol):
    """The reference-composed RK4 step (sync + bundle + algebra + KO per stage, full-array
    updates, enforce, sync) with the FP64 or the INT8-split bundle, against the fused step:
    3 steps of the 20^3 puncture, periodic."""
    say(f"composed {scheme} RK4 step vs fused")
    grid, st, _ = idata.config_single_puncture(seed=3, N=20, device=DEV)
    from bssn3d_hip.evolve import BSSNEvolution
    dt = 0.25 * grid.dx
    p = PhysicsParams()
    a = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="periodic", scheme=scheme)
    b = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="periodic", scheme="fused")
    got = a.evolve(st, dt, 3, t0=0.1).data.cpu().numpy()
    want = b.evolve(st, dt, 3, t0=0.1).data.cpu().numpy()
    e = normwise(got, want)
    print(f"{scheme}: max {e.max():.2e}")
    assert e.max() <= tol


@pytest.mark.parametrize("k", [7, 4])
def test_batched_int8_bundle_equals_single_grid(k):
    """bssn_batch_deriv_bundle_i8_f64 over blocks of uneven extents (3..40 per axis, several
    tiles and partial tiles) and three spacings: every block's 138 arrays equal
    bssn_deriv_bundle_i8_f64 on that block alone, bit for bit."""
    say(f"batched INT8 bundle vs single grid, {k} digits")
    from bssn3d_hip.batch import Block, BlockBatch, fill_punctures
    rng = np.random.default_rng(17 + k)
    blocks = []
    for _ in range(9):
        ext = tuple(int(v) for v in rng.choice([3, 8, 12, 29, 40], size=3))
        h = float(rng.choice([0.25, 0.125, 0.0625]))
        lo = tuple(float(v) for v in rng.uniform(-2, 2, size=3))
        blocks.append(Block(ext, (h, h, h), lo))
    batch = BlockBatch(blocks, ng=3, device=DEV)
    U = batch.empty_pool()
    fill_punctures(batch, U, masses=(0.909, 0.0909), positions=[[0.31, 0.07, -0.11],
                                                                 [-0.52, 0.13, 0.21]], seed=2)
    D = torch.full((batch.bundle_doubles,), np.nan, dtype=torch.float64, device=DEV)
    batch.deriv_bundle_i8(U, D, k)
    torch.cuda.synchronize()
    for b, blk in enumerate(blocks):
        want = hip_bundle(batch.view(U, b).cpu().numpy(), blk.h, k)
        got = batch.bundle_view(D, b).cpu().numpy()
        assert np.array_equal(got, want), f"block {b} {blk}"
