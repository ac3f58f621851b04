"""GPU: the block-floating-point bundle storage through the C ABI.

Bars:
  * per-array maxima (bssn_deriv_bundle_absmax_f64) equal max |d| of the device's own fp64
    bundle bit for bit, and the exponents formed from them equal the restatement's;
  * BFP storage (bssn_deriv_bundle_bfp, int32 and 48-bit limbs) holds exactly the restatement's
    mantissas of that fp64 bundle; the header holds 2^-e;
  * the algebra from BFP storage (bssn_algebra_bfp_f64) equals the fp64 algebra on the
    dequantized bundle bit for bit;
  * on the reference's Gowdy gate data (N=16, FD order 6) the BFP32 RHS error against the
    fp64 RHS is the restatement's (7.33e-10, relative L2, within 2 %) and BFP48 is below
    1e-13; the composed BFP RK4 step tracks the fused step; the tracking gate passes for the
    BFP48 and BFP32 schemes.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import bfp as B
import oracle as O
from conftest import gpu_available, normwise

from bssn3d_hip import _lib
from bssn3d_hip import tracking_gate as tg
from bssn3d_hip import initial_data as idata
from bssn3d_hip.evolve import BSSNEvolution
from bssn3d_hip.grid import Grid
from bssn3d_hip.rhs import BSSNSolver, bfp_buffers, bfp_exponents, rhs_bfp_into, stream_handle
from bssn3d_hip.state import PhysicsParams, scalars

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DEV = "cuda:0"


def say(msg):
    print(f"[case] {msg}", flush=True)


def puncture(N=16, order=4):
    grid, st, _ = idata.config_single_puncture(seed=4, N=N, ng=order // 2 + 1)
    return grid, st.to(DEV)


def gpu_bundle(F, h3, order):
    _, Sx, Sy, Sz = F.shape
    D = torch.empty((138, Sx, Sy, Sz), dtype=torch.float64, device=DEV)
    _lib.check(_lib.load().bssn_deriv_bundle_f64(F.data_ptr(), D.data_ptr(), Sx, Sy, Sz, order,
                                                 _lib.dbl(h3), stream_handle(F.device)))
    return D


def gpu_absmax(F, h3, order):
    _, Sx, Sy, Sz = F.shape
    amax = torch.empty(138, dtype=torch.float64, device=DEV)
    _lib.check(_lib.load().bssn_deriv_bundle_absmax_f64(F.data_ptr(), amax.data_ptr(), Sx, Sy,
                                                        Sz, order, _lib.dbl(h3),
                                                        stream_handle(F.device)))
    return amax.cpu().numpy()


def gpu_store(F, h3, order, k, exps):
    _, Sx, Sy, Sz = F.shape
    Q, _ = bfp_buffers(F, k)
    e = (ctypes.c_int * 138)(*[int(v) for v in exps])
    _lib.check(_lib.load().bssn_deriv_bundle_bfp(F.data_ptr(), Q.data_ptr(), e, k, Sx, Sy, Sz,
                                                 order, _lib.dbl(h3), stream_handle(F.device)))
    return Q


def unpack(Q, k, P):
    q = Q.cpu().numpy()
    inv = q[:138 * 8].view(np.float64)
    body = q[1280:]
    if k <= 32:
        m = body[:138 * P * 4].view(np.int32).reshape(138, P).astype(np.int64)
    else:
        m = B.unpack48(body[:138 * P * 6].view(np.uint16).reshape(138, P, 3))
    return inv, m


@pytest.mark.parametrize("order", [4, 6])
def test_absmax_and_exponents(order):
    say(f"bundle maxima, order {order}")
    grid, F = puncture(16, order)
    h3 = (grid.dx, grid.dy, grid.dz)
    D = gpu_bundle(F, h3, order).cpu().numpy().reshape(138, -1)
    amax = gpu_absmax(F, h3, order)
    assert np.array_equal(amax, np.max(np.abs(D), axis=1))
    for k in (24, 32, 48):
        assert np.array_equal(bfp_exponents(amax, k),
                              [B.block_exponent(D[i], k) for i in range(138)])


@pytest.mark.parametrize("k", [24, 32, 40, 48])
def test_bfp_storage_bitwise(k):
    say(f"BFP{k} storage")
    grid, F = puncture(16, 4)
    h3 = (grid.dx, grid.dy, grid.dz)
    D = gpu_bundle(F, h3, 4).cpu().numpy().reshape(138, -1)
    exps = bfp_exponents(np.max(np.abs(D), axis=1), k)
    Q = gpu_store(F, h3, 4, k, exps)
    inv, m = unpack(Q, k, D.shape[1])
    assert np.array_equal(inv, np.ldexp(1.0, -exps))
    want = np.stack([B.mantissas(D[i], k, int(exps[i])) for i in range(138)])
    assert np.array_equal(m, want.astype(np.int64))


@pytest.mark.parametrize("k", [24, 48])
def test_bfp_storage_nan_saturates(k):
    say(f"BFP{k} storage of a bundle with a NaN, fixed exponents")
    grid, F = puncture(16, 4)
    h3 = (grid.dx, grid.dy, grid.dz)
    D0 = gpu_bundle(F, h3, 4).cpu().numpy().reshape(138, -1)
    exps = bfp_exponents(np.max(np.abs(D0), axis=1), k)     # fixed from the finite state
    F[5, 9, 10, 11] = float("nan")                          # gt2 (gt_xz) at one interior point
    D = gpu_bundle(F, h3, 4).cpu().numpy().reshape(138, -1)
    assert np.isnan(D).any()
    Q = gpu_store(F, h3, 4, k, exps)
    _, m = unpack(Q, k, D.shape[1])
    want = np.stack([B.mantissas(D[i], k, int(exps[i])) for i in range(138)])
    assert np.array_equal(m, want.astype(np.int64))
    assert np.all(m[np.isnan(D)] == -(2 ** (k - 1)))


@pytest.mark.parametrize("k", [32, 48])
def test_bfp_algebra_bitwise(k):
    say(f"algebra from BFP{k}")
    grid, F = puncture(16, 4)
    h3 = (grid.dx, grid.dy, grid.dz)
    _, Sx, Sy, Sz = F.shape
    P = Sx * Sy * Sz
    S16 = scalars(PhysicsParams(), *h3, 0.25 * grid.dx, 0.3)
    D = gpu_bundle(F, h3, 4)
    exps = bfp_exponents(np.max(np.abs(D.cpu().numpy().reshape(138, -1)), axis=1), k)
    Q = gpu_store(F, h3, 4, k, exps)
    out = torch.empty_like(F)
    lib = _lib.load()
    _lib.check(lib.bssn_algebra_bfp_f64(F.data_ptr(), Q.data_ptr(), k, _lib.dbl(S16),
                                        out.data_ptr(), P, stream_handle(F.device)))
    inv, m = unpack(Q, k, P)
    Dq = torch.from_numpy(np.ascontiguousarray(m * inv[:, None])).to(DEV)
    ref = torch.empty_like(F)
    _lib.check(lib.bssn_algebra_f64(F.data_ptr(), Dq.data_ptr(), _lib.dbl(S16), ref.data_ptr(),
                                    P, stream_handle(F.device)))
    assert torch.equal(out, ref)


def test_gowdy_gate_values():
    say("Gowdy N=16 order 6: BFP RHS errors")
    g = Grid.from_domain(16)
    F = idata.gowdy(g, 1.0, device=DEV)
    I = (slice(None),) + (slice(g.ng, -g.ng),) * 3
    ref = BSSNSolver(g, order=6, scheme="twopass").rhs_tensor(F, 0.0, 0.25 * g.dx)[I]
    l2 = lambda x: float(torch.sqrt(torch.mean(x ** 2)))
    errs = {}
    for k in (48, 32, 24):
        out = BSSNSolver(g, order=6, scheme=f"bfp{k}").rhs_tensor(F, 0.0, 0.25 * g.dx)[I]
        errs[k] = l2(out - ref) / l2(ref)
    print({k: f"{v:.3e}" for k, v in errs.items()})
    assert errs[48] < 1e-13 and errs[48] < errs[32] < errs[24]
    assert errs[32] == pytest.approx(7.327e-10, rel=0.02)


@pytest.mark.parametrize("k,tol", [(48, 1e-12), (32, 1e-7)])
def test_bfp_composed_step_tracks_fused(k, tol):
    say(f"BFP{k} composed RK4 step vs fused")
    grid, F = puncture(20, 4)
    dt = 0.25 * grid.dx
    a = BSSNEvolution(grid, PhysicsParams(), order=4, ko_sigma=0.1, scheme=f"bfp{k}")
    b = BSSNEvolution(grid, PhysicsParams(), order=4, ko_sigma=0.1, scheme="fused")
    e = normwise(a.evolve(F, dt, 3, t0=0.1).data.cpu().numpy(),
                 b.evolve(F, dt, 3, t0=0.1).data.cpu().numpy())
    print(f"bfp{k}: max {e.max():.2e}")
    assert e.max() <= tol


@pytest.mark.parametrize("scheme", ["bfp48", "bfp32"])
def test_tracking_gate_bfp(scheme):
    say(f"tracking gate {scheme}")
    rep = tg.run_gate(scheme, order=6, device=DEV)
    print(rep.table())
    assert rep.passed
