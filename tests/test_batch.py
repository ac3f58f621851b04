"""GPU: the batched block form against the single-grid entry points, block by block.

Bar: bit-identical. A block of a batch is evaluated by the same wave program as the single-grid
kernel with the same scales (computed with the same rounding on the device), so every block's
RHS and RK4 step must equal the single-grid result on that block exactly. Cases: blocks of
uneven extents (not multiples of the 4x4x4 brick) and three spacings, FD orders 4 and 6, KO on
and off, edge-padded and periodic blocks, both launch shapes (one workgroup per brick and
persistent queues), a periodic box tiled by blocks with halo copies against the single-grid
periodic step on the whole box, and a captured step replayed from a device-resident time
(ssl factor formed by the device's exp: 1e-12 normwise).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import gpu_available, normwise

from bssn3d_hip import initial_data as idata
from bssn3d_hip.batch import Block, BlockBatch, config_blocks, fill_punctures, tiling_halo
from bssn3d_hip.evolve import BSSNEvolution
from bssn3d_hip.rhs import rhs_into
from bssn3d_hip.state import PhysicsParams, scalars

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DEV = "cuda:0"


def say(msg):
    print(f"[case] {msg}", flush=True)


def small_blocks(seed=0, n=9, ng=3):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        ext = tuple(int(v) for v in rng.choice([ng, 5, 8, 12], size=3))
        h = float(rng.choice([0.25, 0.125, 0.0625]))
        lo = tuple(float(v) for v in rng.uniform(-2, 2, size=3))
        out.append(Block(ext, (h, h, h), lo))
    return out


def make_batch(blocks, ng=3, seed=0):
    batch = BlockBatch(blocks, ng=ng, device=DEV)
    U = batch.empty_pool()
    fill_punctures(batch, U, masses=(0.909, 0.0909), positions=[[0.31, 0.07, -0.11],
                                                                 [-0.52, 0.13, 0.21]], seed=seed)
    return batch, U


@pytest.mark.parametrize("order,ko", [(4, 0.0), (4, 0.1), (6, 0.1), (8, 0.1)])
@pytest.mark.parametrize("persistent", [0, 512])
def test_batch_rhs_equals_single_grid(order, ko, persistent):
    say(f"batch rhs order={order} ko={ko} persistent={persistent}")
    ng = 3 if order == 4 else 4
    batch, U = make_batch(small_blocks(1, ng=ng), ng=ng)
    batch.set_launch(persistent)
    p = PhysicsParams()
    dt, t = 0.013, 0.7
    OUT = batch.empty_pool()
    batch.rhs_into(U, OUT, p, dt, t, order=order, ko_sigma=ko)
    for b, blk in enumerate(batch.blocks):
        F = batch.view(U, b).contiguous()
        ref = torch.empty_like(F)
        rhs_into(F, ref, scalars(p, *blk.h, dt, t), order, ko)
        torch.cuda.synchronize()
        assert torch.equal(batch.view(OUT, b), ref), f"block {b} {blk}"


@pytest.mark.parametrize("bc", ["edge", "periodic"])
@pytest.mark.parametrize("persistent", [0, 256])
def test_batch_step_equals_single_grid(bc, persistent):
    say(f"batch rk4 bc={bc} persistent={persistent}")
    batch, U = make_batch(small_blocks(2), seed=3)
    batch.set_launch(persistent)
    p = PhysicsParams()
    dt, t = 0.01, 0.25
    ref = [batch.view(U, b).clone() for b in range(len(batch.blocks))]
    for k in range(2):
        batch.step_(U, t + k * dt, dt, p, order=4, ko_sigma=0.1, bc=bc)
    for b, blk in enumerate(batch.blocks):
        evo = BSSNEvolution(blk.grid(3), p, order=4, ko_sigma=0.1,
                            bc="periodic" if bc == "periodic" else "none")
        for k in range(2):
            evo.step_(ref[b], t + k * dt, dt)
        torch.cuda.synchronize()
        assert torch.equal(batch.view(U, b), ref[b]), f"block {b} {blk}"


def test_batch_periodic_sync_equals_single_grid():
    say("batch periodic sync")
    batch, U = make_batch(small_blocks(4))
    V = U.clone()
    batch.periodic_sync_(V)
    for b, blk in enumerate(batch.blocks):
        ref = batch.view(U, b).clone()
        BSSNEvolution(blk.grid(3), order=4, bc="periodic").periodic_sync_(ref)
        torch.cuda.synchronize()
        assert torch.equal(batch.view(V, b), ref)


def test_tiled_box_with_halo_equals_single_grid():
    """A periodic 16 x 12 x 8 box cut into 5 blocks of uneven extents, ghosts exchanged by
    the halo map after every stage, equals the single-grid periodic step on the whole box."""
    say("tiled box halo")
    N = (16, 12, 8)
    h = 0.125
    grid, Ug, _ = idata.config_binary(seed=2, N=16, L=2.0, separation=0.5, device=DEV)
    # a non-cubic box: crop the cubic workload to N (then make its ghosts periodic)
    ng = 3
    Ug = Ug[:, :N[0] + 2 * ng, :N[1] + 2 * ng, :N[2] + 2 * ng].contiguous()
    from bssn3d_hip.grid import Grid
    box = Grid(*N, h, h, h, ng=ng, xmin=-1.0, ymin=-1.0, zmin=-1.0)
    evo = BSSNEvolution(box, PhysicsParams(), order=4, ko_sigma=0.1, bc="periodic")
    evo.periodic_sync_(Ug)
    tiles = [(0, 0, 0), (0, 5, 0), (9, 0, 0), (9, 7, 0), (0, 0, 5)]
    exts = [(9, 5, 5), (9, 7, 8), (7, 7, 8), (7, 5, 8), (9, 5, 3)]
    blocks = [Block(e, (h, h, h), tuple(-1.0 + h * np.asarray(t))) for t, e in zip(tiles, exts)]
    batch = BlockBatch(blocks, ng=ng, device=DEV)
    halo = tiling_halo(batch, tiles, N)
    U = batch.empty_pool()
    for b, (t, e) in enumerate(zip(tiles, exts)):
        batch.view(U, b).copy_(Ug[:, t[0]:t[0] + e[0] + 2 * ng, t[1]:t[1] + e[1] + 2 * ng,
                                  t[2]:t[2] + e[2] + 2 * ng])
    dt = 0.25 * h
    for k in range(2):
        batch.step_(U, k * dt, dt, PhysicsParams(), order=4, ko_sigma=0.1, bc="halo", halo=halo)
        evo.step_(Ug, k * dt, dt)
    torch.cuda.synchronize()
    for b, (t, e) in enumerate(zip(tiles, exts)):
        want = Ug[:, t[0]:t[0] + e[0] + 2 * ng, t[1]:t[1] + e[1] + 2 * ng,
                  t[2]:t[2] + e[2] + 2 * ng]
        assert torch.equal(batch.view(U, b), want), f"block {b}"


def test_held_ghosts_step_equals_composed_rk4():
    """bc="halo" without a map: stage inputs carry U's ghosts. Against RK4 composed in torch
    from the single-grid stage RHS (bssn_stage_rhs_f64, ghosts read as data) and enforce;
    1e-12 normwise (the fused update rounds with FMAs); U's ghosts must stay bit-identical."""
    say("held ghosts vs composed RK4")
    from bssn3d_hip import _lib
    from bssn3d_hip.rhs import stream_handle
    batch, U = make_batch(small_blocks(6, n=5), seed=5)
    p = PhysicsParams()
    dt, t = 0.01, 0.2
    U0 = U.clone()
    batch.step_(U, t, dt, p, order=4, ko_sigma=0.1, bc="halo")
    lib = _lib.load()
    ng = 3
    for b, blk in enumerate(batch.blocks):
        F = batch.view(U0, b).clone()
        _, Sx, Sy, Sz = F.shape
        inner = (slice(None),) + (slice(ng, -ng),) * 3

        def f(Y, tt):
            out = torch.empty_like(Y)
            _lib.check(lib.bssn_stage_rhs_f64(Y.data_ptr(), out.data_ptr(), Sx, Sy, Sz,
                                              _lib.dbl(scalars(p, *blk.h, dt, tt)), 4, 0.1, ng, 0,
                                              stream_handle(Y.device)))
            return out

        def upd(c, k):
            Y = F.clone()
            Y[inner] = F[inner] + c * k[inner]
            return Y
        k1 = f(F, t)
        k2 = f(upd(0.5 * dt, k1), t + 0.5 * dt)
        k3 = f(upd(0.5 * dt, k2), t + 0.5 * dt)
        k4 = f(upd(dt, k3), t + dt)
        new = upd(dt / 6.0, k1 + 2.0 * k2 + 2.0 * k3 + k4)
        enf = new.clone()
        _lib.check(lib.bssn_enforce_f64(enf.data_ptr(), Sx * Sy * Sz, p.chi_floor, p.alpha_floor,
                                        stream_handle(enf.device)))
        new[inner] = enf[inner]
        got = batch.view(U, b)
        torch.cuda.synchronize()
        e = normwise(got[inner].cpu().numpy(), new[inner].cpu().numpy())
        assert np.all(e <= 1e-12), (b, e.max())
        mask = torch.ones_like(got, dtype=torch.bool)
        mask[inner] = False
        assert torch.equal(got[mask], batch.view(U0, b)[mask]), f"ghosts of block {b} changed"


def test_captured_steps_replay():
    say("graph replay")
    batch, U = make_batch(small_blocks(5, n=6))
    p = PhysicsParams()
    dt, t0 = 0.01, 0.4
    batch.set_launch(256)
    V = U.clone()
    tdev = torch.full((1,), t0, dtype=torch.float64, device=DEV)
    g = batch.capture_steps(U, tdev, dt, nsteps=2, params=p)
    for r in range(3):
        g.replay()
        torch.cuda.synchronize()
        for k in range(2):                      # eager steps between the replays
            batch.step_(V, t0 + (2 * r + k) * dt, dt, p)
        torch.cuda.synchronize()
        for b in range(len(batch.blocks)):
            e = normwise(batch.view(U, b).cpu().numpy(), batch.view(V, b).cpu().numpy())
            assert np.all(e <= 1e-12), (r, b, e.max())
    assert abs(float(tdev.item()) - (t0 + 6 * dt)) < 1e-14


def test_autotune_and_workload_geometry():
    say("autotune")
    blocks = config_blocks(seed=0, nblk=40)
    batch = BlockBatch(blocks, ng=3, device=DEV)
    U = batch.empty_pool()
    fill_punctures(batch, U, seed=0)
    res = batch.autotune(U, dt=0.25 * 0.0625, candidates=[0, 512, 1024], reps=1)
    assert set(res) == {0, 512, 1024} and batch.persistent_waves in res
    assert batch.interior_points == sum(int(np.prod(b.n)) for b in blocks)
    assert bool(torch.isfinite(U).all())
