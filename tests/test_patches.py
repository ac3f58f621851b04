"""GPU: the batched block-AMR workload at its full size -- 640 uneven blocks (interior extents
from {8, 12, 16, 24, 32} per axis) on three nested levels (h = 1/4, 1/8, 1/16) around the 1:10
puncture binary, stepped by one batched launch per RK stage with the ghost fill program after
every stage (bssn_batch_rk4_step_prog_f64).

Bars:
  * the device ghost fill on all 640 blocks reproduces a degree-5-per-axis polynomial at every
    refined-level ghost (1e-13 of its scale) and copies the root ghosts from their periodic
    images exactly;
  * the batched RHS (+KO) of every block is finite, and bit-identical to the single-grid fused
    RHS (bssn_rhs_f64, each block alone with its own spacings) on a seeded sample of 48 blocks
    -- the per-block bar of the batched form;
  * one hierarchy RK4 step equals the same step composed from the primitives (ghost fill,
    batched RHS + KO at the interior points, the RK4 combinations in torch, enforce), per block
    and field within 1e-12 normwise (the fused update rounds with FMAs);
  * steps captured into a HIP graph and replayed (time on the device), with eager work between
    the replays, equal eager steps to 1e-12 (the device's exp forms the SSL factor);
  * a linked step (ghost copies read at their sources, only the prolongations and the copies
    they need between stages) equals the program step bit for bit on every block -- ghosts
    included when it ends with the full program, interiors when it does not -- and captured
    linked steps replay to the eager ones.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import gpu_available, normwise

from bssn3d_hip import _lib
from bssn3d_hip.patches import config_hierarchy, fill_binary
from bssn3d_hip.rhs import rhs_into, stream_handle
from bssn3d_hip.state import PhysicsParams, scalars

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DEV = "cuda:0"


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def hier():
    say("building the 640-block hierarchy")
    H, pos = config_hierarchy(seed=0, device=DEV)
    assert len(H) == 640
    H.program                                   # build and upload the ghost program
    return H, pos


def _interior_mask(H) -> torch.Tensor:
    m = torch.zeros(H.layout.pool_doubles, dtype=torch.bool, device=DEV)
    ng = H.ng
    for b in range(len(H)):
        H.view(m, b)[:, ng:-ng, ng:-ng, ng:-ng] = True
    return m


def test_ghost_fill_reproduces_polynomials(hier):
    H, _ = hier
    say("640-block ghost fill vs a degree-5 polynomial")
    ng = H.ng
    root = H.boxes[0]
    period = [root.extent[a] * root.h for a in range(3)]
    U = H.empty_pool()
    U.fill_(float("nan"))

    def poly(b, wrap):
        g = H.blocks[b].grid(ng)
        c = []
        for a, v in enumerate((g.x, g.y, g.z)):
            t = torch.as_tensor(v, device=DEV)
            if wrap:
                t = H.x0[a] + torch.remainder(t - H.x0[a], period[a])
            c.append(t)
        x, y, z = c[0].view(-1, 1, 1), c[1].view(1, -1, 1), c[2].view(1, 1, -1)
        px = 0.3 + 0.02 * x - 0.005 * x ** 2 + 0.001 * x ** 3 - 2e-5 * x ** 4 + 3e-6 * x ** 5
        py = 1.1 - 0.01 * y + 0.003 * y ** 2 - 4e-4 * y ** 3 + 5e-6 * y ** 4 - 2e-7 * y ** 5
        pz = -0.7 + 0.015 * z + 0.002 * z ** 2 - 3e-4 * z ** 3 + 1e-5 * z ** 4 + 4e-7 * z ** 5
        return px * py * pz

    want = {}
    for b in range(len(H)):
        f = poly(b, H.level[b] == 0)
        want[b] = f
        v = H.view(U, b)
        for k in range(24):
            v[k, ng:-ng, ng:-ng, ng:-ng] = (1.0 + 0.01 * k) * f[ng:-ng, ng:-ng, ng:-ng]
    H.sync_(U)
    torch.cuda.synchronize()
    for b in range(len(H)):                     # every ghost of every block written
        assert not bool(torch.isnan(H.view(U, b)).any()), b
    scale = max(float(f.abs().max()) for f in want.values())
    worst = [0.0, 0.0]
    for b in range(len(H)):
        v = H.view(U, b)
        for k in (0, 11, 23):
            err = float((v[k] - (1.0 + 0.01 * k) * want[b]).abs().max()) / scale
            lv = min(H.level[b], 1)
            worst[lv] = max(worst[lv], err)
    print(f"root ghosts {worst[0]:.1e}, refined-level ghosts {worst[1]:.1e}")
    assert worst[0] == 0.0 and worst[1] <= 1e-13


def test_tile_program_equals_point_program(hier):
    H, _ = hier
    say("640-block ghost fill: prolongation tiles vs point entries")
    from bssn3d_hip.amr import DeviceProgram
    U = fill_binary(H, H.empty_pool(), sync=False)
    V = U.clone()
    pts = DeviceProgram(*H.program_entries(), H.device, tiles=False)
    assert H.program.n_tile > 0 and pts.n_tile == 0
    # tiles holding more than 64 entries (neighbouring blocks' ghosts at the same fine position
    # each have one) take the kernel's second pass over the tile's points
    assert int(H.program.tiles["count"].max()) > 64
    H.sync_(U)
    _lib.check(_lib.load().bssn_ghost_fill_f64(V.data_ptr(), V.numel(), pts.ptr,
                                               stream_handle(V.device)))
    torch.cuda.synchronize()
    for b in range(len(H)):
        assert torch.equal(H.view(U, b), H.view(V, b)), b


def test_batched_rhs_equals_single_grid_sample(hier):
    H, _ = hier
    say("640-block batched RHS+KO vs bssn_rhs_f64 on 48 sampled blocks")
    U = fill_binary(H, H.empty_pool())
    p = PhysicsParams()
    dt, t = 0.25 * H.boxes[-1].h, 0.3
    OUT = H.empty_pool()
    H.batch.rhs_into(U, OUT, p, dt, t, order=4, ko_sigma=0.1)
    torch.cuda.synchronize()
    for b in range(len(H)):
        assert bool(torch.isfinite(H.view(OUT, b)).all()), b
    rng = np.random.default_rng(11)
    sample = sorted(set(rng.choice(len(H), size=48, replace=False).tolist()) | {0, 639})
    for b in sample:
        blk = H.blocks[b]
        F = H.view(U, b).contiguous()
        ref = torch.empty_like(F)
        rhs_into(F, ref, scalars(p, *blk.h, dt, t), 4, 0.1)
        torch.cuda.synchronize()
        assert torch.equal(H.view(OUT, b), ref), f"block {b} {blk}"


def test_step_equals_composed_primitives(hier):
    H, _ = hier
    say("640-block hierarchy RK4 step vs the composed step")
    U = fill_binary(H, H.empty_pool())
    p = PhysicsParams()
    dt, t = 0.25 * H.boxes[-1].h, 0.2
    U0 = U.clone()
    H.step_(U, t, dt, p, order=4, ko_sigma=0.1)
    inner = _interior_mask(H)
    lib = _lib.load()

    def f(Y, tt):
        out = H.empty_pool()
        H.batch.rhs_into(Y, out, p, dt, tt, order=4, ko_sigma=0.1)
        return out

    def upd(base, c, k):
        Y = base.clone()
        Y[inner] = base[inner] + c * k[inner]
        return H.sync_(Y)

    k1 = f(U0, t)
    k2 = f(upd(U0, 0.5 * dt, k1), t + 0.5 * dt)
    k3 = f(upd(U0, 0.5 * dt, k2), t + 0.5 * dt)
    k4 = f(upd(U0, dt, k3), t + dt)
    new = U0.clone()
    new[inner] = U0[inner] + (dt / 6.0) * (k1[inner] + 2.0 * k2[inner] + 2.0 * k3[inner]
                                           + k4[inner])
    del k1, k2, k3, k4
    for b in range(len(H)):
        v = H.view(new, b)
        _, Sx, Sy, Sz = v.shape
        _lib.check(lib.bssn_enforce_f64(v.data_ptr(), Sx * Sy * Sz, p.chi_floor, p.alpha_floor,
                                        stream_handle(v.device)))
    H.sync_(new)
    torch.cuda.synchronize()
    worst = np.zeros(24)
    for b in range(len(H)):
        e = normwise(H.view(U, b).cpu().numpy(), H.view(new, b).cpu().numpy())
        worst = np.maximum(worst, e)
    print("per field worst " + " ".join(f"{v:.1e}" for v in worst))
    assert np.all(worst <= 1e-12)


def test_graph_replay_equals_eager(hier):
    H, _ = hier
    say("640-block hierarchy: captured steps replayed")
    U = fill_binary(H, H.empty_pool())
    V = U.clone()
    p = PhysicsParams()
    dt, t0 = 0.25 * H.boxes[-1].h, 0.1
    H.tune_launch(U, dt, candidates=[0, 1024], reps=1, params=p, order=4, ko_sigma=0.1)
    tdev = torch.full((1,), t0, dtype=torch.float64, device=DEV)
    g = H.capture_steps(U, tdev, dt, 2, params=p, order=4, ko_sigma=0.1)
    for r in range(2):
        g.replay()
        torch.cuda.synchronize()
        for k in range(2):
            H.step_(V, t0 + (2 * r + k) * dt, dt, p, order=4, ko_sigma=0.1)
        torch.cuda.synchronize()
    assert abs(float(tdev.item()) - (t0 + 4 * dt)) < 1e-14
    worst = 0.0
    for b in range(len(H)):
        worst = max(worst, float(normwise(H.view(U, b).cpu().numpy(),
                                          H.view(V, b).cpu().numpy()).max()))
    print(f"graph vs eager {worst:.2e}")
    assert worst <= 1e-12


def _blocks_equal(H, A, B, interior_only=False):
    ng = H.ng
    for b in range(len(H)):
        a, c = H.view(A, b), H.view(B, b)
        if interior_only:
            a, c = a[:, ng:-ng, ng:-ng, ng:-ng], c[:, ng:-ng, ng:-ng, ng:-ng]
        if not torch.equal(a, c):
            return b
    return None


def test_linked_step_equals_program_step(hier):
    H, _ = hier
    say("640-block hierarchy: linked steps vs program steps")
    U = fill_binary(H, H.empty_pool())
    p = PhysicsParams()
    dt, t = 0.25 * H.boxes[-1].h, 0.15
    A, B = U.clone(), U.clone()
    H.step_(A, t, dt, p, order=4, ko_sigma=0.1)
    H.step_(B, t, dt, p, order=4, ko_sigma=0.1, ghosts="linked")
    torch.cuda.synchronize()
    assert _blocks_equal(H, A, B) is None
    # a linked step that leaves the copied ghosts stale: the next linked step reads none of them
    H.step_(A, t + dt, dt, p, order=4, ko_sigma=0.1)
    H.step_(A, t + 2 * dt, dt, p, order=4, ko_sigma=0.1)
    H.step_(B, t + dt, dt, p, order=4, ko_sigma=0.1, ghosts="linked", sync_end=False)
    H.step_(B, t + 2 * dt, dt, p, order=4, ko_sigma=0.1, ghosts="linked", sync_end=False)
    torch.cuda.synchronize()
    assert _blocks_equal(H, A, B, interior_only=True) is None
    H.sync_(B)
    torch.cuda.synchronize()
    assert _blocks_equal(H, A, B) is None


def test_linked_graph_replay_equals_eager(hier):
    H, _ = hier
    say("640-block hierarchy: captured linked steps replayed")
    U = fill_binary(H, H.empty_pool())
    V = U.clone()
    p = PhysicsParams()
    dt, t0 = 0.25 * H.boxes[-1].h, 0.05
    tdev = torch.full((1,), t0, dtype=torch.float64, device=DEV)
    g = H.capture_steps(U, tdev, dt, 2, params=p, order=4, ko_sigma=0.1, ghosts="linked",
                        sync_end=False)
    for r in range(2):
        g.replay()
        torch.cuda.synchronize()
        for k in range(2):
            H.step_(V, t0 + (2 * r + k) * dt, dt, p, order=4, ko_sigma=0.1, ghosts="linked",
                    sync_end=False)
        torch.cuda.synchronize()
    assert abs(float(tdev.item()) - (t0 + 4 * dt)) < 1e-14
    worst = 0.0
    ng = H.ng
    for b in range(len(H)):
        a = H.view(U, b)[:, ng:-ng, ng:-ng, ng:-ng].cpu().numpy()
        c = H.view(V, b)[:, ng:-ng, ng:-ng, ng:-ng].cpu().numpy()
        worst = max(worst, float(normwise(a, c).max()))
    print(f"linked graph vs eager {worst:.2e}")
    assert worst <= 1e-12
