"""GPU: the block hierarchy through the C ABI against the restatement in oracle/amr.py.

Bars:
  * prolongation (bssn_prolong_f64), the indicator (bssn_batch_indicator_f64) and the ghost
    fill program (bssn_ghost_fill_f64) bit-identical to the restatement (no contraction in the
    prolongation on either side; the indicator and copies are exact);
  * a root-only hierarchy step bit-identical to the single-grid periodic RK4 step (interiors);
  * two shared-dt RK4 steps of a three-level hierarchy (FD orders 4 and 6) against the
    restatement's step (whole padded blocks), per field within 1e-12 or four times the
    single-grid step's own error on the same data; captured steps replayed as a graph equal to
    eager steps;
  * regrid: new children hold the prolongation of their parent, coarsening injects them back
    (bit-identical), nesting holds, capacity growth keeps the data; a run with regrids stays
    finite and refines where the feature is.
"""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest
import torch

import amr as OA
import oracle as O
from conftest import gpu_available, normwise

from bssn3d_hip import _lib
from bssn3d_hip import initial_data as idata
from bssn3d_hip.amr import (REFINE, AMRHierarchy, RegridPolicy, octant, parent_tile,
                            prolong_entries)
from bssn3d_hip.evolve import BSSNEvolution
from bssn3d_hip.grid import Grid
from bssn3d_hip.rhs import stream_handle
from bssn3d_hip.state import CHI, PhysicsParams, scalars

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DEV = "cuda:0"
CORNERS = list(itertools.product((0, 1), repeat=3))


def say(msg):
    print(f"[case] {msg}", flush=True)


def three_level(BS=8, ng=3, device=DEV, L=1.0):
    """A hierarchy on 2x2x2 root blocks: root tile (0,0,0) fully refined, 2 children of root
    tile (0,0,1) (one a z face neighbour of a child of the first), 3 level-2 tiles under level-1
    tile (1,1,1)."""
    H = AMRHierarchy(nb_root=(2, 2, 2), BS=BS, ng=ng, levels=3, device=device, L=L,
                     lo=(-L / 2,) * 3)
    T = H.tree
    for c in CORNERS:
        T.plant((1,) + tuple(c))
    for c in ((0, 0, 0), (1, 0, 0)):
        T.plant((1, c[0], c[1], 2 + c[2]))
    for c in ((0, 0, 0), (1, 0, 0), (1, 1, 1)):
        T.plant((2, 2 + c[0], 2 + c[1], 2 + c[2]))
    assert T.nesting_errors() == []
    return H


def smooth_fn(seed=0, amp=1e-3, L=1.0, lo=-0.5):
    """Flat space + seeded periodic sinusoids on the box lo + [0, L)^3 (same function on every
    block)."""
    pert = idata.Perturbation.draw(np.random.default_rng(seed))

    def fn(g: Grid):
        x = torch.as_tensor(g.x, device=DEV).view(-1, 1, 1)
        y = torch.as_tensor(g.y, device=DEV).view(1, -1, 1)
        z = torch.as_tensor(g.z, device=DEV).view(1, 1, -1)
        d = idata.minkowski(g, device=DEV)
        for f in range(24):
            acc = torch.zeros(g.shape, dtype=torch.float64, device=DEV)
            for n in range(pert.k.shape[1]):
                k = [2 * math.pi * int(v) / L for v in pert.k[f, n]]
                acc += torch.sin(k[0] * (x - lo) + k[1] * (y - lo) + k[2] * (z - lo)
                                 + float(pert.phi[f, n]))
            d[f] += amp * acc / pert.k.shape[1]
        return d
    return fn


def host_blocks(H, U):
    P = U.cpu().numpy()
    return {k: H.view(P, k).copy() for k in H.keys}


def hier_of(H):
    return {t: (t[0], H.origin(t)) for t in H.keys}


@pytest.mark.parametrize("BS,ng", [(8, 3), (16, 4)])
def test_prolong_kernel_bitwise(BS, ng):
    say(f"prolongation BS={BS} ng={ng}")
    W = BS + 2 * ng
    P = 24 * W ** 3
    rng = np.random.default_rng(BS)
    parent = rng.standard_normal((24, W, W, W))
    pool = torch.zeros(9 * P, dtype=torch.float64, device=DEV)
    pool[:P] = torch.from_numpy(parent.reshape(-1)).to(DEV)
    pts = np.stack(np.meshgrid(*[np.arange(W)] * 3, indexing="ij"), -1).reshape(-1, 3)
    e = np.concatenate([prolong_entries(pts, (i + 1) * P, 0, c, BS, ng)
                        for i, c in enumerate(CORNERS)])
    d = torch.from_numpy(e.view(np.uint8).copy()).to(DEV)
    _lib.check(_lib.load().bssn_prolong_f64(pool.data_ptr(), pool.numel(), d.data_ptr(), len(e),
                                            stream_handle(pool.device)))
    got = pool.cpu().numpy()
    for i, c in enumerate(CORNERS):
        want = OA.prolongate(parent, c, BS, ng)
        assert np.array_equal(got[(i + 1) * P:(i + 2) * P].reshape(24, W, W, W), want), c


@pytest.mark.parametrize("BS,ng", [(8, 3), (16, 4)])
def test_prolong_blocks_bitwise(BS, ng):
    """Whole-block jobs (separable passes through LDS) write exactly the restatement's
    prolongation at their points and nothing elsewhere, for every octant and several skip
    masks."""
    say(f"block prolongation BS={BS} ng={ng}")
    from amr_emul import job_points
    from bssn3d_hip.amr import block_job
    W = BS + 2 * ng
    P = 24 * W ** 3
    parent = np.random.default_rng(BS + 1).standard_normal((24, W, W, W))
    skips = [0, 64, 0b010101, 0b101010 | 64]
    jobs = [(c, sk) for c in CORNERS for sk in skips]
    pool = torch.full(((len(jobs) + 1) * P,), -7.0, dtype=torch.float64, device=DEV)
    pool[:P] = torch.from_numpy(parent.reshape(-1)).to(DEV)
    e = np.concatenate([block_job((i + 1) * P, 0, c, BS, ng, sk) for i, (c, sk) in
                        enumerate(jobs)])
    d = torch.from_numpy(e.view(np.uint8).copy()).to(DEV)
    _lib.check(_lib.load().bssn_prolong_blocks_f64(pool.data_ptr(), pool.numel(), d.data_ptr(),
                                                   len(e),
                                                   stream_handle(pool.device)))
    got = pool.cpu().numpy()
    for i, (c, sk) in enumerate(jobs):
        blk = got[(i + 1) * P:(i + 2) * P].reshape(24, W, W, W)
        want = OA.prolongate(parent, c, BS, ng)
        pts = job_points(e[i])
        mask = np.zeros((W, W, W), bool)
        mask[pts[:, 0], pts[:, 1], pts[:, 2]] = True
        assert np.array_equal(blk[:, mask], want[:, mask]), (c, sk)
        assert np.all(blk[:, ~mask] == -7.0), (c, sk)


def test_indicator_bitwise():
    say("indicator")
    H = three_level()
    U = torch.from_numpy(np.random.default_rng(3).standard_normal(H.pool_doubles)).to(DEV)
    keys = H.keys
    W = H.W
    U[H.slot_offset(keys[4]) + 5 * W ** 3 + (5 * W + 5) * W + 5] = float("nan")    # interior
    out = torch.empty(len(keys) * 24, dtype=torch.float64, device=DEV)
    _lib.check(_lib.load().bssn_batch_indicator_f64(H.batch._handle, U.data_ptr(),
                                                    out.data_ptr(), stream_handle(U.device)))
    got = out.view(len(keys), 24).cpu().numpy()
    hb = host_blocks(H, U)
    for i, k in enumerate(keys):
        want = OA.indicator(hb[k], H.ng)
        assert np.array_equal(got[i], want, equal_nan=True), k
    assert np.isnan(got[4, 5]) and not np.isnan(got[4, 4])
    lv = H.indicators(U, fields=[0, 1])
    assert lv[keys[0]] == max(got[0, 0], got[0, 1])


def test_sync_program_bitwise():
    say("ghost fill program, three levels")
    H = three_level()
    U = torch.from_numpy(np.random.default_rng(4).standard_normal(H.pool_doubles)).to(DEV)
    want = OA.sync_all(host_blocks(H, U), hier_of(H), H.nb_root, H.BS, H.ng)
    H.sync_(U)
    got = host_blocks(H, U)
    for k in H.keys:
        assert np.array_equal(got[k], want[k]), k


def test_root_only_step_equals_single_grid():
    say("root-only hierarchy step vs single-grid periodic step")
    N, BS, ng = 16, 8, 3
    H = AMRHierarchy(nb_root=(2, 2, 2), BS=BS, ng=ng, levels=1, device=DEV)
    h = H.h0
    g = Grid(N, N, N, h, h, h, ng=ng, xmin=-0.5, ymin=-0.5, zmin=-0.5)
    F = smooth_fn(1)(g)
    U = H.from_global(H.empty_pool(), F)
    p = PhysicsParams()
    dt = 0.25 * h
    ev = BSSNEvolution(g, p, order=4, ko_sigma=0.1, bc="periodic")
    Fg = ev.evolve(F, dt, 3, t0=0.2).data
    H.evolve(U, 0.2, dt, 3, params=p, order=4, ko_sigma=0.1)
    assert torch.equal(H.to_global(U), Fg[:, ng:ng + N, ng:ng + N, ng:ng + N])


def test_hierarchy_step_vs_oracle():
    """Two steps of the three-level hierarchy on the box [-8, 8)^3 (spacings 1, 1/2, 1/4)
    against the restatement's step. Bar per field: within 1e-12, or within four times what the
    single-grid periodic step itself shows against the oracle on the same data at the finest
    spacing (measured ratios up to 2; an indexing or sync mistake shows at 1e-6 and above) (flat space + 1e-3 sinusoids: K and At are ~1e-3 while their right-hand sides
    come from stencils of ~1 values, so FP64 evaluation orders differ there by a few 1e-12
    on a single grid already)."""
    say("three-level hierarchy, 2 RK4 steps vs oracle")
    Lb = 16.0
    fn = smooth_fn(2, L=Lb, lo=-Lb / 2)
    H = three_level(L=Lb)
    U = H.fill(H.empty_pool(), fn)
    p = PhysicsParams()
    dt = 0.25 * H.spacing(2)
    S16 = scalars(p, 1.0, 1.0, 1.0, dt, 0.0)
    ref = host_blocks(H, U)
    t = 0.05
    for _ in range(2):
        ref = OA.rk4_step(ref, hier_of(H), H.nb_root, H.BS, H.ng, H.h0, S16, t, dt, order=4,
                          ko_sigma=0.1)
        t += dt
    H.evolve(U, 0.05, dt, 2, params=p, order=4, ko_sigma=0.1)
    got = host_blocks(H, U)
    per_field = np.zeros(24)
    for k in H.keys:
        per_field = np.maximum(per_field, normwise(got[k], ref[k]))
    # the single-grid path on the same data at the finest spacing
    N, h = 64, H.spacing(2)
    g = Grid(N, N, N, h, h, h, ng=3, xmin=-Lb / 2, ymin=-Lb / 2, zmin=-Lb / 2)
    F = fn(g)
    sg = BSSNEvolution(g, p, order=4, ko_sigma=0.1, bc="periodic").evolve(F, dt, 2, t0=0.05)
    want = O.periodic_sync(O.evolve(F.cpu().numpy(), scalars(p, h, h, h, dt, 0.05), dt, 2,
                                    t0=0.05, order=4, ko_order=6, ko_sigma=0.1, ng=3), 3)
    single = normwise(sg.data.cpu().numpy(), want)
    print("hierarchy " + " ".join(f"{v:.1e}" for v in per_field))
    print("single    " + " ".join(f"{v:.1e}" for v in single))
    assert np.all(per_field <= np.maximum(1e-12, 4.0 * single))


def test_hierarchy_step_order6_vs_oracle():
    """FD order 6 / KO 8 with 4 ghosts (padded blocks of 24, the whole-block prolongation's
    largest), two steps on the box [-8, 8)^3, same bar as at order 4."""
    say("three-level hierarchy, order 6, 2 RK4 steps vs oracle")
    Lb = 16.0
    fn = smooth_fn(5, L=Lb, lo=-Lb / 2)
    H = three_level(BS=16, ng=4, L=Lb)
    U = H.fill(H.empty_pool(), fn)
    p = PhysicsParams()
    dt = 0.25 * H.spacing(2)
    S16 = scalars(p, 1.0, 1.0, 1.0, dt, 0.0)
    ref = host_blocks(H, U)
    t = 0.0
    for _ in range(2):
        ref = OA.rk4_step(ref, hier_of(H), H.nb_root, H.BS, H.ng, H.h0, S16, t, dt, order=6,
                          ko_sigma=0.1)
        t += dt
    H.evolve(U, 0.0, dt, 2, params=p, order=6, ko_sigma=0.1)
    got = host_blocks(H, U)
    per_field = np.zeros(24)
    for k in H.keys:
        per_field = np.maximum(per_field, normwise(got[k], ref[k]))
    N, h = 128, H.spacing(2)
    g = Grid(N, N, N, h, h, h, ng=4, xmin=-Lb / 2, ymin=-Lb / 2, zmin=-Lb / 2)
    F = fn(g)
    sg = BSSNEvolution(g, p, order=6, ko_sigma=0.1, bc="periodic").evolve(F, dt, 2, t0=0.0)
    want = O.periodic_sync(O.evolve(F.cpu().numpy(), scalars(p, h, h, h, dt, 0.0), dt, 2,
                                    t0=0.0, order=6, ko_order=8, ko_sigma=0.1, ng=4), 4)
    single = normwise(sg.data.cpu().numpy(), want)
    print("hierarchy " + " ".join(f"{v:.1e}" for v in per_field))
    print("single    " + " ".join(f"{v:.1e}" for v in single))
    assert np.all(per_field <= np.maximum(1e-12, 4.0 * single))


def test_hierarchy_graph_replay():
    """Three captured steps replayed twice, with eager work on the pool between the replays,
    against six eager steps (the SSL factor from the device's exp: 1e-12)."""
    say("hierarchy steps captured into a graph")
    H = three_level()
    fn = smooth_fn(7)
    U = H.fill(H.empty_pool(), fn)
    V = U.clone()
    p = PhysicsParams()
    dt = 0.25 * H.spacing(2)
    H.evolve(V, 0.1, dt, 6, params=p, order=4, ko_sigma=0.1)
    tdev = torch.full((1,), 0.1, dtype=torch.float64, device=DEV)
    g = H.capture_steps(U, tdev, dt, 3, params=p, order=4, ko_sigma=0.1)
    g.replay()
    torch.cuda.synchronize()
    W = U.clone()
    H.sync_(W)                         # eager work between the replays
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    assert abs(float(tdev) - (0.1 + 6 * dt)) < 1e-14
    worst = 0.0
    a, b = host_blocks(H, U), host_blocks(H, V)
    for k in H.keys:
        worst = max(worst, float(normwise(a[k], b[k]).max()))
    print(f"graph vs eager {worst:.2e}")
    assert worst <= 1e-12


@pytest.mark.parametrize("BS,ng", [(8, 3), (16, 4)])
def test_linked_step_equals_program_step(BS, ng):
    """Linked steps (face-copied ghosts read at their sources, only the prolongations and the
    copies their taps need between stages) equal program steps bit for bit: every point after
    a step that ends with the full program, the interiors after steps that do not."""
    say(f"linked hierarchy steps vs program steps, BS={BS}, ng={ng}")
    H = three_level(BS, ng)
    U = H.fill(H.empty_pool(), smooth_fn(9))
    p = PhysicsParams()
    dt = 0.25 * H.spacing(2)
    order = 4 if ng == 3 else 6
    A, B = U.clone(), U.clone()
    H.step_(A, 0.1, dt, p, order=order, ko_sigma=0.1)
    H.step_(B, 0.1, dt, p, order=order, ko_sigma=0.1, ghosts="linked")
    a, b = host_blocks(H, A), host_blocks(H, B)
    for k in H.keys:
        assert np.array_equal(a[k], b[k]), k
    for n in range(2):
        H.step_(A, 0.1 + (n + 1) * dt, dt, p, order=order, ko_sigma=0.1)
        H.step_(B, 0.1 + (n + 1) * dt, dt, p, order=order, ko_sigma=0.1, ghosts="linked",
                sync_end=False)
    a, b = host_blocks(H, A), host_blocks(H, B)
    I = (slice(None),) + (slice(ng, ng + BS),) * 3
    for k in H.keys:
        assert np.array_equal(a[k][I], b[k][I]), k
    H.sync_(B)
    b = host_blocks(H, B)
    for k in H.keys:
        assert np.array_equal(a[k], b[k]), k


def bump_fn(center=(0.2, 0.2, 0.2), width=0.08, amp=0.05):
    def fn(g: Grid):
        x = torch.as_tensor(g.x, device=DEV).view(-1, 1, 1)
        y = torch.as_tensor(g.y, device=DEV).view(1, -1, 1)
        z = torch.as_tensor(g.z, device=DEV).view(1, 1, -1)
        d = idata.minkowski(g, device=DEV)
        r2 = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
        d[CHI] += amp * torch.exp(-r2 / width ** 2)
        return d
    return fn


def test_regrid_refine_then_coarsen():
    say("regrid: refine a bump's block, then coarsen it back")
    H = AMRHierarchy(nb_root=(2, 2, 2), BS=8, ng=3, levels=2, device=DEV)
    U = H.fill(H.empty_pool(), bump_fn())
    before = host_blocks(H, U)
    ind = H.indicators(U, fields=[CHI])
    hot = [t for t in H.keys if ind[t] > 1e-3]
    assert len(hot) >= 1
    U = H.regrid(U, RegridPolicy(refine_above=1e-3, coarsen_below=1e-9, patience=1, buffer=0))
    T = H.tree
    kids = T.on_level(1)
    assert len(kids) == 8 * len(hot) and T.nesting_errors() == []
    after = host_blocks(H, U)
    I = (slice(None),) + (slice(3, 11),) * 3
    for k in kids:
        want = OA.prolongate(before[parent_tile(k)], octant(k), 8, 3)
        assert np.array_equal(after[k][I], want[I]), k
    # every block now reads low: the children coarsen back (injection restores the parent)
    U = H.regrid(U, RegridPolicy(refine_above=1e9, coarsen_below=1e8, patience=1))
    assert T.on_level(1) == [] and T.nesting_errors() == []
    final = host_blocks(H, U)
    for k in H.keys:
        assert np.array_equal(final[k][I], before[k][I]), k


def test_capacity_growth_keeps_data():
    say("slot capacity growth")
    H = AMRHierarchy(nb_root=(2, 2, 2), BS=8, ng=3, levels=2, caps=[8, 4], max_blocks=64,
                     device=DEV)
    U = H.fill(H.empty_pool(), bump_fn(center=(0.0, 0.0, 0.0)))
    before = host_blocks(H, U)
    U2 = H.apply(U, {(0, 0, 0, 0): REFINE})
    assert H.tree.capacity(1) >= 8 and H.tree.widenings and U2.numel() == H.pool_doubles
    after = host_blocks(H, U2)
    I = (slice(None),) + (slice(3, 11),) * 3
    for k in before:
        assert np.array_equal(after[k][I], before[k][I])
    assert len(H.tree.on_level(1)) == 8


def test_run_with_regrids_follows_bump():
    say("evolve with regrids")
    H = AMRHierarchy(nb_root=(2, 2, 2), BS=8, ng=3, levels=3, device=DEV)
    U = H.fill(H.empty_pool(), bump_fn())
    dt = 0.25 * H.spacing(2)
    seen = []
    pol = RegridPolicy(refine_above=2e-3, coarsen_below=1e-6, patience=1, buffer=0)
    U, t = H.run(U, 0.0, dt, 8, pol, every=2, fields=[CHI],
                 on_regrid=lambda n, u, h: seen.append([h.tree.count(L) for L in range(3)]),
                 params=PhysicsParams(), order=4, ko_sigma=0.1)
    assert abs(t - 8 * dt) < 1e-15 and len(seen) == 3
    assert seen[0][1] > 0 and H.tree.nesting_errors() == []
    assert bool(torch.isfinite(U).all())
    # the bump's centre lies inside a level-1 block
    def holds(k):
        b = H.block(k)
        return all(b.lo[a] <= 0.2 < b.lo[a] + 8 * b.h[a] for a in range(3))
    assert any(holds(k) for k in H.tree.on_level(1))
