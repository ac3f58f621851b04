"""CPU: the hierarchy transfers restated in oracle/amr.py, pinned, and the host-built ghost
program checked against them.

Pins of the restatement (the reference's own transfer tests, 3D/tests/multipatch_amr/
test_transfer.py and test_sync.py, plus what the mathematics fixes):
  * the midpoint weights are (3, -25, 150, 150, -25, 3)/256 exactly;
  * coincident fine nodes copy the parent exactly (all 8 octants); a degree-5 polynomial is
    reproduced at fine nodes whose stencils stay inside the parent; restriction of a
    prolongated child gives back the parent's nodes exactly;
  * the indicator: 0 on constants, 1 on a unit ramp, a brute-force loop on random data, NaN
    propagation;
  * the root sync equals the single-grid periodic sync; a child's halo is its parent's
    prolongation with the interior preserved; same-level face slabs are exact copies;
  * a hierarchy step of the root level alone equals the single-grid periodic RK4 step bit for
    bit (interiors).
The product's ghost program (bssn3d_hip.amr.build_ghost_program), applied with a host
emulation of the two device kernels (tests/amr_emul.py), equals the restatement's sync bit for
bit on a three-level hierarchy with partial levels.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import amr as OA
import oracle as O
from amr_emul import apply_prolongs, expand_jobs, run_program

from bssn3d_hip import initial_data as idata
from bssn3d_hip.amr import (AMRHierarchy, build_ghost_program, prolong_entries,
                            restrict_entries)
from bssn3d_hip.grid import Grid
from bssn3d_hip.state import PhysicsParams, scalars

CORNERS = list(itertools.product((0, 1), repeat=3))


def _rand(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def test_midpoint_weights_exact():
    assert np.array_equal(OA.W_MID, np.array([3, -25, 150, 150, -25, 3]) / 256.0)
    x = np.array(OA.OFFS, dtype=np.float64)
    for d in range(6):                                   # exact on degree <= 5 at 1/2
        assert abs(np.dot(OA.W_MID, x ** d) - 0.5 ** d) < 1e-15


@pytest.mark.parametrize("BS,ng", [(8, 3), (16, 4)])
@pytest.mark.parametrize("corner", CORNERS)
def test_coincident_copy_and_restrict_inverse(BS, ng, corner):
    W = BS + 2 * ng
    parent = _rand((24, W, W, W), 1)
    child = OA.prolongate(parent, corner, BS, ng)
    half = BS // 2
    a = np.arange(half)
    got = child[np.ix_(range(24), ng + 2 * a, ng + 2 * a, ng + 2 * a)]
    p = [ng + corner[q] * half + a for q in range(3)]
    want = parent[np.ix_(range(24), p[0], p[1], p[2])]
    assert np.array_equal(got, want)
    back = OA.restrict_into_parent(np.zeros_like(parent), child, corner, BS, ng)
    assert np.array_equal(back[np.ix_(range(24), p[0], p[1], p[2])], want)


def test_polynomial_exact():
    BS, ng = 8, 3
    W = BS + 2 * ng
    h = 0.37
    c = [np.array([0.4, -1.1, 0.7, 0.2, -0.3, 0.15]), np.array([-0.5, 0.9, -0.2, 0.6, 0.1, -0.25]),
         np.array([0.8, 0.3, -0.6, -0.1, 0.45, 0.05])]
    P = lambda cc, x: sum(cc[d] * x ** d for d in range(6))
    xp = (np.arange(W) - ng) * h
    parent = np.einsum("i,j,k->ijk", P(c[0], xp), P(c[1], xp), P(c[2], xp))[None].repeat(24, 0)
    child = OA.prolongate(parent, (0, 0, 0), BS, ng)
    ff = np.arange(ng + 3, ng + BS - 3)
    xf = (ff - ng) * h / 2
    want = np.einsum("i,j,k->ijk", P(c[0], xf), P(c[1], xf), P(c[2], xf))
    assert np.max(np.abs(child[np.ix_(range(24), ff, ff, ff)] - want[None])) < 1e-9


def test_indicator():
    ng, W = 3, 14
    assert np.all(OA.indicator(np.ones((24, W, W, W)), ng) == 0.0)
    ramp = np.broadcast_to(np.arange(W, dtype=float)[None, :, None, None], (24, W, W, W))
    assert np.all(OA.indicator(np.ascontiguousarray(ramp), ng) == 1.0)
    B = _rand((24, 8, 9, 10), 2)
    got = OA.indicator(B, 2)
    for f in (0, 7, 23):
        m = 0.0
        for x in range(1, 7):
            for y in range(1, 8):
                for z in range(1, 9):
                    for d in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                        q = (x + d[0], y + d[1], z + d[2])
                        if q[0] <= 6 and q[1] <= 7 and q[2] <= 8:
                            m = max(m, abs(B[f][q] - B[f, x, y, z]))
        assert got[f] == m
    B[5, 4, 4, 4] = np.nan
    assert np.isnan(OA.indicator(B, 2)[5]) and not np.isnan(OA.indicator(B, 2)[4])


# --- hierarchies ----------------------------------------------------------------------------

def three_level(BS=8, ng=3):
    """A hierarchy on 2x2x2 root blocks: root tile (0,0,0) fully refined, 2 children of root
    tile (0,0,1) (one a z face neighbour of a child of the first), 3 level-2 tiles under level-1
    tile (1,1,1)."""
    H = AMRHierarchy(nb_root=(2, 2, 2), BS=BS, ng=ng, levels=3, device="cpu")
    T = H.tree
    for c in CORNERS:
        T.plant((1,) + tuple(c))
    for c in ((0, 0, 0), (1, 0, 0)):
        T.plant((1, c[0], c[1], 2 + c[2]))
    for c in ((0, 0, 0), (1, 0, 0), (1, 1, 1)):
        T.plant((2, 2 + c[0], 2 + c[1], 2 + c[2]))
    assert T.nesting_errors() == []
    return H


def _hier(H):
    return {t: (t[0], H.origin(t)) for t in H.keys}


def _views(H, pool):
    return {k: H.view(pool, k) for k in H.keys}


def test_program_emulation_equals_oracle_sync():
    H = three_level()
    pool = _rand(H.pool_doubles, 3)
    want = OA.sync_all({k: v.copy() for k, v in _views(H, pool).items()}, _hier(H), H.nb_root,
                       H.BS, H.ng)
    for use_blocks in (True, False):
        P = pool.copy()
        pr, jb, cp, pe, be, ce = build_ghost_program(H.tree, H.BS, H.ng, H.slot_offset,
                                                     use_blocks)
        assert len(pe) == 3 and pe[0] == 0 and ce[0] > 0
        assert (len(jb) > 0) == use_blocks and (len(pr) > 0) != use_blocks
        run_program(P, pr, jb, cp, pe, be, ce)
        for k, v in _views(H, P).items():
            assert np.array_equal(v, want[k]), (use_blocks, k)


def test_program_covers_every_ghost_once():
    H = three_level()
    pr, jb, cp, pe, be, ce = H.program_entries()
    dst = np.concatenate([pr["dst"], expand_jobs(jb)["dst"], cp["dst"]])
    assert len(np.unique(dst)) == len(dst)
    W, ng, BS = H.W, H.ng, H.BS
    nghost = W ** 3 - BS ** 3
    assert len(dst) == nghost * len(H.keys)
    # copies read interior points only
    src = cp["src"]
    for k in H.keys:
        o = H.slot_offset(k)
        m = (src >= o) & (src < o + W ** 3)
        loc = src[m] - o
        c = np.stack([loc // (W * W), (loc // W) % W, loc % W], -1)
        assert np.all((c >= ng) & (c < ng + BS))


def test_linked_program_reproduces_prolongations():
    """The program of a linked step (bssn3d_hip.amr.linked_entries: every prolongation, only
    the copies their taps read) gives every prolongated ghost -- point entries and whole-block
    jobs alike -- the value the full program gives it, bit for bit, on a pool whose ghosts
    start as NaN; the product's expansion of jobs into points (job_prolong_points) equals the
    emulator's; the link map sends each copied ghost to its source and every other point to
    itself."""
    from bssn3d_hip.amr import job_prolong_points, link_map, linked_entries
    H = three_level()
    W, ng, BS = H.W, H.ng, H.BS
    for use_blocks in (True, False):
        ent = build_ghost_program(H.tree, BS, ng, H.slot_offset, use_blocks)
        pr, jb, cp, pe, be, ce = ent
        lpr, ljb, lcp, lpe, lbe, lce = linked_entries(ent, ng)
        assert len(lpr) == len(pr) and len(ljb) == len(jb) and len(lcp) < len(cp)
        if use_blocks:
            a, b = job_prolong_points(jb), expand_jobs(jb)
            assert all(np.array_equal(a[k], b[k]) for k in ("dst", "parent", "fine", "pS"))
        pool = np.full(H.pool_doubles, np.nan)
        rng = np.random.default_rng(6)
        for k in H.keys:
            v = H.view(pool, k)
            v[:, ng:ng + BS, ng:ng + BS, ng:ng + BS] = rng.standard_normal((24, BS, BS, BS))
        full, linked = pool.copy(), pool.copy()
        run_program(full, *ent)
        run_program(linked, lpr, ljb, lcp, lpe, lbe, lce)
        pts = np.concatenate([pr, expand_jobs(jb)])
        idx = (pts["dst"][:, None] + np.arange(24)[None, :] * pts["dst_stride"][:, None]).ravel()
        assert not np.isnan(full[idx]).any()
        assert np.array_equal(linked[idx], full[idx])
        tiles = H.keys
        off = [H.slot_offset(t) for t in tiles]
        L = link_map(off, [W ** 3] * len(tiles), cp, H.pool_doubles)
        m = np.arange(len(tiles)) * W ** 3
        order = np.argsort(off)
        bi = np.asarray(order)[np.searchsorted(np.asarray(off)[order], cp["dst"], side="right") - 1]
        i = m[bi] + (cp["dst"] - np.asarray(off)[bi])
        assert np.array_equal(L[i, 0], cp["src"]) and np.all(L[i, 1] == W ** 3)
        rest = np.ones(len(L), bool)
        rest[i] = False
        blk = np.repeat(np.arange(len(tiles)), W ** 3)
        q = np.arange(len(L)) - m[blk]
        assert np.array_equal(L[rest, 0], (np.asarray(off)[blk] + q)[rest])


def test_root_sync_equals_single_grid_periodic():
    H = AMRHierarchy(nb_root=(2, 3, 2), BS=8, ng=3, levels=1, device="cpu")
    N = [n * 8 for n in H.nb_root]
    F = _rand((24, N[0] + 6, N[1] + 6, N[2] + 6), 4)
    G = O.periodic_sync(F, 3)
    pool = _rand(H.pool_doubles, 5)
    for k in H.keys:
        b = H.origin(k)
        H.view(pool, k)[:, 3:11, 3:11, 3:11] = F[:, 3 + b[0]:11 + b[0], 3 + b[1]:11 + b[1],
                                                3 + b[2]:11 + b[2]]
    run_program(pool, *H.program_entries())
    for k in H.keys:
        b = H.origin(k)
        assert np.array_equal(H.view(pool, k), G[:, b[0]:b[0] + 14, b[1]:b[1] + 14,
                                                b[2]:b[2] + 14])


def test_child_halo_is_prolongated_parent_interior_kept():
    BS, ng = 8, 3
    W = BS + 2 * ng
    hier = {(0, 0): (0, (0, 0, 0)), (1, 0): (1, (8, 0, 8))}
    parent = _rand((24, W, W, W), 6)
    blocks = {(0, 0): parent, (1, 0): np.full((24, W, W, W), -7.0)}
    out = OA.sync_all(blocks, hier, (1, 1, 1), BS, ng)
    want = OA.prolongate(out[(0, 0)], (1, 0, 1), BS, ng)
    halo = np.ones((W, W, W), bool)
    halo[ng:ng + BS, ng:ng + BS, ng:ng + BS] = False
    assert np.all(out[(1, 0)][:, ~halo] == -7.0)
    assert np.array_equal(out[(1, 0)][:, halo], want[:, halo])


def test_entry_builders_match_oracle_transfers():
    BS, ng = 8, 3
    W = BS + 2 * ng
    P = 24 * W ** 3
    pool = np.concatenate([_rand(P, 7), np.zeros(P)])
    for c in ((1, 0, 1), (0, 1, 1)):
        pts = np.stack(np.meshgrid(*[np.arange(W)] * 3, indexing="ij"), -1).reshape(-1, 3)
        apply_prolongs(pool, prolong_entries(pts, P, 0, c, BS, ng))
        child = pool[P:].reshape(24, W, W, W)
        assert np.array_equal(child, OA.prolongate(pool[:P].reshape(24, W, W, W), c, BS, ng))
        from amr_emul import apply_copies
        par = np.zeros(P)
        both = np.concatenate([par, pool[P:]])
        apply_copies(both, restrict_entries(0, P, c, BS, ng))
        want = OA.restrict_into_parent(np.zeros((24, W, W, W)), child, c, BS, ng)
        assert np.array_equal(both[:P].reshape(24, W, W, W), want)


def test_root_only_hierarchy_step_equals_single_grid():
    BS, ng, nb = 8, 3, (2, 2, 2)
    N = 16
    h = 1.0 / N
    g = Grid(N, N, N, h, h, h, ng=ng, xmin=-0.5, ymin=-0.5, zmin=-0.5)
    F = idata.punctures(g, [], [], amp=1e-3, seed=5).numpy()
    F = O.periodic_sync(F, ng)
    p = PhysicsParams()
    dt = 0.25 * h
    S16 = scalars(p, h, h, h, dt, 0.0)
    want = O.rk4_step(F, S16, 0.1, dt, order=4, ko_order=6, ko_sigma=0.1, ng=ng, periodic=True)
    hier, blocks = {}, {}
    for i, c in enumerate(itertools.product(range(2), repeat=3)):
        b = tuple(c[a] * BS for a in range(3))
        hier[(0, i)] = (0, b)
        blocks[(0, i)] = F[:, b[0]:b[0] + BS + 2 * ng, b[1]:b[1] + BS + 2 * ng,
                           b[2]:b[2] + BS + 2 * ng].copy()
    got = OA.rk4_step(blocks, hier, nb, BS, ng, h, S16, 0.1, dt, order=4, ko_sigma=0.1)
    for k, (lv, b) in hier.items():
        I = (slice(None),) + (slice(ng, ng + BS),) * 3
        w = want[:, ng + b[0]:ng + b[0] + BS, ng + b[1]:ng + b[1] + BS, ng + b[2]:ng + b[2] + BS]
        assert np.array_equal(got[k][I], w), k
