"""CPU: regrid decisions and tile bookkeeping of the block hierarchy (bssn3d_hip.amr), covering
the behaviours the reference's regrid unit cases check (2D/tests/unit/test_amr_regrid.py), in
3D: a refinement needs `patience` consecutive votes, a neutral indicator walks the counter back,
roots never coarsen, the deepest level never refines, coarsening waits for the children, and
refinement spreads to the Chebyshev neighbourhood."""
from __future__ import annotations

import numpy as np
import pytest

from bssn3d_hip.amr import (COARSEN, KEEP, REFINE, AMRHierarchy, Octree, RegridPolicy,
                            SlotArena, child_tiles, face_step, octant, parent_tile)


def tree(roots=(4, 4, 4), depth=3, ceiling=128):
    return Octree(roots, depth, [int(np.prod(roots))] + [ceiling] * (depth - 1), ceiling)


ROOT = (0, 0, 0, 0)


def test_refine_needs_patience_votes():
    T = tree((1, 1, 1))
    pol = RegridPolicy(refine_above=1.0, coarsen_below=0.1, patience=3)
    got = [pol.decide(T, {ROOT: 5.0})[ROOT] for _ in range(3)]
    assert got == [KEEP, KEEP, REFINE]


def test_neutral_indicator_walks_counter_back():
    T = tree((1, 1, 1))
    pol = RegridPolicy(refine_above=1.0, coarsen_below=0.1, patience=3)
    for _ in range(2):
        pol.decide(T, {ROOT: 5.0})
    assert T.votes[ROOT] == 2
    pol.decide(T, {ROOT: 0.5})                      # between the thresholds
    assert T.votes[ROOT] == 1
    assert pol.decide(T, {ROOT: 5.0})[ROOT] == KEEP and T.votes[ROOT] == 2
    for _ in range(4):                               # counters saturate at +-patience
        pol.decide(T, {ROOT: 5.0})
    assert T.votes[ROOT] == 3


def test_roots_keep_and_depth_cap():
    T = tree((1, 1, 1))
    kid = (1, 0, 0, 0)
    T.plant(kid)
    pol = RegridPolicy(refine_above=1.0, coarsen_below=0.5, patience=2)
    for _ in range(2):
        acts = pol.vote(T, {})
    assert acts[ROOT] == KEEP and acts[kid] == COARSEN
    T2 = tree((1, 1, 1))
    T2.plant(kid)
    pol2 = RegridPolicy(refine_above=1.0, coarsen_below=0.5, patience=2, max_level=1)
    for _ in range(2):
        acts = pol2.vote(T2, {ROOT: 9.0, kid: 9.0})
    assert acts[ROOT] == REFINE and acts[kid] == KEEP   # level 1 may not create level 2
    with pytest.raises(ValueError):
        RegridPolicy(refine_above=0.5, coarsen_below=0.5)


def test_only_existing_tiles_get_actions():
    T = tree((1, 1, 1))
    pol = RegridPolicy(refine_above=1.0, coarsen_below=0.1, patience=2)
    for _ in range(4):
        acts = pol.decide(T, {(0, 0, 0, 5): 9.0})
    assert set(acts) == {ROOT} and acts[ROOT] == KEEP     # a root with a low score is kept


def test_refinement_spreads_to_chebyshev_neighbourhood():
    T = tree((4, 4, 4))
    pol = RegridPolicy(refine_above=1.0, coarsen_below=0.1, buffer=1)
    acts = {t: KEEP for t in T.all_tiles()}
    acts[(0, 1, 1, 1)] = REFINE
    got = {t for t, a in pol.settle(T, acts).items() if a == REFINE}
    assert got == {(0, i, j, k) for i in range(3) for j in range(3) for k in range(3)}
    pol0 = RegridPolicy(refine_above=1.0, coarsen_below=0.1, buffer=0)
    assert {t for t, a in pol0.settle(T, acts).items() if a == REFINE} == {(0, 1, 1, 1)}


def test_coarsening_waits_for_children():
    T = tree((1, 1, 1))
    a, b = (1, 0, 0, 0), (2, 0, 0, 0)
    T.plant(a)
    T.plant(b)
    pol = RegridPolicy(refine_above=1.0, coarsen_below=0.1, buffer=0)
    acts = {ROOT: KEEP, a: COARSEN, b: COARSEN}
    out = pol.settle(T, acts)
    assert out[a] == KEEP and out[b] == COARSEN      # finest first, one level per pass
    T.prune(b)
    assert pol.settle(T, {ROOT: KEEP, a: COARSEN})[a] == COARSEN


def test_tile_arithmetic_and_nesting():
    T = tree((2, 2, 2))
    kids = [(1, 0, 0, 0), (1, 1, 0, 0), (1, 0, 1, 1)]
    for c in kids:
        T.plant(c)
    assert [octant(c) for c in kids] == [(0, 0, 0), (1, 0, 0), (0, 1, 1)]
    assert all(parent_tile(c) == (0, 0, 0, 0) for c in kids)
    assert set(child_tiles((0, 1, 0, 1))) == {(1, 2 + a, b, 2 + c) for a in (0, 1)
                                               for b in (0, 1) for c in (0, 1)}
    assert T.neighbour((1, 0, 0, 0), 1) == (1, 1, 0, 0)
    assert T.neighbour((1, 1, 0, 0), 0) == (1, 0, 0, 0)
    assert T.neighbour((1, 0, 0, 0), 3) is None and T.neighbour((1, 0, 1, 1), 0) is None
    assert face_step((0, 0, 0, 0), 0) == (0, -1, 0, 0)
    assert T.neighbour((0, 0, 0, 0), 0) is None          # no wrap in the tile lattice
    assert T.neighbour((0, 0, 0, 0), 1) == (0, 1, 0, 0)
    assert T.nesting_errors() == []
    T.plant((2, 0, 0, 0))                                # parent (1, 0, 0, 0) exists
    T.plant((2, 6, 0, 0))                                # parent (1, 3, 0, 0) does not
    assert T.nesting_errors() == [((2, 6, 0, 0), "parent missing")]
    T.prune((1, 1, 0, 0))
    assert T.kids((0, 0, 0, 0)) == [(1, 0, 0, 0), (1, 0, 1, 1)]
    with pytest.raises(ValueError):
        T.plant((1, 0, 0, 0))                            # exists
    with pytest.raises(ValueError):
        T.plant((0, 2, 0, 0))                            # outside the root lattice


def test_slot_arena_lowest_free_and_widening():
    A = SlotArena(3)
    assert [A.take((1, i, 0, 0)) for i in range(3)] == [0, 1, 2] and A.full()
    A.release(1)
    assert A.take((1, 9, 0, 0)) == 1
    with pytest.raises(RuntimeError):
        A.take((1, 8, 0, 0))
    A.widen(6)
    assert A.capacity == 6 and A.take((1, 8, 0, 0)) == 3
    T = Octree((1, 1, 1), 2, [1, 4], 16)
    assert T.widen(1) == (4, 8) and T.widen(1) == (8, 16) and T.widenings[-1] == (1, 8, 16)
    with pytest.raises(RuntimeError):
        T.widen(1)


def test_hierarchy_layout_bookkeeping():
    H = AMRHierarchy(nb_root=(2, 2, 2), BS=8, ng=3, levels=3, caps=[8, 16, 16], max_blocks=64,
                     device="cpu")
    assert H.W == 14 and H.h0 == 1.0 / 16
    slot = (24 * 14 ** 3 + 63) // 64 * 64
    H.tree.plant((1, 0, 0, 0))
    H.tree.plant((1, 1, 0, 0))
    assert H.slot_offset((1, 1, 0, 0)) == 8 * slot + 1 * slot
    assert H.pool_doubles == 40 * slot
    assert H.keys[:2] == [(0, 0, 0, 0), (0, 0, 0, 1)] and H.keys[-1] == (1, 1, 0, 0)
    b = H.block((0, 1, 1, 1))
    assert b.lo == (0.0, 0.0, 0.0) and b.h == (1.0 / 16,) * 3     # block (1,1,1) of [-.5,.5)
    assert H.spacing(2) == 1.0 / 64
    with pytest.raises(ValueError):
        AMRHierarchy(BS=6, device="cpu")
    with pytest.raises(ValueError):
        AMRHierarchy(ng=2, device="cpu")
