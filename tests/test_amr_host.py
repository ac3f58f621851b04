"""CPU: regrid decisions and topology bookkeeping of the block hierarchy (bssn3d_hip.amr), the
reference's regrid unit cases (2D/tests/unit/test_amr_regrid.py) in 3D."""
from __future__ import annotations

import numpy as np
import pytest

from bssn3d_hip.amr import (COARSEN, KEEP, REFINE, AMRHierarchy, AMRTopology, compute_flags,
                            enforce_nesting_buffer)


def topo_root(nb=(4, 4, 4), BS=8, levels=3, max_blocks=128):
    T = AMRTopology(levels, BS, [int(np.prod(nb))] + [max_blocks] * (levels - 1), max_blocks)
    for i in range(nb[0]):
        for j in range(nb[1]):
            for k in range(nb[2]):
                T.add_block(0, (i * nb[1] + j) * nb[2] + k, (i * BS, j * BS, k * BS))
    return T


def _ind(T, level_vals):
    out = [np.zeros(T.max_blocks) for _ in range(T.levels)]
    for (L, s), v in level_vals.items():
        out[L][s] = v
    return out


def test_streak_must_reach_K_to_refine():
    T = topo_root((1, 1, 1))
    for cycle in range(3):
        fl = compute_flags(_ind(T, {(0, 0): 5.0}), T, refine_threshold=1.0,
                           coarsen_threshold=0.1, hysteresis_K=3)
        assert fl[0, 0] == (REFINE if cycle == 2 else KEEP)


def test_decay_resets_streak():
    T = topo_root((1, 1, 1))
    kw = dict(refine_threshold=1.0, coarsen_threshold=0.1, hysteresis_K=3)
    for _ in range(2):
        compute_flags(_ind(T, {(0, 0): 5.0}), T, **kw)
    assert T.streaks[0, 0] == 2
    compute_flags(_ind(T, {(0, 0): 0.5}), T, **kw)
    assert T.streaks[0, 0] == 1
    fl = compute_flags(_ind(T, {(0, 0): 5.0}), T, **kw)
    assert fl[0, 0] == KEEP and T.streaks[0, 0] == 2


def test_root_cannot_coarsen_and_depth_cap():
    T = topo_root((1, 1, 1))
    T.add_block(1, 0, (0, 0, 0), parent=(0, 0))
    for _ in range(2):
        fl = compute_flags(_ind(T, {}), T, refine_threshold=1.0, coarsen_threshold=0.5,
                           hysteresis_K=2)
    assert fl[0, 0] == KEEP and fl[1, 0] == COARSEN
    T2 = topo_root((1, 1, 1))
    T2.add_block(1, 0, (0, 0, 0), parent=(0, 0))
    for _ in range(2):
        fl = compute_flags(_ind(T2, {(0, 0): 9.0, (1, 0): 9.0}), T2, refine_threshold=1.0,
                           coarsen_threshold=0.5, hysteresis_K=2, max_level=1)
    assert fl[0, 0] == REFINE and fl[1, 0] == KEEP       # level 1 may not create level 2
    with pytest.raises(ValueError):
        compute_flags(_ind(T2, {}), T2, refine_threshold=0.5, coarsen_threshold=0.5)


def test_inactive_slots_never_flag():
    T = topo_root((1, 1, 1))
    for _ in range(4):
        fl = compute_flags(_ind(T, {(0, 5): 9.0}), T, refine_threshold=1.0,
                           coarsen_threshold=0.1, hysteresis_K=2)
    assert fl[0, 5] == KEEP


def test_buffer_dilates_to_chebyshev_neighbours():
    T = topo_root((4, 4, 4))
    fl = np.zeros((T.levels, T.max_blocks), np.int32)
    c = (1 * 4 + 1) * 4 + 1                                  # block (1, 1, 1)
    fl[0, c] = REFINE
    out = enforce_nesting_buffer(fl, T, n_buffer=1)
    got = {s for s in range(T.max_blocks) if out[0, s] == REFINE}
    want = {(i * 4 + j) * 4 + k for i in range(3) for j in range(3) for k in range(3)}
    assert got == want
    out2 = enforce_nesting_buffer(fl, T, n_buffer=0)
    assert {s for s in range(T.max_blocks) if out2[0, s] == REFINE} == {c}


def test_no_orphan_coarsening():
    T = topo_root((1, 1, 1))
    T.add_block(1, 0, (0, 0, 0), parent=(0, 0))
    T.add_block(2, 0, (0, 0, 0), parent=(1, 0))
    fl = np.zeros((T.levels, T.max_blocks), np.int32)
    fl[1, 0] = COARSEN
    fl[2, 0] = COARSEN
    out = enforce_nesting_buffer(fl, T)
    assert out[1, 0] == KEEP and out[2, 0] == COARSEN      # finest first, one level per pass
    T.remove_block(2, 0)
    assert enforce_nesting_buffer(fl, T)[1, 0] == COARSEN


def test_topology_corners_neighbours_nesting():
    T = topo_root((2, 2, 2), BS=8)
    for i, c in enumerate(((0, 0, 0), (1, 0, 0), (0, 1, 1))):
        T.add_block(1, i, tuple(c[a] * 8 for a in range(3)), parent=(0, 0))
    assert [T.child_corner((1, i)) for i in range(3)] == [(0, 0, 0), (1, 0, 0), (0, 1, 1)]
    T.rebuild_neighbors()
    assert T.neighbors[(1, 0, 1)] == (1, 1) and T.neighbors[(1, 1, 0)] == (1, 0)
    assert T.neighbors[(1, 0, 3)] is None and T.neighbors[(1, 2, 0)] is None
    # root neighbours do not wrap: block (0,0,0) has no -x neighbour
    assert T.neighbors[(0, 0, 0)] is None and T.neighbors[(0, 0, 1)] is not None
    assert T.check_proper_nesting() == []
    T.add_block(1, 3, (40, 0, 0), parent=(0, 0))             # outside its parent
    assert T.check_proper_nesting() and T.check_proper_nesting()[0][0:2] == (1, 3)
    T.remove_block(1, 1)
    assert (1, 1) not in T.children[(0, 0)] and T.neighbors.get((1, 0, 1)) in (None, (1, 1))


def test_hierarchy_layout_and_growth_bookkeeping():
    H = AMRHierarchy(nb_root=(2, 2, 2), BS=8, ng=3, levels=3, caps=[8, 16, 16], max_blocks=64,
                     device="cpu")
    assert H.W == 14 and H.h0 == 1.0 / 16
    slot = (24 * 14 ** 3 + 63) // 64 * 64
    assert H.slot_offset(1, 0) == 8 * slot and H.slot_offset(2, 3) == 24 * slot + 3 * slot
    assert H.pool_doubles == 40 * slot
    with pytest.raises(IndexError):
        H.slot_offset(1, 16)
    b = H.block((0, 7))
    assert b.lo == (0.0, 0.0, 0.0) and b.h == (1.0 / 16,) * 3     # block (1,1,1) of [-.5,.5)
    assert H.spacing(2) == 1.0 / 64
    with pytest.raises(ValueError):
        AMRHierarchy(BS=6, device="cpu")
    with pytest.raises(ValueError):
        AMRHierarchy(ng=2, device="cpu")
