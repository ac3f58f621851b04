"""CPU: host logic of the batched block form -- pool layout, the halo copy map of a tiled
periodic box (against numpy's wrap padding of the whole box), the workload geometry and the
argument checks of the batch entry points (which fire before any device work)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from bssn3d_hip import _lib
from bssn3d_hip.batch import (Block, BlockLayout, HALO_DTYPE, apply_entries, config_blocks,
                              puncture_positions, tiling_entries)


def test_layout_offsets_aligned_and_disjoint():
    L = BlockLayout([Block((5, 3, 8), (1, 1, 1)), Block((3, 3, 3), (1, 1, 1))], ng=3, align=64)
    assert L.S == [(11, 9, 14), (9, 9, 9)]
    assert all(o % 64 == 0 for o in L.off)
    assert L.off[1] >= 24 * L.P[0] and L.pool_doubles >= L.off[1] + 24 * L.P[1]


@pytest.mark.parametrize("N,tiles,exts", [
    ((8, 6, 4), [(0, 0, 0), (5, 0, 0), (0, 3, 0), (5, 3, 0)],
     [(5, 3, 4), (3, 3, 4), (5, 3, 4), (3, 3, 4)]),
    ((16, 12, 8), [(0, 0, 0), (0, 5, 0), (9, 0, 0), (9, 7, 0), (0, 0, 5)],
     [(9, 5, 5), (9, 7, 8), (7, 7, 8), (7, 5, 8), (9, 5, 3)]),
])
def test_tiling_halo_equals_wrap_padding(N, tiles, exts):
    """Interiors copied from a global box, halo copies applied: every block's padded array must
    equal the periodic (wrap) padding of the whole box, ghosts included."""
    ng = 3
    L = BlockLayout([Block(e, (1, 1, 1)) for e in exts], ng)
    G = np.random.default_rng(1).standard_normal((24,) + N)
    Gp = np.pad(G, ((0, 0),) + ((ng, ng),) * 3, mode="wrap")
    pool = np.zeros(L.pool_doubles)
    for b, (t, e) in enumerate(zip(tiles, exts)):
        L.view(pool, b)[(slice(None),) + tuple(slice(ng, ng + e[a]) for a in range(3))] = \
            G[(slice(None),) + tuple(slice(t[a], t[a] + e[a]) for a in range(3))]
    E = tiling_entries(L, tiles, N)
    assert E.dtype == HALO_DTYPE
    # every ghost point exactly once; sources are interior points
    assert len(np.unique(E["dst"])) == len(E) == sum(P - np.prod(e) for P, e in zip(L.P, exts))
    assert not np.isin(E["src"], E["dst"]).any()
    apply_entries(pool, E)
    for b, (t, e) in enumerate(zip(tiles, exts)):
        want = Gp[(slice(None),) + tuple(slice(t[a], t[a] + e[a] + 2 * ng) for a in range(3))]
        assert np.array_equal(L.view(pool, b), want), b


def test_tiling_rejects_gaps_and_overlaps():
    L = BlockLayout([Block((4, 4, 4), (1, 1, 1)), Block((4, 4, 4), (1, 1, 1))], 3)
    with pytest.raises(ValueError, match="cover"):
        tiling_entries(L, [(0, 0, 0), (4, 0, 0)], (12, 4, 4))
    with pytest.raises(ValueError, match="overlap"):
        tiling_entries(L, [(0, 0, 0), (2, 0, 0)], (6, 4, 4))


def test_workload_geometry():
    """640 cubic blocks, extents uniform over {8,12,16,24,32}, three spacings; about 6.8 M
    interior points expected (E[n^3] = 10585.6 per block)."""
    blocks = config_blocks(seed=0)
    assert len(blocks) == 640
    assert {b.n[0] for b in blocks} <= {8, 12, 16, 24, 32}
    assert {b.h[0] for b in blocks} == {0.25, 0.125, 0.0625}
    pts = sum(int(np.prod(b.n)) for b in blocks)
    assert 5.5e6 < pts < 8.0e6
    for b in blocks:     # inside [-8, 8)^3 and on its level's grid
        for a in range(3):
            assert -8.0 <= b.lo[a] and b.lo[a] + b.n[a] * b.h[a] <= 8.0 + 1e-12
            assert abs((b.lo[a] + 8.0) / b.h[a] - round((b.lo[a] + 8.0) / b.h[a])) < 1e-9
    p1, p2 = puncture_positions(0)
    assert abs((p1[0] - p2[0]) - 4.0) < 1e-12
    # centre of mass = the seeded off-grid shift, within [h/4, 3h/4) of the finest spacing
    com = (0.909 * np.asarray(p1) + 0.0909 * np.asarray(p2)) / (0.909 + 0.0909)
    assert np.all(com >= 0.25 * 0.0625 - 1e-12) and np.all(com < 0.75 * 0.0625)


def test_batch_argument_errors_before_device_work():
    lib = _lib.load()
    S = (ctypes.c_int * 6)(11, 11, 11, 9, 9, 9)
    assert lib.bssn_batch_plan_bytes(2, S, 3) > 0
    assert lib.bssn_batch_plan_bytes(2, (ctypes.c_int * 6)(11, 11, 11, 6, 9, 9), 3) == 0
    h = (ctypes.c_double * 6)(*([0.1] * 6))
    fake = ctypes.c_void_p(256)
    out = ctypes.c_void_p()
    big = 10 ** 9
    # overlapping blocks
    off = (ctypes.c_longlong * 2)(0, 100)
    assert lib.bssn_batch_create(ctypes.byref(out), fake, big, 2, off, S, h, 3, 10 ** 6, None) == 1
    assert b"overlap" in lib.bssn_last_error()
    # a block past the pool end
    off = (ctypes.c_longlong * 2)(0, 24 * 1331)
    assert lib.bssn_batch_create(ctypes.byref(out), fake, big, 2, off, S, h, 3, 24 * 1331, None) == 1
    assert b"outside" in lib.bssn_last_error()
    # bad spacing, too small a plan, interior smaller than ng
    off = (ctypes.c_longlong * 2)(0, 24 * 1331)
    bad_h = (ctypes.c_double * 6)(0.1, 0.1, 0.0, 0.1, 0.1, 0.1)
    pool = 24 * (1331 + 729)
    assert lib.bssn_batch_create(ctypes.byref(out), fake, big, 2, off, S, bad_h, 3, pool, None) == 1
    assert lib.bssn_batch_create(ctypes.byref(out), fake, 8, 2, off, S, h, 3, pool, None) == 1
    assert b"plan" in lib.bssn_last_error()
    assert lib.bssn_batch_create(ctypes.byref(out), fake, big, 2, off, S, h, 4, pool, None) == 1
    # null handles
    assert lib.bssn_batch_set_launch(None, 0) == 1
    assert lib.bssn_batch_points(None, 1) == -1
    assert lib.bssn_batch_rhs_f64(None, fake, _lib.dbl([1.0] * 16), fake, 4, 0.0, None) == 1
    assert lib.bssn_batch_halo_f64(fake, 100, None, 5, None) == 1
    assert lib.bssn_batch_halo_f64(fake, 0, fake, 5, None) == 1          # empty pool


def test_entry_bounds_predicates():
    """bssn_count_skipped_entries runs, on the host, the test every ghost / transfer kernel
    applies to its entry before touching the pool: an entry that reaches outside the pool (or a
    block job beyond the LDS passes' 24-node limit, or with an octant index other than 0/1) is
    skipped whole."""
    from bssn3d_hip.amr import BLOCK_DTYPE, PROLONG_DTYPE, block_job, prolong_entries
    from bssn3d_hip.batch import HALO_DTYPE
    lib = _lib.load()
    count = lambda kind, e, pool: lib.bssn_count_skipped_entries(kind, e.ctypes.data, len(e), pool)
    pool = 24 * 1000
    h = np.zeros(6, dtype=HALO_DTYPE)
    h["dst_stride"] = h["src_stride"] = 1000
    h["dst"] = [0, 999, 1000, 5, -1, 5]
    h["src"] = [999, 0, 5, 5, 5, 5]
    h["src_stride"][5] = 1044                # field 23 of the source at 5 + 23 * 1044 > pool
    assert count(0, h, pool) == 3            # entries 2, 4 and 5
    assert count(0, h[:2], pool) == 0
    # prolongation points: the parent's whole padded block must lie in the pool
    BS, ng = 8, 3
    W = BS + 2 * ng
    P = 24 * W ** 3
    pts = np.array([[0, 0, 0], [W - 1, W - 1, W - 1]])
    good = prolong_entries(pts, P, 0, (1, 0, 1), BS, ng)
    assert count(1, good, 2 * P) == 0
    assert count(1, good, 2 * P - 1) == 1    # the last destination point is the pool's end
    bad = good.copy()
    bad["parent"] = P + 1                    # parent block runs past the pool
    assert count(1, bad, 2 * P) == 2
    bad = good.copy()
    bad["pS"][0] = (0, 0, 0)
    assert count(1, bad, 2 * P) == 1
    # whole-block jobs
    jobs = np.concatenate([block_job(P, 0, (1, 1, 0), BS, ng, 0),
                           block_job(P, 0, (2, 0, 0), BS, ng, 0),      # octant index 2
                           block_job(P, 0, (0, 0, 0), 20, 3, 0),       # W = 26 > 24
                           block_job(P + 1, 0, (0, 0, 0), BS, ng, 0)])  # child past the end
    assert jobs.dtype == BLOCK_DTYPE and good.dtype == PROLONG_DTYPE
    assert count(2, jobs, 2 * P) == 3 and count(2, jobs[:1], 2 * P) == 0
    assert lib.bssn_count_skipped_entries(6, None, 0, pool) == -1
    # link map entries (pool offset of field 0, field stride): all 24 fields inside the pool
    links = np.array([[0, 10], [pool - 23 * 10 - 1, 10], [pool - 23 * 10, 10], [-1, 10],
                      [5, -1]], dtype=np.int32)
    assert lib.bssn_count_skipped_entries(5, links.ctypes.data, len(links), pool) == 3
    assert lib.bssn_count_skipped_entries(0, None, 4, pool) == -1


def test_pick_launch_keeps_the_first_shape_within_the_noise():
    """A launch shape replaces the first candidate only when it is faster by more than the
    margin; among several such, the fastest wins."""
    from bssn3d_hip.batch import LAUNCH_MARGIN, pick_launch
    assert LAUNCH_MARGIN == 0.005
    assert pick_launch({0: 10.0, 1024: 9.96, 2048: 10.3}) == 0          # 0.4 % faster: noise
    assert pick_launch({0: 10.0, 1024: 9.9, 2048: 9.8}) == 2048          # 2 % faster
    assert pick_launch({1024: 10.0, 0: 10.01}) == 1024                   # first in timing order
    assert pick_launch({0: 10.0, 1024: 9.9}, margin=0.02) == 0
