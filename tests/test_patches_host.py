"""CPU: the nested patch hierarchy of the batched block-AMR workload (bssn3d_hip.patches).

* The workload geometry: 640 blocks on three levels (h = 1/4, 1/8, 1/16), every interior extent
  from {8, 12, 16, 24, 32}, each level tiled exactly by its blocks, 2:1 nesting with a margin,
  both punctures of the 1:10 binary inside the finest box.
* The ghost fill program, applied with the host emulation of the device kernels
  (tests/amr_emul.py), on a small three-level hierarchy whose interiors hold a polynomial of
  degree 5 per axis: every ghost of the refined levels must hold the polynomial at its own
  coordinates (copies are exact, and the 6-point midpoint prolongation reproduces degree 5
  exactly), every root ghost the polynomial at its periodic image; every ghost point is written
  exactly once and nothing else is written. A wrong owner, a shifted fine index, a wrong parent
  or a missed ghost shows as an O(1) error.
"""
from __future__ import annotations

import numpy as np
import pytest

from amr_emul import run_program

from bssn3d_hip.patches import SIZES, config_hierarchy

NF = 24


def test_workload_geometry():
    H, pos = config_hierarchy(seed=0, device="cpu")
    assert len(H) == 640
    assert [b.nblocks for b in H.boxes] == [256, 192, 192]
    assert [b.h for b in H.boxes] == [0.25, 0.125, 0.0625]
    for blk in H.blocks:
        assert all(n in SIZES for n in blk.n)
    for L, box in enumerate(H.boxes):
        own = H.owner_map(L)                      # tiles exactly: every node owned once
        ids = np.flatnonzero(np.asarray(H.level) == L)
        assert set(np.unique(own)) == set(ids.tolist())
        vol = sum(int(np.prod(H.blocks[b].n)) for b in ids)
        assert vol == int(np.prod(box.extent))
    fine = H.boxes[-1]
    for q in pos:
        n = [(q[a] - H.x0[a]) / fine.h for a in range(3)]
        assert all(fine.origin[a] + 8 <= n[a] <= fine.origin[a] + fine.extent[a] - 8
                   for a in range(3))
    # blocks sit at the physical position of their lattice nodes
    for b in (0, 300, 639):
        L = H.level[b]
        assert np.allclose(H.blocks[b].lo, [H.x0[a] + H.node0[b][a] * H.boxes[L].h
                                            for a in range(3)], rtol=0, atol=1e-12)


def _poly(x, y, z):
    px = 0.3 + 0.2 * x - 0.05 * x ** 2 + 0.01 * x ** 3 - 0.002 * x ** 4 + 0.0003 * x ** 5
    py = 1.1 - 0.1 * y + 0.03 * y ** 2 - 0.004 * y ** 3 + 0.0005 * y ** 4 - 0.00002 * y ** 5
    pz = -0.7 + 0.15 * z + 0.02 * z ** 2 - 0.003 * z ** 3 + 0.0001 * z ** 4 + 0.00004 * z ** 5
    return px * py * pz


def test_program_reproduces_polynomials():
    H, _ = config_hierarchy(seed=3, counts=((4, 4, 4), (3, 3, 3), (2, 2, 2)), sizes=(8, 12),
                            device="cpu", hold_punctures=False)
    lay = H.layout
    ng = H.ng
    pool = np.full(lay.pool_doubles, np.nan)
    root = H.boxes[0]
    period = [root.extent[a] * root.h for a in range(3)]
    want = {}
    for b, blk in enumerate(H.blocks):
        g = blk.grid(ng)
        X, Y, Z = np.meshgrid(g.x, g.y, g.z, indexing="ij")
        if H.level[b] == 0:                       # periodic images of the root box
            X = H.x0[0] + np.mod(X - H.x0[0], period[0])
            Y = H.x0[1] + np.mod(Y - H.x0[1], period[1])
            Z = H.x0[2] + np.mod(Z - H.x0[2], period[2])
        f = _poly(X, Y, Z)
        want[b] = f
        v = lay.view(pool, b)
        for k in range(NF):
            v[k][ng:-ng, ng:-ng, ng:-ng] = (k + 1) * f[ng:-ng, ng:-ng, ng:-ng]
    pr, jb, cp, pe, be, ce = H.program_entries()
    dst = np.concatenate([pr["dst"], cp["dst"]])
    assert len(np.unique(dst)) == len(dst)        # every ghost written once
    nghost = sum(int(np.prod(lay.S[b])) - int(np.prod(blk.n)) for b, blk in enumerate(H.blocks))
    assert len(dst) == nghost
    assert pe[0] == 0 and len(pe) == 3 and ce[0] > 0
    run_program(pool, pr, jb, cp, pe, be, ce)
    assert not np.isnan(pool[np.concatenate([np.arange(lay.off[b], lay.off[b] + NF * lay.P[b])
                                             for b in range(len(H))])]).any()
    scale = max(np.abs(f).max() for f in want.values())
    for b in range(len(H)):
        v = lay.view(pool, b)
        for k in (0, 7, 23):
            err = np.abs(v[k] - (k + 1) * want[b]).max() / ((k + 1) * scale)
            tol = 0.0 if H.level[b] == 0 else 1e-13
            assert err <= tol, (b, H.level[b], k, err)


def test_linked_program_and_link_map():
    """A linked step reads a copied ghost at its source and refills only what the
    prolongations need. On a small three-level hierarchy whose interiors hold seeded values:
    the link map sends every copied ghost to its source and every other point to itself; the
    linked program (all prolongations, only the copies into parent ghosts that their taps read)
    gives every prolongated ghost exactly the value the full program gives it; and the copies
    it leaves out are not read by any prolongation (filling them with NaN changes nothing)."""
    H, _ = config_hierarchy(seed=5, counts=((4, 4, 4), (3, 3, 3), (2, 2, 2)), sizes=(8, 12),
                            device="cpu", hold_punctures=False)
    lay = H.layout
    ng = H.ng
    pr, jb, cp, pe, be, ce = H.program_entries()
    lpr, ljb, lcp, lpe, lbe, lce = H.linked_program_entries()
    assert len(lpr) == len(pr) and 0 < len(lcp) < len(cp)
    assert lce[-1] == len(lcp) and all(a <= b for a, b in zip(lce, lce[1:]))
    # link map
    L = H.link_map(cp)
    m = H.link_offsets()
    off = np.asarray(lay.off)
    assert L.shape == (int(np.sum(lay.P)), 2)
    b = np.searchsorted(off, cp["dst"], side="right") - 1
    i = m[b] + (cp["dst"] - off[b])
    assert np.array_equal(L[i, 0], cp["src"]) and np.array_equal(L[i, 1], cp["src_stride"])
    rest = np.ones(len(L), bool)
    rest[i] = False
    blk = np.repeat(np.arange(len(lay.P)), lay.P)
    q = np.arange(len(L)) - m[blk]
    assert np.array_equal(L[rest, 0], (off[blk] + q)[rest])
    assert np.array_equal(L[rest, 1], np.asarray(lay.P)[blk][rest])
    # the linked program reproduces every prolongated ghost
    rng = np.random.default_rng(2)
    base = np.full(lay.pool_doubles, np.nan)
    for k, blk_ in enumerate(H.blocks):
        v = lay.view(base, k)
        v[:, ng:-ng, ng:-ng, ng:-ng] = rng.standard_normal(v[:, ng:-ng, ng:-ng, ng:-ng].shape)
    full = base.copy()
    run_program(full, pr, jb, cp, pe, be, ce)
    linked = base.copy()
    run_program(linked, lpr, ljb, lcp, lpe, lbe, lce)
    for e in (pr["dst"],):
        idx = (e[:, None] + np.arange(NF)[None, :] * pr["dst_stride"][:, None]).ravel()
        assert not np.isnan(full[idx]).any()
        assert np.array_equal(linked[idx], full[idx])
    # copies left out hold NaN in the linked pool and no prolongated ghost became NaN
    skipped = np.setdiff1d(cp["dst"], lcp["dst"])
    assert np.isnan(linked[skipped]).all()


def test_prolongation_tiles_cover_the_points():
    """The tile form of the program's prolongations (prolong_tiles) stands for exactly its
    point entries (tile_prolong_points inverts it); tiles are aligned 4x4x4 blocks of one
    parent; the C-ABI's bounds predicates accept every tile and tile point on the hierarchy's
    pool and reject misaligned, oversized and out-of-pool ones."""
    from bssn3d_hip import _lib
    from bssn3d_hip.amr import prolong_tiles, tile_prolong_points
    H, _ = config_hierarchy(seed=0, device="cpu")
    pr = H.program_entries()[0]
    tiles, pts = prolong_tiles(pr)
    assert 0 < len(tiles) < len(pr) and tiles["count"].sum() == len(pr)
    assert np.all(tiles["origin"] % 4 == 0) and np.all((pts["local"] >= 0) & (pts["local"] < 64))
    back = tile_prolong_points(tiles, pts)
    key = lambda e: np.lexsort((e["fine"][:, 2], e["fine"][:, 1], e["fine"][:, 0], e["dst"]))
    a, b = pr[key(pr)], back[key(back)]
    for k in ("dst", "parent", "dst_stride", "fine", "pS"):
        assert np.array_equal(a[k], b[k]), k
    lib = _lib.load()
    pool = H.layout.pool_doubles
    count = lambda kind, arr: lib.bssn_count_skipped_entries(
        kind, np.ascontiguousarray(arr).ctypes.data, len(arr), pool)
    assert count(3, tiles) == 0 and count(4, pts) == 0
    bad = tiles[:4].copy()
    bad["origin"][1, 0] += 2          # misaligned
    bad["count"][2] = 5000            # too many points
    bad["parent"][3] = pool           # parent past the pool
    assert count(3, bad) == 3
    badq = pts[:3].copy()
    badq["local"][0] = 64
    badq["dst"][1] = pool
    assert count(4, badq) == 2
    assert lib.bssn_count_skipped_entries(6, None, 0, pool) == -1
