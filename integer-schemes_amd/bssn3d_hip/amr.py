"""Block-structured AMR for the BSSN state on the device: a live octree of equal blocks.

The hierarchy is an octree forest over a periodic root tiling. A block is named by its tile
``(level, i, j, k)``: block (i, j, k) of its level's lattice, whose first interior node sits at
level node index BS (i, j, k). Everything else is arithmetic on the tile -- the parent is
(level - 1, i >> 1, j >> 1, k >> 1), the octant inside it (i & 1, j & 1, k & 1), the face
neighbours are the tiles one lattice step away -- so the host keeps only which tiles exist,
the pool slot each one occupies, and a hysteresis counter per tile.

What a hierarchy holds and how its ghosts are filled follows the reference's node-centred
block AMR (it has none for BSSN; its block AMR serves the MCS and wave systems):
  * 2:1 transfers: fine[2k] = coarse[k], fine[2k+1] = 6-point Lagrange midpoint, tensor
    product x, y, z; restriction = injection (3D/src/multipatch/amr/kernels.py:53-140);
  * the hierarchy sync of its shared-dt step: root level periodic, then each finer level
    coarse to fine from its parent's prolongation, with same-level face slabs over the
    interior extent taken from face neighbours (2D/src/mcs2d/amr/evolve.py:719-735,
    3D/src/multipatch/amr/sync.py:84-122);
  * regrid: an indicator per block with a hysteresis of K consecutive votes, no coarsening of
    a block that still has children, refinement spread to nearby blocks of the same level,
    coarsening finest level first and refinement coarsest first
    (2D/src/mcs2d/amr/regrid.py:113-538), in 3D with 8 octants.

Device side: every block lives in a fixed slot of one float64 pool (level L's slots follow level
L-1's); the existing tiles form a BlockBatch (its plan lists only them); the ghost sync is a
ghost fill program (bssn_ghost_fill_f64: one phase per level) run after every RK stage by
bssn_batch_rk4_step_prog_f64. Prolongations run as whole-block jobs (bssn_prolong_blocks_f64,
padded blocks up to 24 per axis) or point lists (bssn_prolong_f64); restrictions as copy lists
(bssn_batch_halo_f64); the indicator is bssn_batch_indicator_f64.
"""
from __future__ import annotations

import ctypes
import itertools
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .batch import HALO_DTYPE, Block, BlockBatch, _device, tune_step_launch
from .grid import Grid
from .rhs import stream_handle
from .state import NUM_VARS, PhysicsParams

REFINE, KEEP, COARSEN = 1, 0, -1
OCTANTS = tuple(itertools.product((0, 1), repeat=3))

PROLONG_DTYPE = np.dtype([("dst", "<i8"), ("parent", "<i8"), ("dst_stride", "<i8"),
                          ("fine", "<i4", (3,)), ("pS", "<i4", (3,))])
assert PROLONG_DTYPE.itemsize == 48
# bssn_prolong_block: a whole-block job (padded extent W = BS + 2 ng <= BLOCK_WMAX)
BLOCK_DTYPE = np.dtype([("dst", "<i8"), ("parent", "<i8"), ("corner", "<i4", (3,)),
                        ("skip", "<i4"), ("BS", "<i4"), ("ng", "<i4")])
assert BLOCK_DTYPE.itemsize == 40
BLOCK_WMAX = 24
SKIP_INTERIOR_WRITE = 64
# bssn_prolong_tile / bssn_prolong_tile_point: prolongation points grouped by 4x4x4 fine tile
TILE_DTYPE = np.dtype([("parent", "<i8"), ("first", "<i8"), ("pS", "<i4", (3,)),
                       ("origin", "<i4", (3,)), ("count", "<i4"), ("pad", "<i4")])
assert TILE_DTYPE.itemsize == 48
TILE_POINT_DTYPE = np.dtype([("dst", "<i8"), ("dst_stride", "<i4"), ("local", "<i4")])
assert TILE_POINT_DTYPE.itemsize == 16

Tile = Tuple[int, int, int, int]


# --- tiles -------------------------------------------------------------------------------------

def parent_tile(t: Tile) -> Tile:
    return (t[0] - 1, t[1] >> 1, t[2] >> 1, t[3] >> 1)


def octant(t: Tile) -> Tuple[int, int, int]:
    """Which of its parent's eight octants tile t covers."""
    return (t[1] & 1, t[2] & 1, t[3] & 1)


def child_tiles(t: Tile) -> List[Tile]:
    L, i, j, k = t
    return [(L + 1, 2 * i + a, 2 * j + b, 2 * k + c) for a, b, c in OCTANTS]


def face_step(t: Tile, face: int) -> Tile:
    """The tile across face ``face`` (0..5 = -x, +x, -y, +y, -z, +z), present or not."""
    axis, sign = face >> 1, 1 if face & 1 else -1
    c = list(t)
    c[1 + axis] += sign
    return tuple(c)


class SlotArena:
    """The pool slots of one level: a tile takes the lowest free slot."""

    def __init__(self, capacity: int):
        self.owner: List[Optional[Tile]] = [None] * int(capacity)

    @property
    def capacity(self) -> int:
        return len(self.owner)

    @property
    def used(self) -> int:
        return sum(o is not None for o in self.owner)

    def full(self) -> bool:
        return all(o is not None for o in self.owner)

    def take(self, tile: Tile) -> int:
        for s, o in enumerate(self.owner):
            if o is None:
                self.owner[s] = tile
                return s
        raise RuntimeError(f"no free slot on level {tile[0]} (capacity {self.capacity})")

    def release(self, slot: int) -> None:
        self.owner[slot] = None

    def widen(self, capacity: int) -> None:
        if capacity < self.capacity:
            raise ValueError("a level's capacity never shrinks")
        self.owner.extend([None] * (capacity - self.capacity))


class Octree:
    """Which tiles exist, their slots, and their regrid vote counters.

    roots: lattice extent of level 0 (every root tile exists); depth: number of levels;
    capacities: slots per level; ceiling: the largest capacity a level may grow to."""

    def __init__(self, roots: Sequence[int], depth: int, capacities: Sequence[int],
                 ceiling: int):
        self.roots = tuple(int(v) for v in roots)
        self.depth = int(depth)
        self.ceiling = int(ceiling)
        if len(capacities) != self.depth:
            raise ValueError("one capacity per level")
        if capacities[0] < int(np.prod(self.roots)) or max(capacities) > self.ceiling:
            raise ValueError("capacities: level 0 holds every root tile; none above the ceiling")
        self.arenas = [SlotArena(c) for c in capacities]
        self.slot_of: Dict[Tile, int] = {}
        self.votes: Dict[Tile, int] = {}
        self.high_water = [0] * self.depth
        self.widenings: List[Tuple[int, int, int]] = []     # (level, old, new capacity)
        for i, j, k in itertools.product(*[range(n) for n in self.roots]):
            self.plant((0, i, j, k))
        self.note_occupancy()

    # membership ----------------------------------------------------------------------------
    def __contains__(self, t) -> bool:
        return t in self.slot_of

    def plant(self, t: Tile) -> int:
        """Create tile t (its parent must exist unless it is a root); returns its slot."""
        t = tuple(int(v) for v in t)
        if t in self.slot_of:
            raise ValueError(f"tile {t} exists")
        if not 0 <= t[0] < self.depth:
            raise ValueError(f"tile {t}: level outside 0..{self.depth - 1}")
        if t[0] == 0 and not all(0 <= t[1 + a] < self.roots[a] for a in range(3)):
            raise ValueError(f"root tile {t} outside the root lattice {self.roots}")
        s = self.arenas[t[0]].take(t)
        self.slot_of[t] = s
        self.votes[t] = 0
        return s

    def prune(self, t: Tile) -> None:
        self.arenas[t[0]].release(self.slot_of.pop(t))
        self.votes.pop(t, None)

    def capacity(self, level: int) -> int:
        return self.arenas[level].capacity

    def widen(self, level: int) -> Tuple[int, int]:
        """Double a level's slot capacity (up to the ceiling); returns (old, new)."""
        old = self.capacity(level)
        new = min(2 * old, self.ceiling)
        if new <= old:
            raise RuntimeError(f"level {level} is at its slot ceiling ({self.ceiling})")
        self.arenas[level].widen(new)
        self.widenings.append((level, old, new))
        return old, new

    # queries -------------------------------------------------------------------------------
    def on_level(self, level: int) -> List[Tile]:
        """The level's tiles in slot order."""
        return [o for o in self.arenas[level].owner if o is not None]

    def all_tiles(self) -> List[Tile]:
        """Every tile, level by level in slot order (the batch order)."""
        return [t for L in range(self.depth) for t in self.on_level(L)]

    def count(self, level: int) -> int:
        return self.arenas[level].used

    def kids(self, t: Tile) -> List[Tile]:
        return [c for c in child_tiles(t) if c in self.slot_of]

    def neighbour(self, t: Tile, face: int) -> Optional[Tile]:
        """The same-level tile across a face if it exists (no wrap: the root level's periodic
        images are the ghost program's business)."""
        n = face_step(t, face)
        return n if n in self.slot_of else None

    def nesting_errors(self) -> List[Tuple[Tile, str]]:
        """Tiles of level >= 1 whose parent tile does not exist (a 2:1 child always lies in its
        parent's footprint by construction)."""
        return [(t, "parent missing") for t in self.slot_of
                if t[0] > 0 and parent_tile(t) not in self.slot_of]

    def note_occupancy(self) -> None:
        for L in range(self.depth):
            self.high_water[L] = max(self.high_water[L], self.count(L))


# --- regrid decisions (host) -------------------------------------------------------------------

@dataclass
class RegridPolicy:
    """When blocks split and merge.

    A tile's counter moves one step toward +patience while its indicator exceeds
    ``refine_above``, one step toward -patience while it is below ``coarsen_below``, and one
    step back toward 0 otherwise; the tile refines when the counter sits at +patience and
    coarsens at -patience. Root tiles never coarsen, tiles on ``max_level`` (default: the
    deepest level) never refine, a tile with children never coarsens (so a hierarchy thins one
    level per pass, finest first), and refinement spreads to every tile of the same level
    within ``buffer`` lattice steps in each axis."""
    refine_above: float
    coarsen_below: float
    patience: int = 3
    buffer: int = 1
    max_level: Optional[int] = None

    def __post_init__(self):
        if not self.refine_above > self.coarsen_below:
            raise ValueError("refine_above must exceed coarsen_below")
        if self.patience < 1:
            raise ValueError("patience must be at least 1")

    def vote(self, tree: Octree, score: Dict[Tile, float]) -> Dict[Tile, int]:
        """Advance every tile's counter by its score (missing = 0) and return the raw action
        of each tile (before the nesting and buffer rules)."""
        P = self.patience
        top = tree.depth - 1 if self.max_level is None else min(self.max_level, tree.depth - 1)
        out = {}
        for t in tree.all_tiles():
            v = float(score.get(t, 0.0))
            c = tree.votes[t]
            if v > self.refine_above:
                c = min(c + 1, P)
            elif v < self.coarsen_below:
                c = max(c - 1, -P)
            else:
                c -= (c > 0) - (c < 0)
            tree.votes[t] = c
            act = REFINE if c == P else (COARSEN if c == -P else KEEP)
            if (act == COARSEN and t[0] == 0) or (act == REFINE and t[0] >= top):
                act = KEEP
            out[t] = act
        return out

    def settle(self, tree: Octree, acts: Dict[Tile, int]) -> Dict[Tile, int]:
        """Apply the nesting and buffer rules to raw actions."""
        acts = dict(acts)
        for t, a in acts.items():
            if a == COARSEN and tree.kids(t):
                acts[t] = KEEP
        if self.buffer > 0:
            grow = [t for t, a in acts.items() if a == REFINE]
            span = range(-self.buffer, self.buffer + 1)
            for L, i, j, k in grow:
                for di, dj, dk in itertools.product(span, span, span):
                    n = (L, i + di, j + dj, k + dk)
                    if n in acts:
                        acts[n] = REFINE
        return acts

    def decide(self, tree: Octree, score: Dict[Tile, float]) -> Dict[Tile, int]:
        return self.settle(tree, self.vote(tree, score))


# --- transfer entries (host) -------------------------------------------------------------------

def _cube(W: int) -> np.ndarray:
    """All points of a W^3 padded block, x slowest."""
    return np.stack(np.meshgrid(np.arange(W), np.arange(W), np.arange(W), indexing="ij"),
                    -1).reshape(-1, 3)


def _flat(c: np.ndarray, W: int) -> np.ndarray:
    return (c[:, 0] * W + c[:, 1]) * W + c[:, 2]


def prolong_entries(points: np.ndarray, dst_off: int, parent_off: int, corner, BS: int,
                    ng: int) -> np.ndarray:
    """bssn_prolong_point entries for the given padded-block points of a child in octant
    ``corner`` of its parent: child point ff is index ng + corner BS + ff of the 2:1 refinement
    of the parent's padded block (the child's first interior node coincides with parent node
    ng + corner BS / 2)."""
    W = BS + 2 * ng
    e = np.empty(points.shape[0], dtype=PROLONG_DTYPE)
    e["dst"] = dst_off + _flat(points, W)
    e["parent"] = parent_off
    e["dst_stride"] = W ** 3
    e["fine"] = ng + np.asarray(corner, dtype=np.int64)[None, :] * BS + points
    e["pS"] = W
    return e


def by_parity(e: np.ndarray) -> np.ndarray:
    """Prolongation entries ordered (stably) by the parity class of their fine indices, so that
    the lanes of one wave evaluate the same number of taps."""
    f = e["fine"]
    cls = (f[:, 0] & 1) * 4 + (f[:, 1] & 1) * 2 + (f[:, 2] & 1)
    return e[np.argsort(cls, kind="stable")]


def prolong_tiles(e: np.ndarray, first: int = 0):
    """Point entries grouped by parent and 4x4x4 fine tile (fine index // 4):
    (tiles, points) for bssn_prolong_tile / bssn_prolong_tile_point, the same values point for
    point; tile point indices start at ``first``."""
    tiles = np.empty(0, dtype=TILE_DTYPE)
    pts = np.empty(0, dtype=TILE_POINT_DTYPE)
    if len(e) == 0:
        return tiles, pts
    f = e["fine"].astype(np.int64)
    t = f >> 2
    order = np.lexsort((t[:, 2], t[:, 1], t[:, 0], e["parent"]))
    e, f, t = e[order], f[order], t[order]
    new = np.ones(len(e), bool)
    new[1:] = (e["parent"][1:] != e["parent"][:-1]) | np.any(t[1:] != t[:-1], axis=1)
    start = np.flatnonzero(new)
    tiles = np.empty(len(start), dtype=TILE_DTYPE)
    tiles["parent"] = e["parent"][start]
    tiles["first"] = first + start
    tiles["pS"] = e["pS"][start]
    tiles["origin"] = (t[start] << 2).astype(np.int32)
    tiles["count"] = np.diff(np.append(start, len(e)))
    tiles["pad"] = 0
    pts = np.empty(len(e), dtype=TILE_POINT_DTYPE)
    pts["dst"] = e["dst"]
    pts["dst_stride"] = e["dst_stride"]
    loc = f & 3
    pts["local"] = ((loc[:, 0] * 4 + loc[:, 1]) * 4 + loc[:, 2]).astype(np.int32)
    return tiles, pts


def tile_prolong_points(tiles: np.ndarray, pts: np.ndarray) -> np.ndarray:
    """The point entries a tile list stands for (the inverse of prolong_tiles up to order)."""
    out = np.empty(len(pts), dtype=PROLONG_DTYPE)
    if len(pts) == 0:
        return out
    own = np.repeat(np.arange(len(tiles)), tiles["count"])
    idx = np.concatenate([np.arange(a, a + c) for a, c in zip(tiles["first"], tiles["count"])])
    idx = idx - int(tiles["first"].min())
    q = pts[idx]
    loc = np.stack([q["local"] >> 4, (q["local"] >> 2) & 3, q["local"] & 3], -1)
    out["dst"] = q["dst"]
    out["dst_stride"] = q["dst_stride"]
    out["parent"] = tiles["parent"][own]
    out["pS"] = tiles["pS"][own]
    out["fine"] = tiles["origin"][own] + loc
    return out


def restrict_entries(parent_off: int, child_off: int, corner, BS: int, ng: int) -> np.ndarray:
    """Injection copies of a child's even interior nodes onto the parent nodes they coincide
    with: parent ng + corner BS/2 + a <- child ng + 2 a, a in [0, BS/2)^3."""
    W = BS + 2 * ng
    half = BS // 2
    a = _cube(half)
    e = np.empty(a.shape[0], dtype=HALO_DTYPE)
    e["dst"] = parent_off + _flat(ng + np.asarray(corner)[None, :] * half + a, W)
    e["src"] = child_off + _flat(ng + 2 * a, W)
    e["dst_stride"] = W ** 3
    e["src_stride"] = W ** 3
    return e


def block_job(dst_off: int, parent_off: int, corner, BS: int, ng: int, skip: int) -> np.ndarray:
    e = np.zeros(1, dtype=BLOCK_DTYPE)
    e["dst"], e["parent"], e["corner"] = dst_off, parent_off, corner
    e["skip"], e["BS"], e["ng"] = skip, BS, ng
    return e


def _slab(face: int, BS: int, ng: int):
    """(destination, source) boxes of a face copy: the ghost slab beyond ``face`` over the
    interior extent of the other two axes, and the neighbour's interior layers it mirrors."""
    axis, upper = face >> 1, face & 1
    dst = [(ng, ng + BS)] * 3
    src = [(ng, ng + BS)] * 3
    dst[axis] = (ng + BS, BS + 2 * ng) if upper else (0, ng)
    src[axis] = (ng, 2 * ng) if upper else (BS, BS + ng)
    grid = lambda r: np.stack(np.meshgrid(*[np.arange(a, b) for a, b in r], indexing="ij"),
                              -1).reshape(-1, 3)
    return grid(dst), grid(src)


def _copies(dst_off: int, dst_pts: np.ndarray, src_off, src_pts: np.ndarray, W: int):
    e = np.empty(dst_pts.shape[0], dtype=HALO_DTYPE)
    e["dst"] = dst_off + _flat(dst_pts, W)
    e["src"] = src_off + _flat(src_pts, W)
    e["dst_stride"] = W ** 3
    e["src_stride"] = W ** 3
    return e


def job_prolong_points(jobs: np.ndarray) -> np.ndarray:
    """Whole-block jobs as the point entries with the same values: the padded-block points a
    job writes (ghosts outside its skipped face slabs, the interior too with bit 6)."""
    out = []
    for j in jobs:
        BS, ng = int(j["BS"]), int(j["ng"])
        W = BS + 2 * ng
        c = _cube(W)
        inside = (c >= ng) & (c < ng + BS)
        keep = np.ones(len(c), bool) if j["skip"] & SKIP_INTERIOR_WRITE else \
            ~np.all(inside, axis=1)
        for ax in range(3):
            slab = ~inside[:, ax] & np.all(np.delete(inside, ax, axis=1), axis=1)
            lo = c[:, ax] < ng
            keep &= ~(slab & lo & bool(j["skip"] & (1 << (2 * ax))))
            keep &= ~(slab & ~lo & bool(j["skip"] & (1 << (2 * ax + 1))))
        out.append(prolong_entries(c[keep], int(j["dst"]), int(j["parent"]),
                                   tuple(int(v) for v in j["corner"]), BS, ng))
    return np.concatenate(out) if out else np.empty(0, dtype=PROLONG_DTYPE)


def prolong_taps(pr: np.ndarray, ng: int) -> np.ndarray:
    """Pool offsets (field 0) of the parent ghost points that prolongation entries read: per
    axis coarse node fine / 2 (even fine) or (fine - 1) / 2 - 2 .. + 3 (odd), clamped to the
    parent's padded block, as bssn_prolong_f64 reads them."""
    found = []
    fine = pr["fine"].astype(np.int64)
    pS = pr["pS"].astype(np.int64)
    par = pr["parent"].astype(np.int64)
    cls = (fine[:, 0] & 1) << 2 | (fine[:, 1] & 1) << 1 | (fine[:, 2] & 1)
    for c in range(8):
        sel = np.flatnonzero(cls == c)
        for lo in range(0, len(sel), 1 << 15):
            e = sel[lo:lo + (1 << 15)]
            taps, ghost = [], []
            for a in range(3):
                f, S = fine[e, a], pS[e, a]
                if (c >> (2 - a)) & 1:
                    t = (f[:, None] - 1) // 2 + np.arange(-2, 4)[None, :]
                else:
                    t = (f // 2)[:, None]
                t = np.clip(t, 0, S[:, None] - 1)
                taps.append(t)
                ghost.append((t < ng) | (t >= S[:, None] - ng))
            tx, ty, tz = taps
            S1, S2 = pS[e, 1][:, None, None, None], pS[e, 2][:, None, None, None]
            flat = par[e][:, None, None, None] + (tx[:, :, None, None] * S1
                                                  + ty[:, None, :, None]) * S2 \
                + tz[:, None, None, :]
            g = ghost[0][:, :, None, None] | ghost[1][:, None, :, None] | \
                ghost[2][:, None, None, :]
            found.append(flat[g])
    return np.unique(np.concatenate(found)) if found else np.empty(0, np.int64)


def linked_entries(entries, ng: int):
    """The ghost program of a linked step: every prolongation (points and whole-block jobs), and
    only the copies whose destination a prolongation reads, in the program's phases."""
    pr, jb, cp, pe, be, ce = entries
    taps = prolong_taps(np.concatenate([pr, job_prolong_points(jb)]), ng)
    need = np.isin(cp["dst"].astype(np.int64), taps)
    kept = np.cumsum(need)
    ce2 = [int(kept[e - 1]) if e > 0 else 0 for e in ce]
    return pr, jb, cp[need], pe, be, ce2


def link_map(off, P, copies: np.ndarray, pool_doubles: int) -> np.ndarray:
    """(n_points, 2) int32 over the padded points of blocks at pool offsets ``off`` with ``P``
    points each (block order, each block row-major): the pool offset of field 0 of the point a
    linked stage reads there and that block's field stride -- a copy entry's source at a copied
    ghost, the point itself elsewhere (interior points, prolongated ghosts)."""
    off = np.asarray(off, dtype=np.int64)
    P = np.asarray(P, dtype=np.int64)
    if pool_doubles >= 2 ** 29:
        raise ValueError("linked steps need a pool below 2^29 doubles (4 GiB)")
    m = np.concatenate([[0], np.cumsum(P)[:-1]])
    n = int(P.sum())
    blk = np.repeat(np.arange(len(P)), P)
    q = np.arange(n, dtype=np.int64) - m[blk]
    out = np.empty((n, 2), dtype=np.int32)
    out[:, 0] = off[blk] + q
    out[:, 1] = P[blk]
    if len(copies):
        dst = copies["dst"].astype(np.int64)
        order = np.argsort(off)
        b = order[np.searchsorted(off[order], dst, side="right") - 1]
        i = m[b] + (dst - off[b])
        out[i, 0] = copies["src"]
        out[i, 1] = copies["src_stride"]
    if _lib.load().bssn_count_skipped_entries(5, out.ctypes.data, len(out), pool_doubles) != 0:
        raise RuntimeError("link map entry outside the pool")
    return out


def build_ghost_program(tree: Octree, BS: int, ng: int, offset_of: Callable[[Tile], int],
                        use_blocks: bool | None = None):
    """The hierarchy sync as ghost fill program entries, one phase per level.
    Phase 0: every ghost point of every root tile is copied from the root tile that owns its
    periodic image. Phase L >= 1, per tile: face slabs copied from face neighbours that exist;
    every other ghost point prolongated from the parent -- one whole-block job per tile when
    the padded extent allows (use_blocks, default when W <= BLOCK_WMAX), else point entries
    grouped by parity.
    Returns (prolongs, jobs, copies, prolong_end, job_end, copy_end) (ends cumulative)."""
    W = BS + 2 * ng
    if use_blocks is None:
        use_blocks = W <= BLOCK_WMAX
    if use_blocks and W > BLOCK_WMAX:
        raise ValueError(f"whole-block prolongation needs BS + 2 ng <= {BLOCK_WMAX}")
    roots = np.asarray(tree.roots, dtype=np.int64)
    if tree.count(0) != int(np.prod(roots)):
        raise ValueError("the root level must tile the box")
    pts = _cube(W)
    halo = pts[np.any((pts < ng) | (pts >= ng + BS), axis=1)]
    root_offset = np.empty(tuple(roots), dtype=np.int64)
    for t in tree.on_level(0):
        root_offset[t[1], t[2], t[3]] = offset_of(t)
    pro, job, cop = [], [], []
    ends = ([], [], [])

    def close_phase():
        for lst, acc in zip((pro, job, cop), ends):
            acc.append(sum(len(x) for x in lst))

    span = roots * BS                       # root nodes per axis (periodic)
    for t in tree.on_level(0):
        node = np.asarray(t[1:]) * BS + halo - ng      # level-0 node index of each ghost
        node %= span[None, :]
        own = node // BS
        cop.append(_copies(offset_of(t), halo, root_offset[own[:, 0], own[:, 1], own[:, 2]],
                           node - own * BS + ng, W))
    close_phase()
    for L in range(1, tree.depth):
        point_lists = []
        for t in tree.on_level(L):
            covered = np.zeros((W, W, W), bool)
            covered[ng:ng + BS, ng:ng + BS, ng:ng + BS] = True
            skip = 0
            for face in range(6):
                n = tree.neighbour(t, face)
                if n is None:
                    continue
                d, s = _slab(face, BS, ng)
                cop.append(_copies(offset_of(t), d, offset_of(n), s, W))
                covered[d[:, 0], d[:, 1], d[:, 2]] = True
                skip |= 1 << face
            par = parent_tile(t)
            if use_blocks:
                job.append(block_job(offset_of(t), offset_of(par), octant(t), BS, ng, skip))
            else:
                rest = pts[~covered.reshape(-1)]
                point_lists.append(prolong_entries(rest, offset_of(t), offset_of(par),
                                                   octant(t), BS, ng))
        if point_lists:
            pro.append(by_parity(np.concatenate(point_lists)))
        close_phase()
    cat = lambda xs, dt: np.concatenate(xs) if xs else np.empty(0, dtype=dt)
    return (cat(pro, PROLONG_DTYPE), cat(job, BLOCK_DTYPE), cat(cop, HALO_DTYPE), *ends)


class DeviceProgram:
    """A ghost fill program in device memory (keeps the entry arrays alive). With tiles=True the
    point prolongations go to the device grouped by 4x4x4 fine tile (prolong_tiles: the same
    values, each tile's coarse window read once and its x and y passes shared)."""

    def __init__(self, prolongs: np.ndarray, jobs: np.ndarray, copies: np.ndarray, pe, be, ce,
                 device, tiles: bool = False):
        self.n_prolong, self.n_block, self.n_copy = len(prolongs), len(jobs), len(copies)
        put = lambda a: upload(a, device) if len(a) else None
        te = [0] * len(pe)
        self.tiles = np.empty(0, dtype=TILE_DTYPE)
        self.tile_points = np.empty(0, dtype=TILE_POINT_DTYPE)
        if tiles and len(prolongs):
            tparts, qparts, lo, nq = [], [], 0, 0
            for i, hi in enumerate(pe):
                tt, qq = prolong_tiles(prolongs[lo:hi], first=nq)
                tparts.append(tt)
                qparts.append(qq)
                nq += len(qq)
                te[i] = sum(len(x) for x in tparts)
                lo = hi
            self.tiles = np.concatenate(tparts)
            self.tile_points = np.concatenate(qparts)
            prolongs = prolongs[:0]
            pe = [0] * len(pe)
        self.n_tile = len(self.tiles)
        self._p, self._b, self._c = put(prolongs), put(jobs), put(copies)
        self._t, self._q = put(self.tiles), put(self.tile_points)
        self._ends = [(ctypes.c_longlong * len(x))(*x) for x in (pe, be, ce, te)]
        addr = lambda t: t.data_ptr() if t is not None else None
        self.struct = _lib.GhostProgram(len(pe), *self._ends[:3], addr(self._p), addr(self._b),
                                        addr(self._c), self._ends[3], addr(self._t),
                                        addr(self._q), len(self.tile_points))

    @property
    def ptr(self):
        return ctypes.byref(self.struct)


def upload(entries: np.ndarray, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(entries).view(np.uint8).copy()).to(device)


# --- the hierarchy -----------------------------------------------------------------------------

class AMRHierarchy:
    """An octree hierarchy of BS^3-node blocks over a periodic box starting at ``lo``: a root
    tiling of nb_root blocks at spacing h0 = L / (nb_root[0] BS) (L is the box's x extent),
    refined 2:1 up to ``levels`` levels; every block has ng ghost layers. The state is one
    float64 pool (``empty_pool``); tile t's block is ``view(U, t)``, a (24, W, W, W) state
    (W = BS + 2 ng). Steps are shared-dt RK4 over all blocks."""

    def __init__(self, nb_root=(2, 2, 2), BS: int = 16, ng: int = 3, levels: int = 3,
                 caps: Sequence[int] | None = None, max_blocks: int = 512, L: float = 1.0,
                 lo=(-0.5, -0.5, -0.5), device="cuda"):
        if BS % 4 or BS < 4:
            raise ValueError("BS must be a positive multiple of 4 (2:1 octants of whole bricks)")
        if ng < 3:
            raise ValueError("ng must be >= 3 (the 6-point prolongation and the stencils)")
        self.nb_root = tuple(int(v) for v in nb_root)
        self.BS, self.ng, self.levels = int(BS), int(ng), int(levels)
        self.W = self.BS + 2 * self.ng
        self.device = _device(device)
        self.L = float(L)
        self.lo = tuple(float(v) for v in lo)
        nroot = int(np.prod(self.nb_root))
        if caps is None:
            caps = [nroot] + [min(max_blocks, 8 * nroot)] * (levels - 1)
        self.tree = Octree(self.nb_root, levels, [int(c) for c in caps], int(max_blocks))
        self.h0 = self.L / (self.nb_root[0] * self.BS)
        self._place()
        self._batch = None
        self._prog = None
        self._linked = None
        self._link = None

    # pool layout ---------------------------------------------------------------------------
    def _place(self):
        self.slot_doubles = (NUM_VARS * self.W ** 3 + 63) // 64 * 64
        self.level_base, pos = [], 0
        for L in range(self.levels):
            self.level_base.append(pos)
            pos += self.tree.capacity(L) * self.slot_doubles
        self.pool_doubles = pos

    def slot_offset(self, t: Tile) -> int:
        """Pool offset of tile t's block."""
        return self.level_base[t[0]] + self.tree.slot_of[tuple(t)] * self.slot_doubles

    def empty_pool(self) -> torch.Tensor:
        return torch.zeros(self.pool_doubles, dtype=torch.float64, device=self.device)

    def view(self, U, t: Tile):
        """Tile t's block of a pool (torch tensor or numpy array) as a (24, W, W, W) view."""
        o = self.slot_offset(t)
        blk = U[o:o + NUM_VARS * self.W ** 3]
        shape = (NUM_VARS, self.W, self.W, self.W)
        return blk.view(*shape) if isinstance(blk, torch.Tensor) else blk.reshape(shape)

    def spacing(self, level: int) -> float:
        return self.h0 / 2 ** level

    def origin(self, t: Tile) -> Tuple[int, int, int]:
        """Level node index of tile t's first interior node."""
        return tuple(self.BS * v for v in t[1:])

    def block(self, t: Tile) -> Block:
        h = self.spacing(t[0])
        o = self.origin(t)
        return Block(n=(self.BS,) * 3, h=(h,) * 3,
                     lo=tuple(self.lo[a] + o[a] * h for a in range(3)), level=t[0])

    def grid(self, t: Tile) -> Grid:
        return self.block(t).grid(self.ng)

    # plans ---------------------------------------------------------------------------------
    def _drop_plans(self):
        if self._batch is not None:
            self._batch.close()
        self._batch = None
        self._prog = None
        self._linked = None
        self._link = None

    @property
    def keys(self) -> List[Tile]:
        """Every tile in batch order (level, then slot)."""
        return self.tree.all_tiles()

    @property
    def batch(self) -> BlockBatch:
        if self._batch is None:
            tiles = self.keys
            self._batch = BlockBatch([self.block(t) for t in tiles], ng=self.ng,
                                     device=self.device,
                                     offsets=[self.slot_offset(t) for t in tiles],
                                     pool_doubles=self.pool_doubles)
        return self._batch

    def program_entries(self, use_blocks: bool | None = None):
        return build_ghost_program(self.tree, self.BS, self.ng, self.slot_offset, use_blocks)

    @property
    def program(self) -> DeviceProgram:
        if self._prog is None:
            pr, jb, cp, pe, be, ce = self.program_entries()
            self._validate(pr, cp, jb)
            self._prog = DeviceProgram(pr, jb, cp, pe, be, ce, self.device)
        return self._prog

    @property
    def linked_program(self) -> DeviceProgram:
        """The program of a linked step: every prolongation (points and whole-block jobs) and
        only the copies into the parent ghosts their taps read."""
        if self._linked is None:
            pr, jb, cp, pe, be, ce = linked_entries(self.program_entries(), self.ng)
            self._validate(pr, cp, jb)
            self._linked = DeviceProgram(pr, jb, cp, pe, be, ce, self.device)
        return self._linked

    @property
    def link(self) -> torch.Tensor:
        """The link map of the tiles in batch order, on the device."""
        if self._link is None:
            tiles = self.keys
            off = [self.slot_offset(t) for t in tiles]
            m = link_map(off, [self.W ** 3] * len(tiles), self.program_entries()[2],
                         self.pool_doubles)
            self._link = torch.from_numpy(m).to(self.device)
        return self._link

    def _validate(self, prolongs: np.ndarray, copies: np.ndarray, jobs: np.ndarray):
        """Every entry must address points of blocks inside this pool (the device kernels do
        not bound-check offsets)."""
        blk = NUM_VARS * self.W ** 3
        point_top = self.pool_doubles - (NUM_VARS - 1) * self.W ** 3
        checks = [(prolongs["dst"], point_top), (prolongs["parent"], self.pool_doubles - blk + 1),
                  (copies["dst"], point_top), (copies["src"], point_top),
                  (jobs["dst"], self.pool_doubles - blk + 1),
                  (jobs["parent"], self.pool_doubles - blk + 1)]
        for arr, top in checks:
            if len(arr) and (arr.min() < 0 or arr.max() >= top):
                raise RuntimeError("ghost program entry outside the pool")
        if len(jobs) and (self.W > BLOCK_WMAX or np.any(jobs["BS"] != self.BS) or
                          np.any(jobs["ng"] != self.ng)):
            raise RuntimeError("whole-block jobs need this hierarchy's BS, ng and W <= 24")
        if len(prolongs) and (prolongs["fine"].min() < 0 or
                              prolongs["fine"].max() >= 2 * self.W):
            raise RuntimeError("prolongation index outside the refined parent")

    def _check_pool(self, U: torch.Tensor):
        if not isinstance(U, torch.Tensor) or U.dtype != torch.float64 or \
                U.numel() != self.pool_doubles or not U.is_contiguous() or U.device != self.device:
            raise ValueError("U must be this hierarchy's contiguous float64 pool")

    # data ----------------------------------------------------------------------------------
    def fill(self, U: torch.Tensor, fn: Callable[[Grid], torch.Tensor], sync: bool = True):
        """Every block <- fn(its Grid) (a (24, W, W, W) state at the block's nodes)."""
        self._check_pool(U)
        for t in self.keys:
            self.view(U, t).copy_(fn(self.grid(t)))
        if sync:
            self.sync_(U)
        return U

    def _root_slices(self, t: Tile, g: int):
        o = self.origin(t)
        return (slice(None),) + tuple(slice(g + o[a], g + o[a] + self.BS) for a in range(3))

    def from_global(self, U: torch.Tensor, F: torch.Tensor, sync: bool = True):
        """Root interiors <- the interior of a single-grid (24, N+2g, ...) state (any ghost
        width g), N = nb_root BS per axis; then the hierarchy is synced."""
        self._check_pool(U)
        g = (F.shape[1] - self.nb_root[0] * self.BS) // 2
        inner = (slice(None),) + (slice(self.ng, self.ng + self.BS),) * 3
        for t in self.tree.on_level(0):
            self.view(U, t)[inner].copy_(F[self._root_slices(t, g)])
        if sync:
            self.sync_(U)
        return U

    def to_global(self, U: torch.Tensor) -> torch.Tensor:
        """The root level's interiors as one (24, N0, N1, N2) array."""
        out = torch.empty((NUM_VARS, *[n * self.BS for n in self.nb_root]), dtype=U.dtype,
                          device=U.device)
        inner = (slice(None),) + (slice(self.ng, self.ng + self.BS),) * 3
        for t in self.tree.on_level(0):
            out[self._root_slices(t, 0)] = self.view(U, t)[inner]
        return out

    # operators -----------------------------------------------------------------------------
    def sync_(self, U: torch.Tensor) -> torch.Tensor:
        self._check_pool(U)
        _lib.check(_lib.load().bssn_ghost_fill_f64(U.data_ptr(), U.numel(), self.program.ptr,
                                                   stream_handle(self.device)))
        return U

    def workspace(self) -> torch.Tensor:
        return self.batch.workspace()

    def step_(self, U: torch.Tensor, t: float, dt: float, params: PhysicsParams | None = None,
              order: int = 4, ko_sigma: float = 0.1, tdev: torch.Tensor | None = None,
              ghosts: str = "program", sync_end: bool = True):
        """One shared-dt RK4 step of every block in place (U synced on entry and on exit).
        ghosts="linked": the stages read face-copied ghosts at their sources
        (bssn_batch_rk4_step_link_f64); with sync_end=False only the ghosts the next linked step
        reads are left up to date."""
        params = params or PhysicsParams()
        self._check_pool(U)
        b = self.batch
        prog = self.program
        tptr = None
        if tdev is not None:
            if tdev.dtype != torch.float64 or tdev.device != self.device or tdev.numel() != 1:
                raise ValueError("tdev must be one float64 value on the hierarchy's device")
            tptr = tdev.data_ptr()
        if ghosts == "linked":
            _lib.check(_lib.load().bssn_batch_rk4_step_link_f64(
                b._handle, U.data_ptr(), self.workspace().data_ptr(),
                _lib.dbl(b.S16(params, dt, t)), int(order), float(ko_sigma),
                self.link.data_ptr(), self.link.shape[0], self.linked_program.ptr,
                prog.ptr if sync_end else None, float(t), float(dt), params.chi_floor,
                params.alpha_floor, tptr,
                stream_handle(self.device)))
            return U
        if ghosts != "program":
            raise ValueError(f"ghosts must be 'program' or 'linked', not {ghosts!r}")
        _lib.check(_lib.load().bssn_batch_rk4_step_prog_f64(
            b._handle, U.data_ptr(), self.workspace().data_ptr(),
            _lib.dbl(b.S16(params, dt, t)), int(order), float(ko_sigma), prog.ptr, float(t),
            float(dt), params.chi_floor, params.alpha_floor, tptr, stream_handle(self.device)))
        return U

    def tune_launch(self, U: torch.Tensor, dt: float, candidates=None, reps: int = 3,
                    **kw) -> dict:
        """Time one hierarchy step per batched launch shape on a copy of U and keep the fastest
        (the first candidate unless another beats it by more than the timing noise,
        batch.tune_step_launch); ``kw`` goes to step_ (ghosts="linked" tunes the linked step).
        Returns {shape: ms} (0 = one workgroup per brick, N = N persistent waves)."""
        return tune_step_launch(self.batch, lambda V: self.step_(V, 0.0, dt, **kw), U,
                                candidates, reps)

    def capture_steps(self, U: torch.Tensor, tdev: torch.Tensor, dt: float, nsteps: int = 1,
                      **kw) -> torch.cuda.CUDAGraph:
        """Capture ``nsteps`` hierarchy steps of U into a HIP graph (time in tdev, advanced by
        each replay). The tiles must not change while the graph is in use."""
        self.workspace()
        prog = self.program
        if kw.get("ghosts") == "linked":
            prog = (prog, self.linked_program, self.link)
        scratch = U.clone()               # first launches (code loading) outside the capture
        self.step_(scratch, 0.0, dt, **kw)
        torch.cuda.synchronize(self.device)
        del scratch
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                for _ in range(nsteps):
                    self.step_(U, 0.0, dt, tdev=tdev, **kw)
        torch.cuda.current_stream(self.device).wait_stream(s)
        self._graph_keep = prog           # the captured launches point at this program
        return g

    def evolve(self, U: torch.Tensor, t: float, dt: float, nsteps: int, **kw) -> float:
        for _ in range(nsteps):
            self.step_(U, t, dt, **kw)
            t = t + dt
        return t

    def indicators(self, U: torch.Tensor,
                   fields: Sequence[int] | None = None) -> Dict[Tile, float]:
        """Per tile: the largest undivided difference over the chosen fields (all by default)."""
        self._check_pool(U)
        tiles = self.keys
        out = torch.empty(len(tiles) * NUM_VARS, dtype=torch.float64, device=self.device)
        _lib.check(_lib.load().bssn_batch_indicator_f64(self.batch._handle, U.data_ptr(),
                                                        out.data_ptr(),
                                                        stream_handle(self.device)))
        v = out.view(len(tiles), NUM_VARS)
        if fields is not None:
            v = v[:, list(fields)]
        return dict(zip(tiles, v.max(dim=1).values.cpu().numpy().tolist()))

    # regrid --------------------------------------------------------------------------------
    def _widen(self, U: torch.Tensor, level: int) -> torch.Tensor:
        """Double a level's slot capacity: a new pool, every level copied to its new base."""
        caps = [self.tree.capacity(L) for L in range(self.levels)]
        base = list(self.level_base)
        self.tree.widen(level)
        self._place()
        V = self.empty_pool()
        for L in range(self.levels):
            n = caps[L] * self.slot_doubles
            V[self.level_base[L]:self.level_base[L] + n] = U[base[L]:base[L] + n]
        self._drop_plans()
        return V

    def _inject(self, U: torch.Tensor, tiles: List[Tile]) -> None:
        """Copy each tile's coincident interior nodes into its parent (one launch)."""
        if not tiles:
            return
        e = np.concatenate([restrict_entries(self.slot_offset(parent_tile(t)),
                                             self.slot_offset(t), octant(t), self.BS, self.ng)
                            for t in tiles])
        _lib.check(_lib.load().bssn_batch_halo_f64(U.data_ptr(), U.numel(),
                                                   upload(e, self.device).data_ptr(), len(e),
                                                   stream_handle(self.device)))

    def _interpolate(self, U: torch.Tensor, tiles: List[Tile]) -> None:
        """Fill each new tile's whole padded block from its parent (one launch)."""
        if not tiles:
            return
        lib = _lib.load()
        st = stream_handle(self.device)
        BS, ng = self.BS, self.ng
        if self.W <= BLOCK_WMAX:
            e = np.concatenate([block_job(self.slot_offset(t), self.slot_offset(parent_tile(t)),
                                          octant(t), BS, ng, SKIP_INTERIOR_WRITE) for t in tiles])
            _lib.check(lib.bssn_prolong_blocks_f64(U.data_ptr(), U.numel(),
                                                   upload(e, self.device).data_ptr(), len(e), st))
        else:
            pts = _cube(self.W)
            e = by_parity(np.concatenate([prolong_entries(pts, self.slot_offset(t),
                                                          self.slot_offset(parent_tile(t)),
                                                          octant(t), BS, ng) for t in tiles]))
            _lib.check(lib.bssn_prolong_f64(U.data_ptr(), U.numel(),
                                            upload(e, self.device).data_ptr(), len(e), st))

    def apply(self, U: torch.Tensor, acts: Dict[Tile, int], grow: bool = True) -> torch.Tensor:
        """Carry out regrid actions. Coarsening runs finest level first (each tile's nodes are
        injected into its parent, then the tile is removed); refinement coarsest level first
        (a refining tile gets all eight children, each interpolated from it over its whole
        padded block; a full level doubles its slots when ``grow``). Returns the pool (a new one
        if a level grew); the hierarchy is synced if anything changed."""
        tree = self.tree
        changed = False
        for L in range(self.levels - 1, 0, -1):
            gone = [t for t in tree.on_level(L) if acts.get(t) == COARSEN]
            self._inject(U, gone)
            for t in gone:
                tree.prune(t)
            changed |= bool(gone)
        for L in range(self.levels - 1):
            born = []
            for t in tree.on_level(L):
                if acts.get(t) != REFINE:
                    continue
                for c in child_tiles(t):
                    if c in tree:
                        continue
                    if tree.arenas[L + 1].full():
                        if not grow:
                            raise RuntimeError(f"level {L + 1} has no free slot")
                        torch.cuda.synchronize(self.device)
                        U = self._widen(U, L + 1)
                    tree.plant(c)
                    born.append(c)
            self._interpolate(U, born)      # offsets after any widening
            changed |= bool(born)
        if changed:
            self._drop_plans()
            self.sync_(U)
        tree.note_occupancy()
        return U

    def regrid(self, U: torch.Tensor, policy: RegridPolicy,
               fields: Sequence[int] | None = None) -> torch.Tensor:
        """Indicators, the policy's decisions, and their application. Returns the pool."""
        return self.apply(U, policy.decide(self.tree, self.indicators(U, fields)))

    def run(self, U: torch.Tensor, t: float, dt: float, nsteps: int, policy: RegridPolicy,
            every: int, fields: Sequence[int] | None = None, on_regrid=None, **step_kw):
        """nsteps shared-dt steps with a regrid after every ``every`` steps (not after the last
        one). on_regrid(steps_done, U, self) is called after each regrid. Returns (U, t)."""
        done = 0
        while done < nsteps:
            n = min(every, nsteps - done)
            t = self.evolve(U, t, dt, n, **step_kw)
            done += n
            if done < nsteps:
                U = self.regrid(U, policy, fields)
                if on_regrid is not None:
                    on_regrid(done, U, self)
        return U, t
