"""Block-structured AMR for the BSSN state on the device: a live, node-centred hierarchy.

Reference patterns (the reference has no BSSN AMR; its block AMR serves the MCS and wave
systems):
  * node-centred 2:1 transfers and the indicator: 3D/src/multipatch/amr/kernels.py:53-158
    (prolongation = tensor product of fine[2k] = coarse[k], fine[2k+1] = 6-point Lagrange
    midpoint; restriction = injection; indicator = largest undivided difference);
  * topology bookkeeping: 3D/src/multipatch/amr/state.py:109-275 (bbox in level node indices,
    parent/children, 6-face neighbours by bbox adjacency with no wrap on fine levels, proper
    nesting);
  * ghost sync of the shared-dt N-level step: 2D/src/mcs2d/amr/evolve.py:719-735 and
    3D/src/multipatch/amr/sync.py:84-122 (root level periodic; each finer level, coarse to
    fine, from a prolongation of its parent, then same-level face slabs over the interior
    extent from neighbours);
  * regrid: 2D/src/mcs2d/amr/regrid.py:113-486 (hysteresis flags, no-orphan nesting and
    Chebyshev buffer dilation, coarsen finest-first by restriction, refine coarse-first by
    prolongation, slot capacities that grow on demand), in 3D (8 octants).

Device side: every block lives in a fixed slot of one float64 pool (level L's slots follow
level L-1's); the active blocks form a BlockBatch (its plan lists only active blocks, so idle
slots cost nothing); the ghost sync is a ghost fill program (bssn_ghost_fill_f64: one phase per
level, prolongations and copies) run after every RK stage by bssn_batch_rk4_step_prog_f64.
Prolongations run as whole-block jobs (bssn_prolong_blocks_f64: separable x, y, z passes
through LDS, padded blocks up to 24 per axis) or point lists (bssn_prolong_f64); restrictions
as copy lists (bssn_batch_halo_f64); the indicator is bssn_batch_indicator_f64.

One deliberate difference: a slot's hysteresis streak restarts at 0 when a block is created in
it (the reference keeps the streak of the block that last held the slot).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .batch import HALO_DTYPE, Block, BlockBatch, _device
from .grid import Grid
from .rhs import stream_handle
from .state import NUM_VARS, PhysicsParams

REFINE, KEEP, COARSEN = 1, 0, -1
RATIO = 2
NFACE = 6

PROLONG_DTYPE = np.dtype([("dst", "<i8"), ("parent", "<i8"), ("dst_stride", "<i8"),
                          ("fine", "<i4", (3,)), ("pS", "<i4", (3,))])
assert PROLONG_DTYPE.itemsize == 48
# bssn_prolong_block: a whole-block job (padded extent W = BS + 2 ng <= BLOCK_WMAX)
BLOCK_DTYPE = np.dtype([("dst", "<i8"), ("parent", "<i8"), ("corner", "<i4", (3,)),
                        ("skip", "<i4"), ("BS", "<i4"), ("ng", "<i4")])
assert BLOCK_DTYPE.itemsize == 40
BLOCK_WMAX = 24
SKIP_INTERIOR_WRITE = 64


# --- topology (host) --------------------------------------------------------------------------

@dataclass
class AMRTopology:
    """Host bookkeeping of a hierarchy (3D/src/multipatch/amr/state.py:109-275).
    bbox_ijk[(L, s)] = level-L node index of block (L, s)'s first interior node."""
    levels: int
    BS: int
    caps: List[int]
    max_blocks: int
    active: np.ndarray = None
    streaks: np.ndarray = None
    parent: Dict = field(default_factory=dict)
    children: Dict = field(default_factory=dict)
    bbox_ijk: Dict = field(default_factory=dict)
    neighbors: Dict = field(default_factory=dict)
    peak_active: List[int] = None
    grow_history: List = field(default_factory=list)

    def __post_init__(self):
        if self.active is None:
            self.active = np.zeros((self.levels, self.max_blocks), bool)
        if self.streaks is None:
            self.streaks = np.zeros((self.levels, self.max_blocks), np.int32)
        if self.peak_active is None:
            self.peak_active = [0] * self.levels

    def n_active(self, level: int) -> int:
        return int(self.active[level].sum())

    def record_occupancy(self):
        for L in range(self.levels):
            self.peak_active[L] = max(self.peak_active[L], self.n_active(L))

    def recommended_caps(self, margin: float = 1.5) -> tuple:
        rec = []
        for L in range(self.levels):
            want = self.peak_active[L] if L == 0 else int(np.ceil(self.peak_active[L] * margin))
            rec.append(min(max(want, 1), self.max_blocks))
        return tuple(rec)

    def free_slots(self, level: int) -> np.ndarray:
        return np.flatnonzero(~self.active[level, :self.caps[level]])

    def add_block(self, level: int, slot: int, bbox: tuple, parent: Optional[tuple] = None):
        self.active[level, slot] = True
        self.streaks[level, slot] = 0
        self.bbox_ijk[(level, slot)] = tuple(int(v) for v in bbox)
        if parent is not None:
            self.parent[(level, slot)] = parent
            self.children.setdefault(parent, []).append((level, slot))

    def remove_block(self, level: int, slot: int):
        key = (level, slot)
        self.active[level, slot] = False
        par = self.parent.pop(key, None)
        if par is not None and par in self.children:
            self.children[par] = [c for c in self.children[par] if c != key]
        self.children.pop(key, None)
        self.bbox_ijk.pop(key, None)
        self.neighbors = {k: v for k, v in self.neighbors.items() if k[:2] != key and v != key}

    def blocks(self, level: int | None = None) -> List[Tuple[int, int]]:
        levels = range(self.levels) if level is None else [level]
        return [(L, int(s)) for L in levels for s in np.flatnonzero(self.active[L])]

    def child_corner(self, key) -> tuple:
        """Octant (cx, cy, cz) of block ``key`` inside its parent."""
        L, _ = key
        cb = self.bbox_ijk[key]
        pb = self.bbox_ijk[self.parent[key]]
        half = self.BS // RATIO
        return tuple((cb[a] // RATIO - pb[a]) // half for a in range(3))

    def rebuild_neighbors(self):
        """Same-level 6-face neighbours from bbox adjacency (no wrap): faces 0..5 =
        (-x, +x, -y, +y, -z, +z) (state.py:167-190)."""
        self.neighbors = {}
        BS = self.BS
        for L in range(self.levels):
            pos = {self.bbox_ijk[(L, s)]: s for (_, s) in self.blocks(L)}
            for (i0, j0, k0), s in pos.items():
                cand = ((i0 - BS, j0, k0), (i0 + BS, j0, k0), (i0, j0 - BS, k0),
                        (i0, j0 + BS, k0), (i0, j0, k0 - BS), (i0, j0, k0 + BS))
                for f, p in enumerate(cand):
                    nb = pos.get(p)
                    self.neighbors[(L, s, f)] = (L, nb) if nb is not None else None

    def check_proper_nesting(self) -> list:
        """Every active L >= 1 block has an active parent whose interior contains its footprint
        (state.py:192-216)."""
        bad = []
        half = self.BS // RATIO
        for L in range(1, self.levels):
            for (_, s) in self.blocks(L):
                par = self.parent.get((L, s))
                if par is None:
                    bad.append((L, s, "no parent"))
                    continue
                if not self.active[par]:
                    bad.append((L, s, "parent inactive"))
                    continue
                cb, pb = self.bbox_ijk[(L, s)], self.bbox_ijk[par]
                f0 = [cb[a] // RATIO for a in range(3)]
                if not all(pb[a] <= f0[a] and f0[a] + half <= pb[a] + self.BS for a in range(3)):
                    bad.append((L, s, "footprint outside parent"))
        return bad


# --- regrid decisions (host) ------------------------------------------------------------------

def compute_flags(indicators_per_level: Sequence[np.ndarray], topo: AMRTopology, *,
                  refine_threshold: float, coarsen_threshold: float, hysteresis_K: int = 3,
                  max_level: Optional[int] = None) -> np.ndarray:
    """Indicator -> REFINE / KEEP / COARSEN per (level, slot) with a streak hysteresis
    (regrid.py:113-190): streak +1 above refine_threshold (cap +K), -1 below coarsen_threshold
    (floor -K), else decays by 1 toward 0; REFINE at +K, COARSEN at -K; level 0 never
    coarsens; no block at or below the deepest refinable level refines. Updates topo.streaks."""
    if not refine_threshold > coarsen_threshold:
        raise ValueError("refine_threshold must be > coarsen_threshold")
    deepest = topo.levels - 1 if max_level is None else min(max_level, topo.levels - 1)
    flags = np.zeros((topo.levels, topo.max_blocks), dtype=np.int32)
    K = hysteresis_K
    for L in range(topo.levels):
        ind = np.asarray(indicators_per_level[L], dtype=np.float64)
        act = topo.active[L]
        s = topo.streaks[L]
        above = (ind > refine_threshold) & act
        below = (ind < coarsen_threshold) & act
        new = np.where(above, np.minimum(s + 1, K),
                       np.where(below, np.maximum(s - 1, -K),
                                np.where(s > 0, s - 1, np.where(s < 0, s + 1, 0))))
        topo.streaks[L] = new
        fl = np.where(new >= K, REFINE, np.where(new <= -K, COARSEN, KEEP))
        if L == 0:
            fl = np.where(fl == COARSEN, KEEP, fl)
        if L >= deepest:
            fl = np.where(fl == REFINE, KEEP, fl)
        flags[L] = fl
    return flags


def enforce_nesting_buffer(flags: np.ndarray, topo: AMRTopology, *, n_buffer: int = 1) -> np.ndarray:
    """No coarsening of a block with active children; REFINE spreads to same-level active
    blocks within n_buffer blocks in Chebyshev distance (regrid.py:195-250, in 3D)."""
    flags = flags.copy()
    for key, kids in list(topo.children.items()):
        if topo.active[key] and any(topo.active[k] for k in kids) and flags[key] == COARSEN:
            flags[key] = KEEP
    reach = n_buffer * topo.BS
    for L in range(topo.levels):
        ref = [topo.bbox_ijk[(L, s)] for (_, s) in topo.blocks(L) if flags[L, s] == REFINE]
        if not ref:
            continue
        R = np.asarray(ref)
        for (_, s) in topo.blocks(L):
            if flags[L, s] == REFINE:
                continue
            b = np.asarray(topo.bbox_ijk[(L, s)])
            if np.any(np.all(np.abs(R - b) <= reach, axis=1)):
                flags[L, s] = REFINE
    return flags


# --- transfer entries (host) ------------------------------------------------------------------

def _grid_points(W: int) -> np.ndarray:
    return np.stack(np.meshgrid(np.arange(W), np.arange(W), np.arange(W), indexing="ij"),
                    -1).reshape(-1, 3)


def _lin(c: np.ndarray, W: int) -> np.ndarray:
    return (c[:, 0] * W + c[:, 1]) * W + c[:, 2]


def prolong_entries(points: np.ndarray, dst_off: int, parent_off: int, corner, BS: int,
                    ng: int) -> np.ndarray:
    """Prolongation entries for the given points (full-block indices) of a child block at
    octant ``corner`` of its parent: the child point ff sits at index ng + corner BS + ff of
    the 2:1 refinement of the parent's padded block (kernels.py:78-85)."""
    W = BS + 2 * ng
    e = np.empty(points.shape[0], dtype=PROLONG_DTYPE)
    e["dst"] = dst_off + _lin(points, W)
    e["parent"] = parent_off
    e["dst_stride"] = W ** 3
    e["fine"] = ng + np.asarray(corner, dtype=np.int64)[None, :] * BS + points
    e["pS"] = W
    return e


def by_parity(e: np.ndarray) -> np.ndarray:
    """Prolongation entries ordered by the parity class of their fine indices (stable), so that
    the device's lanes in one wave evaluate the same number of taps."""
    f = e["fine"]
    cls = (f[:, 0] & 1) * 4 + (f[:, 1] & 1) * 2 + (f[:, 2] & 1)
    return e[np.argsort(cls, kind="stable")]


def restrict_entries(parent_off: int, child_off: int, corner, BS: int, ng: int) -> np.ndarray:
    """Injection of a child's coincident interior nodes into its parent's octant
    (kernels.py:113-140): parent ng + corner BS/2 + a <- child ng + 2 a, a in [0, BS/2)^3."""
    W = BS + 2 * ng
    half = BS // RATIO
    a = _grid_points(half)
    p = ng + np.asarray(corner)[None, :] * half + a
    c = ng + RATIO * a
    e = np.empty(a.shape[0], dtype=HALO_DTYPE)
    e["dst"] = parent_off + _lin(p, W)
    e["src"] = child_off + _lin(c, W)
    e["dst_stride"] = W ** 3
    e["src_stride"] = W ** 3
    return e


def _face_slab(f: int, BS: int, ng: int):
    """Destination and source full-block index ranges of face f's ghost slab (sync.py:100-108):
    my ghost slab along the face axis, interior extent on the other two."""
    ax, hi = f // 2, f % 2
    rng = [(ng, ng + BS)] * 3
    dst, src = list(rng), list(rng)
    if hi == 0:
        dst[ax], src[ax] = (0, ng), (BS, BS + ng)
    else:
        dst[ax], src[ax] = (ng + BS, 2 * ng + BS), (ng, 2 * ng)
    return dst, src


def _box(r):
    return np.stack(np.meshgrid(*[np.arange(a, b) for a, b in r], indexing="ij"), -1).reshape(-1, 3)


def block_job(dst_off: int, parent_off: int, corner, BS: int, ng: int, skip: int) -> np.ndarray:
    e = np.zeros(1, dtype=BLOCK_DTYPE)
    e["dst"], e["parent"], e["corner"] = dst_off, parent_off, corner
    e["skip"], e["BS"], e["ng"] = skip, BS, ng
    return e


def build_ghost_program(topo: AMRTopology, nb_root, ng: int, slot_offset: Callable,
                        use_blocks: bool | None = None):
    """Entries of the hierarchy sync, phase by phase: phase 0 = periodic copies for every ghost
    point of every root block; phase L >= 1 = for every active level-L block, face copies from
    its same-level neighbours and prolongations from its parent for all its other ghost points
    -- one whole-block job per block when the padded extent allows (use_blocks, default when
    W <= BLOCK_WMAX), else one entry per point.
    Returns (prolongs, blocks, copies, prolong_end, block_end, copy_end)."""
    BS = topo.BS
    W = BS + 2 * ng
    P = W ** 3
    if use_blocks is None:
        use_blocks = W <= BLOCK_WMAX
    if use_blocks and W > BLOCK_WMAX:
        raise ValueError(f"whole-block prolongation needs BS + 2 ng <= {BLOCK_WMAX}")
    pts = _grid_points(W)
    ghost = pts[np.any((pts < ng) | (pts >= ng + BS), axis=1)]
    nb = np.asarray(nb_root, dtype=np.int64)
    N = nb * BS
    root_off = -np.ones(tuple(nb), dtype=np.int64)
    for k in topo.blocks(0):
        b = np.asarray(topo.bbox_ijk[k]) // BS
        root_off[tuple(b)] = slot_offset(*k)
    if np.any(root_off < 0) or len(topo.blocks(0)) != int(np.prod(nb)):
        raise ValueError("the root level must tile the box")
    prolongs, jobs, copies, pe, be, ce = [], [], [], [], [], []
    # phase 0: root periodic
    for key in topo.blocks(0):
        b0 = np.asarray(topo.bbox_ijk[key])
        g = (b0[None, :] + ghost - ng) % N[None, :]
        blk = g // BS
        loc = g - blk * BS + ng
        e = np.empty(ghost.shape[0], dtype=HALO_DTYPE)
        e["dst"] = slot_offset(0, key[1]) + _lin(ghost, W)
        e["src"] = root_off[blk[:, 0], blk[:, 1], blk[:, 2]] + _lin(loc, W)
        e["dst_stride"] = P
        e["src_stride"] = P
        copies.append(e)
    pe.append(sum(len(x) for x in prolongs))
    be.append(sum(len(x) for x in jobs))
    ce.append(sum(len(x) for x in copies))
    topo.rebuild_neighbors()
    for L in range(1, topo.levels):
        level_pro = []
        for key in topo.blocks(L):
            s = key[1]
            done = np.zeros((W, W, W), bool)
            done[ng:ng + BS, ng:ng + BS, ng:ng + BS] = True
            skip = 0
            for f in range(NFACE):
                nb = topo.neighbors.get((L, s, f))
                if nb is None:
                    continue
                skip |= 1 << f
                dr, sr = _face_slab(f, BS, ng)
                d, sp = _box(dr), _box(sr)
                e = np.empty(d.shape[0], dtype=HALO_DTYPE)
                e["dst"] = slot_offset(L, s) + _lin(d, W)
                e["src"] = slot_offset(L, nb[1]) + _lin(sp, W)
                e["dst_stride"] = P
                e["src_stride"] = P
                copies.append(e)
                done[d[:, 0], d[:, 1], d[:, 2]] = True
            par = topo.parent[key]
            if use_blocks:
                jobs.append(block_job(slot_offset(L, s), slot_offset(*par),
                                      topo.child_corner(key), BS, ng, skip))
            else:
                rest = pts[~done.reshape(-1)]
                level_pro.append(prolong_entries(rest, slot_offset(L, s), slot_offset(*par),
                                                 topo.child_corner(key), BS, ng))
        if level_pro:
            prolongs.append(by_parity(np.concatenate(level_pro)))
        pe.append(sum(len(x) for x in prolongs))
        be.append(sum(len(x) for x in jobs))
        ce.append(sum(len(x) for x in copies))
    cat = lambda xs, dt: np.concatenate(xs) if xs else np.empty(0, dtype=dt)
    return (cat(prolongs, PROLONG_DTYPE), cat(jobs, BLOCK_DTYPE), cat(copies, HALO_DTYPE),
            pe, be, ce)


class _DeviceProgram:
    """A ghost fill program uploaded to the device (keeps the arrays alive)."""

    def __init__(self, prolongs: np.ndarray, jobs: np.ndarray, copies: np.ndarray, pe, be, ce,
                 device):
        self.n_prolong, self.n_block, self.n_copy = len(prolongs), len(jobs), len(copies)
        up = lambda a: _upload(a, device) if len(a) else None
        self._p, self._b, self._c = up(prolongs), up(jobs), up(copies)
        self._pe = (ctypes.c_longlong * len(pe))(*pe)
        self._be = (ctypes.c_longlong * len(be))(*be)
        self._ce = (ctypes.c_longlong * len(ce))(*ce)
        ptr = lambda t: t.data_ptr() if t is not None else None
        self.struct = _lib.GhostProgram(len(pe), self._pe, self._be, self._ce, ptr(self._p),
                                        ptr(self._b), ptr(self._c))

    @property
    def ptr(self):
        return ctypes.byref(self.struct)


def _upload(entries: np.ndarray, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(entries).view(np.uint8).copy()).to(device)


# --- the hierarchy ----------------------------------------------------------------------------

class AMRHierarchy:
    """A node-centred block hierarchy on a periodic box from lo: a root tiling of nb_root blocks
    of BS^3 nodes at one spacing h0 = L / (nb_x BS) (L = the box's x extent), refined 2:1 by
    octants up to
    ``levels`` levels; every block holds BS^3 interior nodes and ng ghost layers. The state is
    one float64 pool (``empty_pool``); block (L, s) is ``view(U, L, s)``, a (24, W, W, W) state
    (W = BS + 2 ng). Steps are shared-dt RK4 over all active blocks."""

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
        caps = [int(c) for c in caps]
        if caps[0] < nroot or max(caps) > max_blocks or len(caps) != levels:
            raise ValueError("caps: one per level, caps[0] >= root blocks, each <= max_blocks")
        self.h0 = self.L / (self.nb_root[0] * self.BS)       # L is the box's x extent
        self.box = tuple(n * self.BS * self.h0 for n in self.nb_root)
        self.topo = AMRTopology(levels, self.BS, caps, int(max_blocks))
        nbx, nby, nbz = self.nb_root
        for bi in range(nbx):
            for bj in range(nby):
                for bk in range(nbz):
                    self.topo.add_block(0, (bi * nby + bj) * nbz + bk,
                                        (bi * self.BS, bj * self.BS, bk * self.BS))
        self.topo.record_occupancy()
        self._layout()
        self._batch = None
        self._prog = None

    # --- pool layout ----------------------------------------------------------------------
    def _layout(self):
        self.slot_doubles = (NUM_VARS * self.W ** 3 + 63) // 64 * 64
        self.level_base = [0]
        for c in self.topo.caps[:-1]:
            self.level_base.append(self.level_base[-1] + c * self.slot_doubles)
        self.pool_doubles = self.level_base[-1] + self.topo.caps[-1] * self.slot_doubles

    def slot_offset(self, L: int, s: int) -> int:
        if not 0 <= s < self.topo.caps[L]:
            raise IndexError(f"slot {s} outside level {L}'s capacity {self.topo.caps[L]}")
        return self.level_base[L] + s * self.slot_doubles

    def empty_pool(self) -> torch.Tensor:
        return torch.zeros(self.pool_doubles, dtype=torch.float64, device=self.device)

    def view(self, U, L: int, s: int):
        """Block (L, s) of a pool (torch tensor or numpy array) as a (24, W, W, W) view."""
        o = self.slot_offset(L, s)
        blk = U[o:o + NUM_VARS * self.W ** 3]
        shape = (NUM_VARS, self.W, self.W, self.W)
        return blk.view(*shape) if isinstance(blk, torch.Tensor) else blk.reshape(shape)

    def spacing(self, L: int) -> float:
        return self.h0 / (RATIO ** L)

    def block(self, key) -> Block:
        L, _ = key
        h = self.spacing(L)
        b = self.topo.bbox_ijk[key]
        return Block(n=(self.BS,) * 3, h=(h,) * 3,
                     lo=tuple(self.lo[a] + b[a] * h for a in range(3)), level=L)

    def grid(self, key) -> Grid:
        return self.block(key).grid(self.ng)

    # --- plans ----------------------------------------------------------------------------
    def _invalidate(self):
        if self._batch is not None:
            self._batch.close()
        self._batch = None
        self._prog = None

    @property
    def keys(self) -> List[Tuple[int, int]]:
        """Active blocks in batch order (level, then slot)."""
        return self.topo.blocks()

    @property
    def batch(self) -> BlockBatch:
        if self._batch is None:
            keys = self.keys
            self._batch = BlockBatch([self.block(k) for k in keys], ng=self.ng,
                                     device=self.device,
                                     offsets=[self.slot_offset(*k) for k in keys],
                                     pool_doubles=self.pool_doubles)
        return self._batch

    def program_entries(self):
        return build_ghost_program(self.topo, self.nb_root, self.ng, self.slot_offset)

    @property
    def program(self) -> _DeviceProgram:
        if self._prog is None:
            pr, jb, cp, pe, be, ce = self.program_entries()
            self._check_entries(pr, cp, jb)
            self._prog = _DeviceProgram(pr, jb, cp, pe, be, ce, self.device)
        return self._prog

    def _check_entries(self, prolongs: np.ndarray, copies: np.ndarray, jobs=None):
        span = (NUM_VARS - 1) * self.W ** 3
        hi = self.pool_doubles - span
        blk_hi = self.pool_doubles - NUM_VARS * self.W ** 3
        for name, arr, top in (("dst", prolongs["dst"], hi), ("parent", prolongs["parent"], hi),
                               ("dst", copies["dst"], hi), ("src", copies["src"], hi)) + \
                (() if jobs is None else (("job dst", jobs["dst"], blk_hi + 1),
                                          ("job parent", jobs["parent"], blk_hi + 1))):
            if len(arr) and (arr.min() < 0 or arr.max() >= top):
                raise RuntimeError(f"ghost program {name} offset outside the pool")
        if jobs is not None and len(jobs) and (self.W > BLOCK_WMAX or
                                               np.any(jobs["BS"] != self.BS) or
                                               np.any(jobs["ng"] != self.ng)):
            raise RuntimeError("whole-block jobs need this hierarchy's BS, ng and W <= 24")
        if len(prolongs) and (prolongs["fine"].min() < 0 or
                              prolongs["fine"].max() >= 2 * self.W):
            raise RuntimeError("prolongation index outside the refined parent")

    def _check_pool(self, U: torch.Tensor):
        if not isinstance(U, torch.Tensor) or U.dtype != torch.float64 or \
                U.numel() != self.pool_doubles or not U.is_contiguous() or U.device != self.device:
            raise ValueError("U must be this hierarchy's contiguous float64 pool")

    # --- data -----------------------------------------------------------------------------
    def fill(self, U: torch.Tensor, fn: Callable[[Grid], torch.Tensor], sync: bool = True):
        """Every active block <- fn(its Grid) (a (24, W, W, W) state at the block's nodes)."""
        self._check_pool(U)
        for k in self.keys:
            self.view(U, *k).copy_(fn(self.grid(k)))
        if sync:
            self.sync_(U)
        return U

    def from_global(self, U: torch.Tensor, F: torch.Tensor, sync: bool = True):
        """Root interiors <- the interior of a single-grid (24, N+2g, N+2g, N+2g) state (any
        ghost width g), N = nb BS; then the hierarchy is synced."""
        self._check_pool(U)
        g = (F.shape[1] - self.nb_root[0] * self.BS) // 2
        BS, ng = self.BS, self.ng
        for k in self.topo.blocks(0):
            b = self.topo.bbox_ijk[k]
            self.view(U, *k)[:, ng:ng + BS, ng:ng + BS, ng:ng + BS].copy_(
                F[:, g + b[0]:g + b[0] + BS, g + b[1]:g + b[1] + BS, g + b[2]:g + b[2] + BS])
        if sync:
            self.sync_(U)
        return U

    def to_global(self, U: torch.Tensor) -> torch.Tensor:
        """The root level's interiors as one (24, N0, N1, N2) array."""
        BS, ng = self.BS, self.ng
        N = [n * BS for n in self.nb_root]
        out = torch.empty((NUM_VARS, *N), dtype=U.dtype, device=U.device)
        for k in self.topo.blocks(0):
            b = self.topo.bbox_ijk[k]
            out[:, b[0]:b[0] + BS, b[1]:b[1] + BS, b[2]:b[2] + BS] = \
                self.view(U, *k)[:, ng:ng + BS, ng:ng + BS, ng:ng + BS]
        return out

    # --- operators ------------------------------------------------------------------------
    def sync_(self, U: torch.Tensor) -> torch.Tensor:
        self._check_pool(U)
        _lib.check(_lib.load().bssn_ghost_fill_f64(U.data_ptr(), self.program.ptr,
                                                   stream_handle(self.device)))
        return U

    def workspace(self) -> torch.Tensor:
        return self.batch.workspace()

    def step_(self, U: torch.Tensor, t: float, dt: float, params: PhysicsParams | None = None,
              order: int = 4, ko_sigma: float = 0.1, tdev: torch.Tensor | None = None):
        """One shared-dt RK4 step of every active block in place (U synced on entry and on
        exit)."""
        params = params or PhysicsParams()
        self._check_pool(U)
        b = self.batch
        prog = self.program
        tptr = None
        if tdev is not None:
            if tdev.dtype != torch.float64 or tdev.device != self.device or tdev.numel() != 1:
                raise ValueError("tdev must be one float64 value on the hierarchy's device")
            tptr = tdev.data_ptr()
        _lib.check(_lib.load().bssn_batch_rk4_step_prog_f64(
            b._handle, U.data_ptr(), self.workspace().data_ptr(),
            _lib.dbl(b.S16(params, dt, t)), int(order), float(ko_sigma), prog.ptr, float(t),
            float(dt), params.chi_floor, params.alpha_floor, tptr, stream_handle(self.device)))
        return U

    def capture_steps(self, U: torch.Tensor, tdev: torch.Tensor, dt: float, nsteps: int = 1,
                      **kw) -> torch.cuda.CUDAGraph:
        """Capture ``nsteps`` hierarchy steps of U into a HIP graph (time in tdev, advanced by
        each replay). The topology must not change while the graph is in use."""
        self.workspace()
        prog = self.program
        warm = U.clone()                  # first launches (code loading) outside the capture
        self.step_(warm, 0.0, dt, **kw)
        torch.cuda.synchronize(self.device)
        del warm
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

    def indicators(self, U: torch.Tensor, fields: Sequence[int] | None = None) -> List[np.ndarray]:
        """Per level, per slot: the largest undivided difference over the chosen fields (all by
        default) of each active block, 0 for idle slots."""
        self._check_pool(U)
        keys = self.keys
        out = torch.empty(len(keys) * NUM_VARS, dtype=torch.float64, device=self.device)
        _lib.check(_lib.load().bssn_batch_indicator_f64(self.batch._handle, U.data_ptr(),
                                                        out.data_ptr(),
                                                        stream_handle(self.device)))
        v = out.view(len(keys), NUM_VARS)
        if fields is not None:
            v = v[:, list(fields)]
        v = v.max(dim=1).values.cpu().numpy()
        res = [np.zeros(self.topo.max_blocks) for _ in range(self.levels)]
        for (L, s), x in zip(keys, v):
            res[L][s] = x
        return res

    # --- regrid ---------------------------------------------------------------------------
    def _grow(self, U: torch.Tensor, L: int) -> torch.Tensor:
        """Enlarge level L's slot capacity (regrid.py:283-308): a new pool, levels copied."""
        old = self.topo.caps[L]
        new = min(max(old + max(8, old // 2), old + 1), self.topo.max_blocks)
        if new <= old:
            raise RuntimeError(f"AMR slot ceiling reached at level {L}: {old} == max_blocks")
        old_base, old_caps = list(self.level_base), list(self.topo.caps)
        self.topo.caps[L] = new
        self.topo.grow_history.append((L, old, new))
        self._layout()
        V = self.empty_pool()
        for l in range(self.levels):
            n = old_caps[l] * self.slot_doubles
            V[self.level_base[l]:self.level_base[l] + n] = U[old_base[l]:old_base[l] + n]
        self._invalidate()
        return V

    def apply_flags(self, U: torch.Tensor, flags: np.ndarray, auto_grow: bool = True) -> torch.Tensor:
        """Coarsen (finest level first: restrict into the parent, free the slot), then refine
        (coarsest first: up to 8 children per flagged block, each prolongated from its parent
        over the whole padded block) (regrid.py:255-420). Returns the pool (a new one if a
        level grew); the hierarchy is synced."""
        topo, BS, ng = self.topo, self.BS, self.ng
        lib = _lib.load()
        st = stream_handle(self.device)
        changed = False
        for L in range(self.levels - 1, 0, -1):
            ents = []                     # this level's restrictions: one copy launch
            for (_, s) in topo.blocks(L):
                key = (L, s)
                if flags[L, s] != COARSEN or key not in topo.parent:
                    continue
                par = topo.parent[key]
                ents.append(restrict_entries(self.slot_offset(*par), self.slot_offset(L, s),
                                             topo.child_corner(key), BS, ng))
                topo.remove_block(L, s)
                changed = True
            if ents:
                e = np.concatenate(ents)
                d = _upload(e, self.device)
                _lib.check(lib.bssn_batch_halo_f64(U.data_ptr(), d.data_ptr(), len(e), st))
        pts = _grid_points(self.W)
        for L in range(self.levels - 1):
            cl = L + 1
            new = []                      # this level's new children: one prolongation launch
            for (_, s) in topo.blocks(L):
                if flags[L, s] != REFINE or not topo.active[L, s]:
                    continue
                pb = topo.bbox_ijk[(L, s)]
                existing = {topo.bbox_ijk[k] for k in topo.blocks(cl)}
                for cx in (0, 1):
                    for cy in (0, 1):
                        for cz in (0, 1):
                            c = (cx, cy, cz)
                            cb = tuple(pb[a] * RATIO + c[a] * BS for a in range(3))
                            if cb in existing:
                                continue
                            free = topo.free_slots(cl)
                            if len(free) == 0:
                                if not auto_grow:
                                    raise RuntimeError(f"AMR slots exhausted at level {cl}")
                                torch.cuda.synchronize(self.device)
                                U = self._grow(U, cl)
                                free = topo.free_slots(cl)
                            slot = int(free[0])
                            topo.add_block(cl, slot, cb, parent=(L, s))
                            new.append((slot, s, c))
                            changed = True
            if new:        # offsets after any growth; parents are level L, children level L+1
                if self.W <= BLOCK_WMAX:
                    e = np.concatenate([block_job(self.slot_offset(cl, slot),
                                                  self.slot_offset(L, ps), c, BS, ng,
                                                  SKIP_INTERIOR_WRITE) for slot, ps, c in new])
                    d = _upload(e, self.device)
                    _lib.check(lib.bssn_prolong_blocks_f64(U.data_ptr(), d.data_ptr(), len(e),
                                                           st))
                else:
                    e = by_parity(np.concatenate([
                        prolong_entries(pts, self.slot_offset(cl, slot), self.slot_offset(L, ps),
                                        c, BS, ng) for slot, ps, c in new]))
                    d = _upload(e, self.device)
                    _lib.check(lib.bssn_prolong_f64(U.data_ptr(), d.data_ptr(), len(e), st))
        if changed:
            self._invalidate()
            topo.rebuild_neighbors()
            self.sync_(U)
        topo.record_occupancy()
        return U

    def regrid(self, U: torch.Tensor, *, refine_threshold: float, coarsen_threshold: float,
               hysteresis_K: int = 3, n_buffer: int = 1, max_level: Optional[int] = None,
               fields: Sequence[int] | None = None) -> torch.Tensor:
        """One regrid pass: indicators -> flags -> nesting/buffer -> apply (regrid.py:425-460).
        Returns the pool (a new one if a level grew)."""
        ind = self.indicators(U, fields)
        flags = compute_flags(ind, self.topo, refine_threshold=refine_threshold,
                              coarsen_threshold=coarsen_threshold, hysteresis_K=hysteresis_K,
                              max_level=max_level)
        flags = enforce_nesting_buffer(flags, self.topo, n_buffer=n_buffer)
        return self.apply_flags(U, flags)

    def evolve_with_regrid(self, U: torch.Tensor, t: float, dt: float, n_steps: int,
                           regrid_every: int, *, refine_threshold: float,
                           coarsen_threshold: float, hysteresis_K: int = 3, n_buffer: int = 1,
                           max_level: Optional[int] = None, fields: Sequence[int] | None = None,
                           on_regrid=None, **step_kw):
        """n_steps shared-dt steps with a regrid every regrid_every steps (none after the last
        chunk) (regrid.py:463-538). Returns (U, t)."""
        done = 0
        while done < n_steps:
            chunk = min(regrid_every, n_steps - done)
            t = self.evolve(U, t, dt, chunk, **step_kw)
            done += chunk
            if done < n_steps:
                U = self.regrid(U, refine_threshold=refine_threshold,
                                coarsen_threshold=coarsen_threshold, hysteresis_K=hysteresis_K,
                                n_buffer=n_buffer, max_level=max_level, fields=fields)
                if on_regrid is not None:
                    on_regrid(done, U, self)
        return U, t
