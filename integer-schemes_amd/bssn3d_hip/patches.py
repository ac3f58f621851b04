"""A nested patch hierarchy of uneven blocks: the batched block-AMR workload.

Each refinement level is one box on its own node lattice (spacing h0 / 2^L), cut into blocks by
independent segmentations of its three axes: a block's interior extent along each axis is one
segment, so blocks have uneven, non-cubic extents (each drawn from a set of sizes), and the
blocks of a level tile its box exactly. Level L + 1's box lies inside level L's with a margin
of coarse nodes on every side, at even level-(L + 1) node indices, so the 2:1 lattices nest.
Level 0 is periodic.

Ghost points are filled by one ghost fill program (bssn_ghost_fill_f64), a phase per level:
  level 0   every ghost point is copied from the block owning its periodic image;
  level L   a ghost point inside the level's box is copied from the same-level block that
            owns it (faces, edges and corners alike), a point outside the box is prolongated
            from the level-(L - 1) block whose interior holds its coarse node floor(p / 2)
            (the 2:1 node-centred tensor-product prolongation of bssn_prolong_f64, reading the
            parent's padded block, whose ghosts the previous phase filled).
The reference's shared-dt N-level step syncs the same way -- root periodic, each finer level
from its parent, then from same-level neighbours (2D/src/mcs2d/amr/evolve.py:719-735,
3D/src/multipatch/amr/sync.py:84-122) -- for equal blocks in fixed slot arrays
(amr/state.py:27-31); here the blocks are uneven and all of them step in one batched launch per
RK stage (bssn_batch_rk4_step_prog_f64), with the program after every stage.

Linked steps (bssn_batch_rk4_step_link_f64, ``step_(..., ghosts="linked")``) skip the copies: a
link map sends each copied ghost point to its source, which the stage kernels read in place,
and after every stage only the prolongations run, with the copies into the parent ghosts that
their taps read (``linked_program``). Interior values equal the program step's bit for bit.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .amr import (BLOCK_DTYPE, PROLONG_DTYPE, DeviceProgram, by_parity, link_map,
                  linked_entries)
from .batch import HALO_DTYPE, Block, BlockBatch, BlockLayout, tune_step_launch
from .rhs import stream_handle
from .state import PhysicsParams

SIZES = (8, 12, 16, 24, 32)


@dataclass(frozen=True)
class LevelBox:
    """One level: spacing h, the level node index of the box's first interior node, and the
    segment lengths cutting each axis (the box extent is their sum)."""
    h: float
    origin: Tuple[int, int, int]
    cuts: Tuple[Tuple[int, ...], Tuple[int, ...], Tuple[int, ...]]

    @property
    def extent(self) -> Tuple[int, int, int]:
        return tuple(int(sum(c)) for c in self.cuts)

    @property
    def nblocks(self) -> int:
        return int(np.prod([len(c) for c in self.cuts]))


def _starts(cuts: Sequence[int]) -> np.ndarray:
    return np.concatenate([[0], np.cumsum(cuts)[:-1]]).astype(np.int64)


class PatchHierarchy:
    """Blocks of nested level boxes in one pool (a BlockBatch), with the ghost fill program that
    syncs them and the shared-dt RK4 step over all of them.

    boxes: the levels, coarsest first; x0: physical coordinates of level node 0 (common to all
    levels: node p of level L sits at x0 + p h_L); ng: ghost width."""

    def __init__(self, boxes: Sequence[LevelBox], x0: Sequence[float], ng: int = 3,
                 device="cuda", margin: int = 6):
        self.boxes = list(boxes)
        self.x0 = tuple(float(v) for v in x0)
        self.ng = int(ng)
        for L in range(1, len(self.boxes)):
            self._check_nesting(L, margin)
        blocks, self.level, self.node0 = [], [], []
        for L, box in enumerate(self.boxes):
            st = [_starts(c) for c in box.cuts]
            for i, j, k in itertools.product(*[range(len(c)) for c in box.cuts]):
                n = (box.cuts[0][i], box.cuts[1][j], box.cuts[2][k])
                o = (box.origin[0] + int(st[0][i]), box.origin[1] + int(st[1][j]),
                     box.origin[2] + int(st[2][k]))
                lo = tuple(self.x0[a] + o[a] * box.h for a in range(3))
                blocks.append(Block(n, (box.h,) * 3, lo, L))
                self.level.append(L)
                self.node0.append(o)
        self.layout = BlockLayout(blocks, ng=self.ng)
        self._device = device
        self._batch = None
        self._prog = None
        self._linked = None
        self._link = None
        self.tiles = True

    @property
    def batch(self) -> BlockBatch:
        """The device plan of the blocks (created on first use)."""
        if self._batch is None:
            self._batch = BlockBatch(self.layout.blocks, ng=self.ng, device=self._device)
        return self._batch

    @property
    def device(self) -> torch.device:
        return self.batch.device

    # geometry ------------------------------------------------------------------------------
    def _check_nesting(self, L: int, margin: int):
        fine, coarse = self.boxes[L], self.boxes[L - 1]
        if abs(fine.h * 2 - coarse.h) > 1e-15 * coarse.h:
            raise ValueError(f"level {L}: spacing must be half of level {L - 1}'s")
        for a in range(3):
            if fine.origin[a] % 2 or fine.extent[a] % 2:
                raise ValueError(f"level {L}: origin and extent must be even (2:1 nodes)")
            lo = fine.origin[a] // 2
            hi = (fine.origin[a] + fine.extent[a]) // 2
            if lo - coarse.origin[a] < margin or coarse.origin[a] + coarse.extent[a] - hi < margin:
                raise ValueError(f"level {L} is not nested in level {L - 1} with a margin of "
                                 f"{margin} coarse nodes along axis {a}")

    @property
    def blocks(self) -> List[Block]:
        return self.layout.blocks

    def __len__(self) -> int:
        return len(self.layout.blocks)

    def owner_map(self, L: int) -> np.ndarray:
        """owner[i, j, k] = index of the block whose interior holds level-L box node
        origin + (i, j, k)."""
        box = self.boxes[L]
        own = np.empty(box.extent, dtype=np.int64)
        first = self.level.index(L)
        st = [_starts(c) for c in box.cuts]
        b = first
        for i, j, k in itertools.product(*[range(len(c)) for c in box.cuts]):
            own[st[0][i]:st[0][i] + box.cuts[0][i], st[1][j]:st[1][j] + box.cuts[1][j],
                st[2][k]:st[2][k] + box.cuts[2][k]] = b
            b += 1
        return own

    # the ghost fill program ----------------------------------------------------------------
    def program_entries(self):
        """(prolongs, jobs, copies, prolong_end, job_end, copy_end) of the hierarchy sync."""
        ng = self.ng
        lay = self.layout
        node0 = np.asarray(self.node0, dtype=np.int64)
        off = np.asarray(lay.off, dtype=np.int64)
        Sb = np.asarray(lay.S, dtype=np.int64)
        Pb = np.asarray(lay.P, dtype=np.int64)
        level = np.asarray(self.level)
        owners = [self.owner_map(L) for L in range(len(self.boxes))]
        pro_lists, cop_lists = [], []
        ends = ([], [], [])

        def flat(blk, loc):
            return off[blk] + (loc[:, 0] * Sb[blk, 1] + loc[:, 1]) * Sb[blk, 2] + loc[:, 2]

        for L, box in enumerate(self.boxes):
            E = np.asarray(box.extent)
            O = np.asarray(box.origin)
            pro, cop = [], []
            for b in np.flatnonzero(level == L):
                S = Sb[b]
                c = np.stack(np.meshgrid(*[np.arange(v) for v in S], indexing="ij"),
                             -1).reshape(-1, 3)
                c = c[np.any((c < ng) | (c >= S - ng), axis=1)]
                p = node0[b][None, :] + c - ng                 # level node index of each ghost
                dst = off[b] + (c[:, 0] * S[1] + c[:, 1]) * S[2] + c[:, 2]
                if L == 0:
                    q = (p - O) % E
                    inside = np.ones(len(p), bool)
                else:
                    q = p - O
                    inside = np.all((q >= 0) & (q < E), axis=1)
                if inside.any():
                    qi = q[inside]
                    k = owners[L][qi[:, 0], qi[:, 1], qi[:, 2]]
                    e = np.empty(len(k), dtype=HALO_DTYPE)
                    e["dst"] = dst[inside]
                    e["dst_stride"] = Pb[b]
                    e["src"] = flat(k, O + qi - node0[k] + ng)
                    e["src_stride"] = Pb[k]
                    cop.append(e)
                if (~inside).any():
                    po = p[~inside]
                    pbox = self.boxes[L - 1]
                    kc = np.floor_divide(po, 2) - np.asarray(pbox.origin)
                    if np.any(kc < 0) or np.any(kc >= np.asarray(pbox.extent)):
                        raise RuntimeError(f"level {L}: a ghost's coarse node is outside level "
                                           f"{L - 1}")
                    par = owners[L - 1][kc[:, 0], kc[:, 1], kc[:, 2]]
                    e = np.empty(len(par), dtype=PROLONG_DTYPE)
                    e["dst"] = dst[~inside]
                    e["dst_stride"] = Pb[b]
                    e["parent"] = off[par]
                    e["fine"] = po - 2 * (node0[par] - ng)
                    e["pS"] = Sb[par]
                    pro.append(e)
            if pro:
                pro_lists.append(by_parity(np.concatenate(pro)))
            cop_lists.extend(cop)
            ends[0].append(sum(len(x) for x in pro_lists))
            ends[1].append(0)
            ends[2].append(sum(len(x) for x in cop_lists))
        cat = lambda xs, dt: np.concatenate(xs) if xs else np.empty(0, dtype=dt)
        return (cat(pro_lists, PROLONG_DTYPE), np.empty(0, dtype=BLOCK_DTYPE),
                cat(cop_lists, HALO_DTYPE), *ends)

    def _device_program(self, entries) -> DeviceProgram:
        pr, jb, cp, pe, be, ce = entries
        prog = DeviceProgram(pr, jb, cp, pe, be, ce, self.device, tiles=self.tiles)
        lib = _lib.load()
        pool = self.layout.pool_doubles
        for kind, arr in ((1, pr), (0, cp), (3, prog.tiles), (4, prog.tile_points)):
            if len(arr) and lib.bssn_count_skipped_entries(
                    kind, np.ascontiguousarray(arr).ctypes.data, len(arr), pool) != 0:
                raise RuntimeError("ghost program entry outside the pool")
        return prog

    @property
    def program(self) -> DeviceProgram:
        """The ghost fill program on the device (prolongations grouped by 4x4x4 fine tile
        unless ``tiles`` is False)."""
        if self._prog is None:
            self._prog = self._device_program(self.program_entries())
        return self._prog

    # linked steps --------------------------------------------------------------------------
    def link_offsets(self) -> np.ndarray:
        """m_b: first padded point of block b in the link map (P_0 + ... + P_(b-1))."""
        P = np.asarray(self.layout.P, dtype=np.int64)
        return np.concatenate([[0], np.cumsum(P)[:-1]])

    def link_map(self, copies: np.ndarray | None = None) -> np.ndarray:
        """(n_points, 2) int32: for every padded point of every block (block order, each block
        row-major), the pool offset of field 0 of the point its value is read from and that
        block's field stride -- the copy entry's source for a copied ghost, the point itself
        otherwise (interior points, prolongated ghosts)."""
        if copies is None:
            copies = self.program_entries()[2]
        return link_map(self.layout.off, self.layout.P, copies, self.layout.pool_doubles)

    def linked_program_entries(self):
        """The program of a linked step: every prolongation, and only the copies whose
        destination a prolongation reads, in the program's phases."""
        return linked_entries(self.program_entries(), self.ng)

    @property
    def linked_program(self) -> DeviceProgram:
        if self._linked is None:
            self._linked = self._device_program(self.linked_program_entries())
        return self._linked

    @property
    def link(self) -> torch.Tensor:
        """The link map on the device."""
        if self._link is None:
            self._link = torch.from_numpy(self.link_map()).to(self.device)
        return self._link

    # operators -----------------------------------------------------------------------------
    def empty_pool(self) -> torch.Tensor:
        return self.batch.empty_pool()

    def view(self, U, b: int):
        return self.layout.view(U, b)

    def sync_(self, U: torch.Tensor) -> torch.Tensor:
        self.batch._check_pool(U)
        _lib.check(_lib.load().bssn_ghost_fill_f64(U.data_ptr(), U.numel(), self.program.ptr,
                                                   stream_handle(self.device)))
        return U

    def step_(self, U: torch.Tensor, t: float, dt: float, params: PhysicsParams | None = None,
              order: int = 4, ko_sigma: float = 0.1, tdev: torch.Tensor | None = None,
              ghosts: str = "program", sync_end: bool = True):
        """One shared-dt RK4 step of every block in place; U must be synced on entry (every
        step leaves it synced). ghosts="linked": stages read copied ghosts at their sources
        (bssn_batch_rk4_step_link_f64); with sync_end=False the step leaves only the ghosts the
        next linked step reads (prolongated ones and the parents' ghosts they need) up to
        date."""
        params = params or PhysicsParams()
        b = self.batch
        b._check_pool(U)
        tptr = None
        if tdev is not None:
            if tdev.dtype != torch.float64 or tdev.device != self.device or tdev.numel() != 1:
                raise ValueError("tdev must be one float64 value on the hierarchy's device")
            tptr = tdev.data_ptr()
        if ghosts == "linked":
            _lib.check(_lib.load().bssn_batch_rk4_step_link_f64(
                b._handle, U.data_ptr(), b.workspace().data_ptr(),
                _lib.dbl(b.S16(params, dt, t)), int(order), float(ko_sigma),
                self.link.data_ptr(), self.link.shape[0], self.linked_program.ptr,
                self.program.ptr if sync_end else None, float(t), float(dt), params.chi_floor,
                params.alpha_floor, tptr, stream_handle(self.device)))
            return U
        if ghosts != "program":
            raise ValueError(f"ghosts must be 'program' or 'linked', not {ghosts!r}")
        _lib.check(_lib.load().bssn_batch_rk4_step_prog_f64(
            b._handle, U.data_ptr(), b.workspace().data_ptr(), _lib.dbl(b.S16(params, dt, t)),
            int(order), float(ko_sigma), self.program.ptr, float(t), float(dt),
            params.chi_floor, params.alpha_floor, tptr, stream_handle(self.device)))
        return U

    def capture_steps(self, U: torch.Tensor, tdev: torch.Tensor, dt: float, nsteps: int = 1,
                      **kw) -> torch.cuda.CUDAGraph:
        """Capture ``nsteps`` steps of U into a HIP graph (time in tdev, advanced by each
        replay)."""
        self.batch.workspace()
        prog = self.program
        if kw.get("ghosts") == "linked":
            prog = (prog, self.linked_program, self.link)
        scratch = U.clone()
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
        self._graph_keep = prog
        return g

    def tune_launch(self, U: torch.Tensor, dt: float, candidates=None, reps: int = 3,
                    **kw) -> dict:
        """Time one step per batched launch shape on a copy of U and keep the fastest (the first
        candidate unless another beats it by more than the timing noise, batch.tune_step_launch);
        ``kw`` goes to step_, so ghosts="linked" tunes the linked step. Returns {shape: ms}
        (0 = one workgroup per brick, N = N persistent waves)."""
        return tune_step_launch(self.batch, lambda V: self.step_(V, 0.0, dt, **kw), U,
                                candidates, reps)


# --- the workload -----------------------------------------------------------------------------

def draw_cuts(rng: np.random.Generator, count: int, sizes=SIZES) -> Tuple[int, ...]:
    return tuple(int(v) for v in rng.choice(sizes, size=count))


def binary_positions(masses=(0.909, 0.0909), separation: float = 4.0):
    """Two punctures on the x axis, ``separation`` apart about their centre of mass."""
    M = float(sum(masses))
    return [[separation * masses[1] / M, 0.0, 0.0], [-separation * masses[0] / M, 0.0, 0.0]]


def config_hierarchy(seed: int = 0, counts=((8, 8, 4), (8, 6, 4), (8, 6, 4)), h0: float = 0.25,
                     sizes=SIZES, masses=(0.909, 0.0909), separation: float = 4.0,
                     margin: int = 6, ng: int = 3, device="cuda", hold_punctures: bool = True):
    """The batched block-AMR workload: three levels (h0, h0/2, h0/4 = 1/4, 1/8, 1/16 by
    default) around a 1:10 puncture binary 4M apart, each level's box cut per axis into
    ``counts[L]`` segments drawn by numpy default_rng(seed) from ``sizes`` (640 blocks for the
    defaults). Level 0 is centred on the punctures' midpoint; each finer box is centred on it
    too, and redrawn until it nests in its parent with ``margin`` coarse nodes to spare and
    (finest level, unless hold_punctures is false) holds both punctures 8 nodes from its faces.
    Returns (PatchHierarchy, puncture positions)."""
    rng = np.random.default_rng(seed)
    pos = binary_positions(masses, separation)
    mid = 0.5 * (np.asarray(pos[0]) + np.asarray(pos[1]))
    boxes = []
    cuts0 = tuple(draw_cuts(rng, counts[0][a], sizes) for a in range(3))
    E0 = [sum(c) for c in cuts0]
    # level-0 node 0 sits at x0; the box [0, E0) is centred on the midpoint
    x0 = [float(mid[a] - 0.5 * E0[a] * h0) for a in range(3)]
    boxes.append(LevelBox(h0, (0, 0, 0), cuts0))
    for L in range(1, len(counts)):
        h = h0 / 2 ** L
        par = boxes[-1]
        for _ in range(1000):
            cuts = tuple(draw_cuts(rng, counts[L][a], sizes) for a in range(3))
            E = [sum(c) for c in cuts]
            # centre node on this level, then the box start rounded to an even node
            o = [2 * int(round(((mid[a] - x0[a]) / h - 0.5 * E[a]) / 2.0)) for a in range(3)]
            ok = all(o[a] // 2 - par.origin[a] >= margin and
                     par.origin[a] + par.extent[a] - (o[a] + E[a]) // 2 >= margin
                     for a in range(3))
            if ok and hold_punctures and L == len(counts) - 1:
                for q in pos:
                    n = [(q[a] - x0[a]) / h for a in range(3)]
                    ok &= all(o[a] + 8 <= n[a] <= o[a] + E[a] - 8 for a in range(3))
            if ok:
                break
        else:
            raise RuntimeError(f"no nested draw for level {L}")
        boxes.append(LevelBox(h, tuple(o), cuts))
    return PatchHierarchy(boxes, x0, ng=ng, device=device, margin=margin), pos


def fill_binary(H: PatchHierarchy, U: torch.Tensor, masses=(0.909, 0.0909), positions=None,
                amp: float = 1e-4, seed: int = 0, sync: bool = True) -> torch.Tensor:
    """Brill-Lindquist data of the binary plus the seeded perturbations of
    initial_data.punctures (the same modes on every block, so the data are one smooth function
    across blocks), then the hierarchy sync."""
    from .initial_data import punctures
    positions = positions if positions is not None else binary_positions(masses)
    root = H.boxes[0]
    domain = (H.x0, tuple(root.extent[a] * root.h for a in range(3)))
    H.batch.fill(U, lambda b, blk, grid: punctures(grid, list(masses), positions, amp=amp,
                                                   seed=seed + 1000, device=U.device,
                                                   domain=domain))
    if sync:
        H.sync_(U)
    return U
