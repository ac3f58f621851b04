"""Batched block form: many blocks of different extents and spacings in one pool, every block
evaluated by one launch per RK stage (C ABI: the bssn_batch_* entry points of include/bssn_hip.h).

Reference counterpart: the slot-batched block-AMR step make_n_level_step
(2D/src/mcs2d/amr/evolve.py:659-775). There, each level is a fixed-shape slot array
(mb_L, NF, BS+2NG, BS+2NG) (amr/state.py:27-31), the RHS is vmapped over all slots, inactive
ones included, and each RK stage is RHS -> interior axpy -> halo sync (:745-771). Here a block
has any extent; the device walks a list of 4x4x4 bricks over the blocks that exist, either one
workgroup per brick or with persistent waves pulling bricks from per-XCD queues (the launch
shape is chosen by ``BlockBatch.autotune``); a step can be captured into a HIP graph and
replayed with its time in device memory.

Ghost handling per step (``bc``):
  "edge"     each block edge-padded (= the single-grid step with bc != "periodic")
  "periodic" each block periodic on its own interior (= the single-grid periodic step)
  "halo"     ghosts are data: the stage inputs start with U's ghosts (held, Dirichlet data from
             the last sync), and an optional halo copy map (HaloMap) refills them from other
             blocks after every stage, the reference's same-level neighbour sync
             (amr/evolve.py:719-735); prolongation from a coarser parent is not part of this
             form.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Callable, Iterable, Sequence

import numpy as np
import torch

from . import _lib
from .grid import Grid
from .rhs import stream_handle
from .state import NUM_VARS, PhysicsParams, scalars

BC = {"edge": 0, "periodic": 1, "halo": 2}

# one entry of a halo copy map (bssn_halo_copy): pool offsets of field 0 and the field strides
HALO_DTYPE = np.dtype([("dst", "<i8"), ("src", "<i8"), ("dst_stride", "<i8"),
                       ("src_stride", "<i8")])


def _device(device) -> torch.device:
    """A device with its index (so that "cuda" and "cuda:0" compare equal)."""
    d = torch.device(device)
    if d.type == "cuda" and d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    return d


@dataclass(frozen=True)
class Block:
    """One block: interior extents n, spacings h, coordinates lo of its first interior point,
    and (for hierarchies) the refinement level it belongs to."""
    n: tuple
    h: tuple
    lo: tuple = (0.0, 0.0, 0.0)
    level: int = 0

    def grid(self, ng: int) -> Grid:
        return Grid(self.n[0], self.n[1], self.n[2], self.h[0], self.h[1], self.h[2], ng=ng,
                    xmin=self.lo[0], ymin=self.lo[1], zmin=self.lo[2])


class BlockLayout:
    """Pool geometry of a set of blocks (host only).

    Block b occupies pool[off[b] : off[b] + 24 P_b] as a (24, S0, S1, S2) state, S = n + 2 ng;
    offsets are aligned to ``align`` doubles, or given (``offsets`` with ``pool_doubles``: a
    hierarchy whose blocks live in fixed slots)."""

    def __init__(self, blocks: Sequence[Block], ng: int = 3, align: int = 64,
                 offsets: Sequence[int] | None = None, pool_doubles: int | None = None):
        if not blocks:
            raise ValueError("a batch needs at least one block")
        self.blocks = list(blocks)
        self.ng = int(ng)
        self.S = [tuple(int(v) + 2 * self.ng for v in b.n) for b in self.blocks]
        self.P = [s[0] * s[1] * s[2] for s in self.S]
        if offsets is not None:
            if pool_doubles is None or len(offsets) != len(self.blocks):
                raise ValueError("explicit offsets need one per block and pool_doubles")
            self.off = [int(o) for o in offsets]
            self.pool_doubles = int(pool_doubles)
            return
        off, pos = [], 0
        for P in self.P:
            off.append(pos)
            pos += NUM_VARS * P
            pos = (pos + align - 1) // align * align
        self.off = off
        self.pool_doubles = pos

    def view(self, pool, b: int):
        """Block b of a pool (torch tensor or numpy array) as a (24, S0, S1, S2) view."""
        o = self.off[b]
        blk = pool[o:o + NUM_VARS * self.P[b]]
        return blk.view(NUM_VARS, *self.S[b]) if isinstance(blk, torch.Tensor) else \
            blk.reshape(NUM_VARS, *self.S[b])


class BlockBatch(BlockLayout):
    """A pool of blocks and its device plan (block table and brick lists)."""

    def __init__(self, blocks: Sequence[Block], ng: int = 3, device="cuda", align: int = 64,
                 offsets: Sequence[int] | None = None, pool_doubles: int | None = None):
        super().__init__(blocks, ng, align, offsets, pool_doubles)
        self.device = _device(device)
        off = self.off
        n = len(self.blocks)
        self._S = (ctypes.c_int * (3 * n))(*[v for s in self.S for v in s])
        self._off = (ctypes.c_longlong * n)(*off)
        self._h = (ctypes.c_double * (3 * n))(*[float(v) for b in self.blocks for v in b.h])
        lib = _lib.load()
        nbytes = lib.bssn_batch_plan_bytes(n, self._S, self.ng)
        if nbytes == 0:
            raise ValueError("invalid block extents for this ghost width")
        self._plan = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        handle = ctypes.c_void_p()
        _lib.check(lib.bssn_batch_create(ctypes.byref(handle), self._plan.data_ptr(), nbytes, n,
                                         self._off, self._S, self._h, self.ng, self.pool_doubles,
                                         stream_handle(self.device)))
        self._handle = handle
        self._work = None
        self._int8_buf = None
        self.persistent_waves = 0

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            _lib.load().bssn_batch_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- pools ----------------------------------------------------------------------------
    @property
    def interior_points(self) -> int:
        return int(_lib.load().bssn_batch_points(self._handle, 1))

    @property
    def padded_points(self) -> int:
        return int(_lib.load().bssn_batch_points(self._handle, 0))

    def empty_pool(self) -> torch.Tensor:
        return torch.zeros(self.pool_doubles, dtype=torch.float64, device=self.device)

    def fill(self, pool: torch.Tensor, fn: Callable[[int, Block, Grid], torch.Tensor]) -> None:
        for b, blk in enumerate(self.blocks):
            self.view(pool, b).copy_(fn(b, blk, blk.grid(self.ng)))

    def _check_pool(self, *pools):
        for p in pools:
            if not isinstance(p, torch.Tensor) or p.dtype != torch.float64 or \
                    p.numel() < self.pool_doubles or not p.is_contiguous() or p.device != self.device:
                raise ValueError("pool must be a contiguous float64 tensor of pool_doubles on the "
                                 "batch's device")

    # --- launch shape ---------------------------------------------------------------------
    def set_launch(self, persistent_waves: int) -> None:
        _lib.check(_lib.load().bssn_batch_set_launch(self._handle, int(persistent_waves)))
        self.persistent_waves = int(persistent_waves)

    # --- operators ------------------------------------------------------------------------
    @staticmethod
    def S16(params: PhysicsParams, dt: float, t: float):
        # the spacing entries are replaced per block on the device
        return scalars(params, 1.0, 1.0, 1.0, dt, t)

    def rhs_into(self, F: torch.Tensor, OUT: torch.Tensor, params: PhysicsParams, dt: float,
                 t: float = 0.0, order: int = 4, ko_sigma: float = 0.0) -> None:
        """OUT <- RHS (+ KO) of every block at every padded point (blocks edge-padded)."""
        self._check_pool(F, OUT)
        _lib.check(_lib.load().bssn_batch_rhs_f64(
            self._handle, F.data_ptr(), _lib.dbl(self.S16(params, dt, t)), OUT.data_ptr(),
            int(order), float(ko_sigma), stream_handle(self.device)))

    def workspace(self) -> torch.Tensor:
        n = _lib.load().bssn_batch_step_workspace_bytes(self._handle, 0) // 8
        if self._work is None or self._work.numel() < n:
            self._work = torch.empty(n, dtype=torch.float64, device=self.device)
        return self._work

    def step_(self, U: torch.Tensor, t: float, dt: float, params: PhysicsParams | None = None,
              order: int = 4, ko_sigma: float = 0.1, bc: str = "periodic",
              halo: "HaloMap | None" = None, tdev: torch.Tensor | None = None) -> torch.Tensor:
        """One RK4 step of every block, in place. ``tdev`` (a 1-element float64 device tensor)
        makes the step read its time there and advance it by dt (for graph replay)."""
        params = params or PhysicsParams()
        if bc not in BC:
            raise ValueError(f"bc must be one of {sorted(BC)}")
        self._check_pool(U)
        hptr, hn = (halo.data.data_ptr(), halo.n) if halo is not None and bc == "halo" else (None, 0)
        tptr = None
        if tdev is not None:
            if tdev.dtype != torch.float64 or tdev.device != self.device or tdev.numel() != 1:
                raise ValueError("tdev must be one float64 value on the batch's device")
            tptr = tdev.data_ptr()
        _lib.check(_lib.load().bssn_batch_rk4_step_f64(
            self._handle, U.data_ptr(), self.workspace().data_ptr(),
            _lib.dbl(self.S16(params, dt, t)), int(order), float(ko_sigma), BC[bc], hptr, hn,
            float(t), float(dt), params.chi_floor, params.alpha_floor, tptr,
            stream_handle(self.device)))
        return U

    def periodic_sync_(self, U: torch.Tensor) -> torch.Tensor:
        self._check_pool(U)
        _lib.check(_lib.load().bssn_batch_periodic_sync_f64(self._handle, U.data_ptr(),
                                                            stream_handle(self.device)))
        return U

    def halo_(self, U: torch.Tensor, halo: "HaloMap") -> torch.Tensor:
        self._check_pool(U)
        _lib.check(_lib.load().bssn_batch_halo_f64(U.data_ptr(), U.numel(), halo.data.data_ptr(),
                                                   halo.n, stream_handle(self.device)))
        return U

    # --- the precision form over the batch ----------------------------------------------
    @property
    def bundle_doubles(self) -> int:
        return int(_lib.load().bssn_batch_bundle_doubles(self._handle))

    def bundle_offsets(self):
        """Offset of each block's 138 derivative arrays in a batch bundle (the library's
        layout: block after block, 64-double aligned)."""
        out, pos = [], 0
        for P in self.P:
            out.append(pos)
            pos += (138 * P + 63) // 64 * 64
        return out

    def bundle_view(self, D: torch.Tensor, b: int) -> torch.Tensor:
        o = self.bundle_offsets()[b]
        return D[o:o + 138 * self.P[b]].view(138, *self.S[b])

    def deriv_bundle_i8(self, F: torch.Tensor, D: torch.Tensor, slices: int = 7) -> None:
        """D <- the INT8-split derivative bundle of every block (one launch per pass)."""
        self._check_pool(F)
        if D.dtype != torch.float64 or D.numel() < self.bundle_doubles or D.device != self.device:
            raise ValueError("D must hold bundle_doubles float64 values on the batch's device")
        _lib.check(_lib.load().bssn_batch_deriv_bundle_i8_f64(
            self._handle, F.data_ptr(), D.data_ptr(), 4, int(slices), stream_handle(self.device)))

    def _int8_stage(self, Y, out, D, t, dt, params, halo, ko_sigma, slices):
        """out <- f(Y) of every block: ghosts of Y from the halo map, the batched INT8 bundle,
        then per block the FP64 algebra and the Kreiss-Oliger term (ghosts read as data)."""
        if halo is not None:
            self.halo_(Y, halo)
        self.deriv_bundle_i8(Y, D, slices)
        lib = _lib.load()
        st = stream_handle(self.device)
        offs = self.bundle_offsets()
        for b, blk in enumerate(self.blocks):
            Sx, Sy, Sz = self.S[b]
            P = self.P[b]
            Fb, Ob = self.view(Y, b), self.view(out, b)
            Db = D[offs[b]:offs[b] + 138 * P]
            _lib.check(lib.bssn_algebra_f64(Fb.data_ptr(), Db.data_ptr(),
                                            _lib.dbl(scalars(params, *blk.h, dt, t)),
                                            Ob.data_ptr(), P, st))
            if ko_sigma > 0.0:
                _lib.check(lib.bssn_ko_add_f64(Fb.data_ptr(), Ob.data_ptr(), Sx, Sy, Sz, 4,
                                               float(ko_sigma), _lib.dbl(blk.h), 0, self.ng, st))

    def int8_step_(self, U: torch.Tensor, t: float, dt: float,
                   params: PhysicsParams | None = None, halo: "HaloMap | None" = None,
                   ko_sigma: float = 0.1, slices: int = 7) -> torch.Tensor:
        """One RK4 step of every block with the INT8-split derivatives (FD order 4), composed
        as the single-grid INT8 step is (per stage: ghost refill, bundle, algebra, KO; RK4
        combinations over the whole pool; enforce; a final ghost refill), the ghosts of each
        stage input taken from other blocks by the halo map (U synced on entry)."""
        params = params or PhysicsParams()
        self._check_pool(U)
        if self._int8_buf is None or self._int8_buf[0].numel() < self.pool_doubles:
            n = self.pool_doubles
            self._int8_buf = (torch.empty(n, dtype=torch.float64, device=self.device),
                              torch.empty(n, dtype=torch.float64, device=self.device),
                              torch.empty(n, dtype=torch.float64, device=self.device),
                              torch.empty(self.bundle_doubles, dtype=torch.float64,
                                          device=self.device))
        k, acc, Y, D = self._int8_buf
        Un = U[:self.pool_doubles]
        h = 0.5 * dt
        self._int8_stage(U, k, D, t, dt, params, None, ko_sigma, slices)
        acc.copy_(k)
        torch.add(Un, k, alpha=h, out=Y)
        self._int8_stage(Y, k, D, t + h, dt, params, halo, ko_sigma, slices)
        acc.add_(k, alpha=2.0)
        torch.add(Un, k, alpha=h, out=Y)
        self._int8_stage(Y, k, D, t + h, dt, params, halo, ko_sigma, slices)
        acc.add_(k, alpha=2.0)
        torch.add(Un, k, alpha=dt, out=Y)
        self._int8_stage(Y, k, D, t + dt, dt, params, halo, ko_sigma, slices)
        acc.add_(k)
        Un.add_(acc, alpha=dt / 6.0)
        lib = _lib.load()
        st = stream_handle(self.device)
        for b in range(len(self.blocks)):
            v = self.view(U, b)
            _lib.check(lib.bssn_enforce_f64(v.data_ptr(), self.P[b], params.chi_floor,
                                            params.alpha_floor, st))
        if halo is not None:
            self.halo_(U, halo)
        return U

    # --- tuning and capture ---------------------------------------------------------------
    def autotune(self, U: torch.Tensor, dt: float, params: PhysicsParams | None = None,
                 order: int = 4, ko_sigma: float = 0.1, bc: str = "periodic",
                 halo: "HaloMap | None" = None, candidates: Iterable[int] | None = None,
                 reps: int = 3) -> dict:
        """Time one step per launch shape (0 = one workgroup per brick, N = N persistent waves)
        on a copy of U and keep the fastest (the first candidate unless another beats it by more
        than the timing noise, pick_launch). Returns {shape: ms}."""
        return tune_step_launch(self, lambda V: self.step_(V, 0.0, dt, params, order, ko_sigma,
                                                           bc, halo), U, candidates, reps)

    def capture_steps(self, U: torch.Tensor, tdev: torch.Tensor, dt: float, nsteps: int = 1,
                      params: PhysicsParams | None = None, order: int = 4,
                      ko_sigma: float = 0.1, bc: str = "periodic",
                      halo: "HaloMap | None" = None) -> torch.cuda.CUDAGraph:
        """Capture ``nsteps`` steps of U into a HIP graph; each replay advances U by nsteps
        steps from the time held in tdev (which it advances)."""
        self.workspace()
        warm = U.clone()          # first launches (code loading) happen outside the capture
        self.step_(warm, 0.0, dt, params, order, ko_sigma, bc, halo)
        torch.cuda.synchronize(self.device)
        del warm
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                for _ in range(nsteps):
                    self.step_(U, 0.0, dt, params, order, ko_sigma, bc, halo, tdev=tdev)
        torch.cuda.current_stream(self.device).wait_stream(s)
        return g


class HaloMap:
    """Device array of bssn_halo_copy entries."""

    def __init__(self, entries: np.ndarray, device):
        entries = np.ascontiguousarray(entries, dtype=HALO_DTYPE)
        self.n = int(entries.shape[0])
        raw = torch.from_numpy(entries.view(np.int64).reshape(-1).copy())
        self.data = raw.to(device)


def tiling_entries(layout: BlockLayout, tiles: Sequence[Sequence[int]],
                   N: Sequence[int]) -> np.ndarray:
    """Halo copies of blocks that tile a periodic box of N interior points per axis at one
    spacing: block b covers global interior indices tiles[b] + [0, n_b). Every ghost point of
    every block is copied from the block owning its periodically wrapped global index."""
    ng = layout.ng
    N = np.asarray(N, dtype=np.int64)
    owner = -np.ones(tuple(N), dtype=np.int64)
    for b, (t, blk) in enumerate(zip(tiles, layout.blocks)):
        sl = tuple(slice(int(t[a]), int(t[a]) + int(blk.n[a])) for a in range(3))
        if np.any(owner[sl] >= 0):
            raise ValueError("tiles overlap")
        owner[sl] = b
    if np.any(owner < 0):
        raise ValueError("tiles do not cover the box")
    T = np.asarray(tiles, dtype=np.int64)
    Sarr = np.asarray(layout.S, dtype=np.int64)
    offs = np.asarray(layout.off, dtype=np.int64)
    Ps = np.asarray(layout.P, dtype=np.int64)
    out = []
    for b in range(len(layout.blocks)):
        S = Sarr[b]
        c = np.stack(np.meshgrid(*[np.arange(v) for v in S], indexing="ij"), -1).reshape(-1, 3)
        c = c[np.any((c < ng) | (c >= S - ng), axis=1)]
        g = (c - ng + T[b]) % N
        k = owner[g[:, 0], g[:, 1], g[:, 2]]
        loc = g - T[k] + ng
        e = np.empty(c.shape[0], dtype=HALO_DTYPE)
        e["dst"] = offs[b] + (c[:, 0] * S[1] + c[:, 1]) * S[2] + c[:, 2]
        e["dst_stride"] = Ps[b]
        e["src"] = offs[k] + (loc[:, 0] * Sarr[k, 1] + loc[:, 1]) * Sarr[k, 2] + loc[:, 2]
        e["src_stride"] = Ps[k]
        out.append(e)
    return np.concatenate(out)


def apply_entries(pool: np.ndarray, entries: np.ndarray) -> None:
    """The halo copies on a host pool (what bssn_batch_halo_f64 does on the device)."""
    for f in range(NUM_VARS):
        pool[entries["dst"] + f * entries["dst_stride"]] = \
            pool[entries["src"] + f * entries["src_stride"]]


def tiling_halo(batch: BlockBatch, tiles: Sequence[Sequence[int]], N: Sequence[int]) -> HaloMap:
    return HaloMap(tiling_entries(batch, tiles, N), batch.device)


# --- the batched workload -------------------------------------------------------------------

def config_blocks(seed: int = 0, nblk: int = 640, sizes=(8, 12, 16, 24, 32),
                  spacings=(0.25, 0.125, 0.0625), L: float = 16.0):
    """Block geometry of the batched workload: nblk cubic blocks with interior extent drawn
    uniformly from ``sizes`` and spacing (refinement level) from ``spacings`` by
    numpy default_rng(seed); each block is placed at a random grid-aligned position inside
    [-L/2, L/2)^3 of its level. About 6.8 M interior points for the defaults."""
    rng = np.random.default_rng(seed)
    blocks = []
    for _ in range(nblk):
        n = int(rng.choice(sizes))
        lev = int(rng.integers(0, len(spacings)))
        h = float(spacings[lev])
        cells = int(round(L / h))
        lo = []
        for _a in range(3):
            i0 = int(rng.integers(0, max(cells - n, 0) + 1))
            lo.append(-0.5 * L + i0 * h)
        blocks.append(Block((n, n, n), (h, h, h), tuple(lo), lev))
    return blocks


def puncture_positions(seed: int = 0, masses=(0.909, 0.0909), separation: float = 4.0,
                       h_fine: float = 0.0625):
    """Two punctures on the x axis about their centre of mass, ``separation`` apart, shifted
    off the finest grid planes by a seeded sub-spacing offset."""
    from .initial_data import offgrid_shift
    rng = np.random.default_rng(seed + 7)
    M = float(sum(masses))
    s = offgrid_shift(rng, h_fine)
    p1 = np.array([separation * masses[1] / M, 0.0, 0.0]) + s
    p2 = np.array([-separation * masses[0] / M, 0.0, 0.0]) + s
    return [p1.tolist(), p2.tolist()]


def fill_punctures(batch: BlockBatch, pool: torch.Tensor, masses=(0.909, 0.0909),
                   positions=None, amp: float = 1e-4, seed: int = 0) -> None:
    """Brill-Lindquist data of ``masses`` at ``positions`` on every block (ghosts included,
    from the closed form), plus the seeded perturbations of initial_data.punctures."""
    from .initial_data import punctures
    positions = positions if positions is not None else puncture_positions(seed, masses)
    batch.fill(pool, lambda b, blk, grid: punctures(grid, list(masses), positions, amp=amp,
                                                     seed=seed + 1000 + b, device=pool.device))


# a launch shape other than the first candidate must beat it by more than this fraction (the
# run-to-run noise of a few timed steps), so that near-ties do not flip the shape between runs
LAUNCH_MARGIN = 0.005


def pick_launch(res: dict, margin: float = LAUNCH_MARGIN) -> int:
    """The launch shape to keep from {shape: ms} (in the order timed): the first one unless
    another is faster by more than ``margin``, then the fastest."""
    first = next(iter(res))
    best = min(res, key=res.get)
    return best if res[best] < res[first] * (1.0 - margin) else first


def tune_step_launch(batch: "BlockBatch", step, U: torch.Tensor, candidates=None,
                     reps: int = 3) -> dict:
    """Time step(V) for each launch shape of ``batch`` (0 = one workgroup per brick, N = N
    persistent waves; default 0 and 4, 8, 16 waves per CU) on a copy V of U, set the shape
    pick_launch keeps and return {shape: ms}."""
    if candidates is None:
        cus = torch.cuda.get_device_properties(batch.device).multi_processor_count
        candidates = [0] + [cus * k for k in (4, 8, 16)]
    V = U.clone()
    res = {}
    for c in candidates:
        c = int(c) // 8 * 8
        batch.set_launch(c)
        step(V)
        torch.cuda.synchronize(batch.device)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            step(V)
        e1.record()
        torch.cuda.synchronize(batch.device)
        res[c] = e0.elapsed_time(e1) / reps
    batch.set_launch(pick_launch(res))
    del V
    return res


def timed(fn, reps: int, device) -> float:
    """Mean ms of fn() over reps launches, by events on the current stream."""
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize(device)
    return e0.elapsed_time(e1) / reps
