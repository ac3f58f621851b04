"""BSSN time evolution on the device: fused RK4 / MSRK steps (3D/src/bssn3d/evolve.py:24-194).

BSSNEvolution(grid, params, order, ko_sigma, bc, precision, scheme, ko_order, integrator) keeps
the reference's constructor, rhs/enforce/step/evolve methods and semantics:
  * rhs(state, t, dt)  = periodic sync (bc="periodic") + RHS + Kreiss-Oliger
  * step(state, t, dt) = classic RK4 with stage times t, t+dt/2, t+dt/2, t+dt, the CAHD
                          factor on the full dt at every stage, then enforce
  * integrator "rk4_2_2": one RK4 start step, then the 2-step / 3-stage MSRK recurrence
With scheme "fused" each RK stage is ONE fused kernel (RHS + KO + stage update; enforcement
folded into the last stage). With bc="periodic" stencils read periodically wrapped data, the
interior is advanced and the ghosts are refilled by a sync at the end of the step, so the
returned state equals the reference's after the sync its next step would perform. Other bc
values advance every padded point with edge-padded stencils (the reference leaves such ghosts
unsynced too). bc="sommerfeld" gives the ghost zones the radiative outgoing-wave condition
(bssn_rk4_step_sommerfeld_f64; the reference has none for BSSN and this follows its MCS
solver's, 3D/src/mcs3d/main.py:230-250) while the interiors advance as above.
With scheme "twopass", "int8" (the precision form: derivatives from int8 digits) or a "bfp*"
scheme (the bundle stored in block floating point between the passes) the step is
the reference's own composition (evolve.py:86-90, 125-132): per stage a periodic sync, the
materialised 138-derivative bundle, the algebra and KO over the full padded arrays, then the
RK4 combinations over the full arrays, enforce and (periodic) a final sync.
ghosts="carried" (bc="periodic") keeps the reference's ghost values exactly: each stage input is
synced, every padded point is advanced with edge-padded stencils from it, the base state keeps
its own ghosts and nothing is synced at the end. The interior is the same as with the default
ghosts="synced" (bit for bit, fused scheme); the ghosts, which the reference's constraint norms
and long-run samples read, are the reference's.
"""
from __future__ import annotations

import torch

from . import _lib
from .grid import Grid
from .rhs import (BFP_BITS, BSSNSolver, as_tensor, bfp_buffers, rhs_bfp_into, rhs_int8_into,
                  rhs_twopass_into, stream_handle)
from .state import BSSNState, PhysicsParams

INTEGRATORS = ("rk4", "rk4_2_2")
GHOSTS = ("synced", "carried")


class BSSNEvolution:
    def __init__(self, grid: Grid, params: PhysicsParams | None = None, order: int = 6,
                 ko_sigma: float = 0.0, bc: str = "periodic", precision: str = "fp64",
                 scheme: str = "fused", ko_order: int | None = None, integrator: str = "rk4",
                 ghosts: str = "synced"):
        self.grid = grid
        self.params = params or PhysicsParams()
        self.solver = BSSNSolver(grid, self.params, order, precision=precision, scheme=scheme,
                                 ko_order=ko_order)
        self.order = order
        self.ng = grid.ng
        self.ko_sigma = float(ko_sigma)
        self.bc = bc
        if integrator not in INTEGRATORS:
            raise ValueError(f"integrator {integrator!r}: only 'rk4' or 'rk4_2_2' are supported")
        self.integrator = integrator
        if self.solver.scheme != "fused" and integrator != "rk4":
            raise ValueError(f"the {self.solver.scheme} scheme steps with rk4 only")
        if bc == "sommerfeld" and (integrator != "rk4" or self.solver.scheme != "fused"):
            raise ValueError('bc="sommerfeld" steps with the fused scheme and rk4')
        if ghosts not in GHOSTS:
            raise ValueError(f"ghosts {ghosts!r}: use one of {GHOSTS}")
        self.ghosts = ghosts
        self._work = None
        self._D = None
        self._Q = None

    @property
    def periodic(self) -> bool:
        return self.bc == "periodic"

    @property
    def carried(self) -> bool:
        return self.periodic and self.ghosts == "carried"

    @property
    def _ghost_mode(self) -> int:
        """The step entry points' ghost argument: 0 edge, 1 wrapped + final sync, 2 carried."""
        return 2 if self.carried else int(self.periodic)

    @property
    def radiative(self) -> bool:
        return self.bc == "sommerfeld"

    def _xmin3(self):
        g = self.grid
        return _lib.dbl([g.xmin, g.ymin, g.zmin])

    def _S16(self, t: float, dt: float):
        return self.solver.S16(t, dt)

    def _workspace(self, F: torch.Tensor, integrator: int) -> torch.Tensor:
        _, Sx, Sy, Sz = F.shape
        nbytes = _lib.load().bssn_step_workspace_bytes(integrator, Sx, Sy, Sz)
        n = nbytes // 8
        if self._work is None or self._work.numel() < n or self._work.device != F.device:
            self._work = torch.empty(n, dtype=torch.float64, device=F.device)
        return self._work

    # --- pieces -----------------------------------------------------------------------------
    def rhs(self, state, t: float = 0.0, dt: float | None = None) -> BSSNState:
        """d/dt state as the step sees it: sync (periodic) + RHS + KO (evolve.py:86-90).
        Periodic: interior values only (ghost entries of the result are zero); with carried
        ghosts, the rates at every padded point of the synced state (edge-padded stencils)."""
        F = as_tensor(state)
        out = torch.zeros_like(F)
        _, Sx, Sy, Sz = F.shape
        dt = self.solver.dt if dt is None else dt
        lib = _lib.load()
        S16 = self._S16(t, dt)
        wrap = self.periodic
        if self.carried:
            F = self.periodic_sync_(F.clone())
            wrap = False
        _lib.check(lib.bssn_stage_rhs_f64(
            F.data_ptr(), out.data_ptr(), Sx, Sy, Sz, _lib.dbl(S16), self.order,
            self.ko_sigma, self.ng, int(wrap), stream_handle(F.device)))
        if self.radiative:
            _lib.check(lib.bssn_sommerfeld_rhs_f64(
                F.data_ptr(), out.data_ptr(), Sx, Sy, Sz, self.order, _lib.dbl(S16[13:16]),
                self.ng, self._xmin3(), stream_handle(F.device)))
        return BSSNState(out)

    def enforce(self, state) -> BSSNState:
        F = as_tensor(state).clone()
        _lib.check(_lib.load().bssn_enforce_f64(
            F.data_ptr(), F[0].numel(), self.params.chi_floor, self.params.alpha_floor,
            stream_handle(F.device)))
        return BSSNState(F)

    def periodic_sync_(self, F: torch.Tensor) -> torch.Tensor:
        _, Sx, Sy, Sz = F.shape
        _lib.check(_lib.load().bssn_periodic_sync_f64(F.data_ptr(), Sx, Sy, Sz, self.ng,
                                                      stream_handle(F.device)))
        return F

    # --- steps (in place on a tensor) ---------------------------------------------------------
    def step_(self, U: torch.Tensor, t: float, dt: float,
              tdev: torch.Tensor | None = None) -> torch.Tensor:
        """One RK4 step of U in place. With ``tdev`` (one float64 on U's device) the step
        takes its time from there (``t`` is ignored) and advances it by dt."""
        if self.solver.scheme != "fused":
            if tdev is not None:
                raise ValueError("device-resident time needs the fused scheme")
            return self._composed_step_(U, t, dt)
        _, Sx, Sy, Sz = U.shape
        work = self._workspace(U, 0)
        lib = _lib.load()
        if self.radiative:
            if tdev is not None:
                raise ValueError("device-resident time is for the periodic and edge steps")
            _lib.check(lib.bssn_rk4_step_sommerfeld_f64(
                U.data_ptr(), work.data_ptr(), Sx, Sy, Sz, _lib.dbl(self._S16(t, dt)),
                self.order, self.ko_sigma, self.ng, self._xmin3(), float(t), float(dt),
                self.params.chi_floor, self.params.alpha_floor, stream_handle(U.device)))
            return U
        if tdev is None:
            _lib.check(lib.bssn_rk4_step_f64(
                U.data_ptr(), work.data_ptr(), Sx, Sy, Sz, _lib.dbl(self._S16(t, dt)),
                self.order, self.ko_sigma, self.ng, self._ghost_mode, float(t), float(dt),
                self.params.chi_floor, self.params.alpha_floor, stream_handle(U.device)))
            return U
        if tdev.dtype != torch.float64 or tdev.numel() != 1 or tdev.device != U.device:
            raise ValueError("tdev must be one float64 value on the state's device")
        _lib.check(lib.bssn_rk4_step_devtime_f64(
            U.data_ptr(), work.data_ptr(), Sx, Sy, Sz, _lib.dbl(self._S16(0.0, dt)),
            self.order, self.ko_sigma, self.ng, self._ghost_mode, tdev.data_ptr(), float(dt),
            self.params.chi_floor, self.params.alpha_floor, stream_handle(U.device)))
        return U

    # --- the composed step of the two-pass schemes ----------------------------------------
    def _bundle(self, F: torch.Tensor) -> torch.Tensor:
        if self._D is None or self._D.shape[1:] != F.shape[1:] or self._D.device != F.device:
            self._D = torch.empty((138,) + tuple(F.shape[1:]), dtype=torch.float64,
                                  device=F.device)
        return self._D

    def _stage_rhs_twopass(self, Y: torch.Tensor, out: torch.Tensor, t: float, dt: float):
        """f(Y, t): periodic sync of Y's ghosts, then bundle + algebra + KO at every point."""
        if self.periodic:
            self.periodic_sync_(Y)
        S = self._S16(t, dt)
        if self.solver.scheme in BFP_BITS:
            k = BFP_BITS[self.solver.scheme]
            if self._Q is None or self._Q[2] != tuple(Y.shape) or self._Q[0].device != Y.device:
                self._Q = bfp_buffers(Y, k) + (tuple(Y.shape),)
            rhs_bfp_into(Y, out, self._Q[0], self._Q[1], S, self.order, self.ko_sigma, k)
            return out
        D = self._bundle(Y)
        if self.solver.scheme == "int8":
            rhs_int8_into(Y, out, D, S, self.order, self.ko_sigma)
        else:
            rhs_twopass_into(Y, out, D, S, self.order, self.ko_sigma)
        return out

    def _composed_step_(self, U: torch.Tensor, t: float, dt: float) -> torch.Tensor:
        k = torch.empty_like(U)
        acc = torch.empty_like(U)
        Y = torch.empty_like(U)
        h = 0.5 * dt
        if self.carried:                                         # k1 from a synced copy:
            Y.copy_(U)                                           # U keeps its own ghosts
            self._stage_rhs_twopass(Y, k, t, dt)
        else:
            self._stage_rhs_twopass(U, k, t, dt)                 # k1 (syncs U's ghosts)
        acc.copy_(k)
        torch.add(U, k, alpha=h, out=Y)
        self._stage_rhs_twopass(Y, k, t + h, dt)                 # k2
        acc.add_(k, alpha=2.0)
        torch.add(U, k, alpha=h, out=Y)
        self._stage_rhs_twopass(Y, k, t + h, dt)                 # k3
        acc.add_(k, alpha=2.0)
        torch.add(U, k, alpha=dt, out=Y)
        self._stage_rhs_twopass(Y, k, t + dt, dt)                # k4
        acc.add_(k)
        U.add_(acc, alpha=dt / 6.0)
        _, Sx, Sy, Sz = U.shape
        _lib.check(_lib.load().bssn_enforce_f64(U.data_ptr(), Sx * Sy * Sz,
                                                self.params.chi_floor, self.params.alpha_floor,
                                                stream_handle(U.device)))
        if self.periodic and not self.carried:
            self.periodic_sync_(U)
        return U

    def capture_steps(self, U: torch.Tensor, tdev: torch.Tensor, dt: float,
                      nsteps: int) -> torch.cuda.CUDAGraph:
        """Record ``nsteps`` RK4 steps of U (time in tdev) into a HIP graph; each replay
        advances U and tdev by nsteps steps."""
        self._workspace(U, 0)
        warm = U.clone()                      # first launches happen outside the capture
        self.step_(warm, 0.0, dt)
        torch.cuda.synchronize(U.device)
        del warm
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=U.device)
        s.wait_stream(torch.cuda.current_stream(U.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                for _ in range(nsteps):
                    self.step_(U, 0.0, dt, tdev=tdev)
        torch.cuda.current_stream(U.device).wait_stream(s)
        return g

    def msrk_step_(self, U: torch.Tensor, fprev: torch.Tensor, fnext: torch.Tensor, t: float,
                   dt: float) -> torch.Tensor:
        """One rk4_2_2 step of U in place; fprev (f(y_{n-1})) is clobbered, fnext <- f(y_n)."""
        _, Sx, Sy, Sz = U.shape
        work = self._workspace(U, 1)
        _lib.check(_lib.load().bssn_msrk_step_f64(
            U.data_ptr(), fprev.data_ptr(), fnext.data_ptr(), work.data_ptr(), Sx, Sy, Sz,
            _lib.dbl(self._S16(t, dt)), self.order, self.ko_sigma, self.ng, self._ghost_mode,
            float(t), float(dt), self.params.chi_floor, self.params.alpha_floor,
            stream_handle(U.device)))
        return U

    def step(self, state, t: float, dt: float) -> BSSNState:
        U = as_tensor(state).clone()
        return BSSNState(self.step_(U, t, dt))

    def evolve(self, state, dt: float, nsteps: int, t0: float = 0.0,
               graph_chunk: int = 0) -> BSSNState:
        """nsteps steps from t0 (evolve.py:155-194); the input is not modified.
        graph_chunk > 0 (RK4): capture that many steps once into a HIP graph and replay it,
        with the time advanced on the device -- the analogue of the reference's single
        compiled fori_loop; the SSL factor is then formed by the device's exp."""
        U = as_tensor(state).clone()
        if nsteps <= 0:
            return BSSNState(U)
        if self.integrator == "rk4":
            if graph_chunk > 0 and nsteps >= graph_chunk:
                tdev = torch.full((1,), float(t0), dtype=torch.float64, device=U.device)
                g = self.capture_steps(U, tdev, dt, graph_chunk)
                for _ in range(nsteps // graph_chunk):
                    g.replay()
                for _ in range(nsteps % graph_chunk):
                    self.step_(U, 0.0, dt, tdev=tdev)
                return BSSNState(U)
            t = t0
            for _ in range(nsteps):
                self.step_(U, t, dt)
                t = t + dt
            return BSSNState(U)
        # MSRK: f(y_0) fills the history, one RK4 start step, then the recurrence
        fprev = self.rhs(U, t0, dt).data
        fnext = torch.empty_like(U)
        self.step_(U, t0, dt)
        t = t0 + dt
        for _ in range(1, nsteps):
            self.msrk_step_(U, fprev, fnext, t, dt)
            fprev, fnext = fnext, fprev
            t = t + dt
        return BSSNState(U)
