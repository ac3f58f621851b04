"""Extreme-mass-ratio stress run: the BSSN step over many steps at large size, FP64 against the
INT8-split form, with the constraint norms over time and the throughput of both.

Workload (the fifth benchmark configuration): 384^3 interior, h = 1/24 on [-8, 8)^3, two
Brill-Lindquist punctures of masses 0.990 and 0.0099 separated by 3 along x (shifted together
off the grid planes by a seeded sub-h offset), seeded 1e-4 sinusoid perturbations on every
field, FD order 4 / KO order 6 with sigma 0.1, periodic ghosts, RK4 at dt = h / 4, 200 steps.

Schemes run one after the other on the same initial data:
  "fused"  all-FP64 fused RK4 stages (the reference arithmetic, parity <= 1e-12 per step);
  "int8"   the RK4 step composed from the INT8-split derivative bundle (7 int8 digits on the
           INT8 matrix cores) and the FP64 algebra (the reference's per-stage composition);
  "batched" the same grid tiled into equal blocks, advanced by the batched fused kernel with
           halo copies after every stage (bit-identical to "fused"; the block schedule's
           throughput at this size);
  "int8_batched"
           the two combined: the same block tiling, every stage's derivatives from the batched
           INT8-split bundle of all blocks (bssn_batch_deriv_bundle_i8_f64), the FP64 algebra
           and KO per block, halo copies after every stage (BlockBatch.int8_step_).

The Hamiltonian and momentum constraint norms (interior RMS, FD order 4) are sampled every
``sample_every`` steps; the step time excludes the sampling. The INT8 run is compared with the
FP64 run sample by sample in the tracking gate's floored relative measure |a - b| / (|b| + floor)
(the reference's apples gate, 3D/src/bssn3d/apples_gate.py:57-59).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from typing import Dict, List

import numpy as np
import torch

from . import initial_data as idata
from .batch import Block, BlockBatch, _device, tiling_halo
from .constraints import ConstraintSolver
from .evolve import BSSNEvolution
from .state import ALPHA, PhysicsParams

MASSES = (0.990, 0.0099)
SEPARATION = 3.0
BOX = 16.0


def initial_state(N: int = 384, seed: int = 0, device="cuda"):
    return idata.config_binary(seed=seed, N=N, L=BOX, masses=MASSES, separation=SEPARATION,
                               device=device)


def _sample(U, cs: ConstraintSolver, step: int, t: float) -> Dict:
    H, M = cs.l2(U)
    return dict(step=step, t=t, H=H, M=M, max_alpha=float(U[ALPHA].abs().max()),
                finite=bool(torch.isfinite(U).all()))


def run_scheme(scheme: str, N: int = 384, steps: int = 200, sample_every: int = 20,
               seed: int = 0, device="cuda", tiles: int = 6, verbose: bool = False) -> Dict:
    """One scheme's run: samples [{step, t, H, M, max_alpha, finite}], step time, Mpts/s."""
    dev = _device(device)
    grid, U, _ = initial_state(N, seed, dev)
    p = PhysicsParams()
    dt = 0.25 * grid.dx
    cs = ConstraintSolver(grid, order=4)
    npts = N ** 3
    if scheme in ("batched", "int8_batched"):
        if N % tiles:
            raise ValueError("N must be a multiple of the tile count")
        n = N // tiles
        ng = grid.ng
        blocks = [Block(n=(n, n, n), h=(grid.dx,) * 3,
                        lo=(grid.xmin + i * n * grid.dx, grid.ymin + j * n * grid.dx,
                            grid.zmin + k * n * grid.dx))
                  for i in range(tiles) for j in range(tiles) for k in range(tiles)]
        tl = [(i * n, j * n, k * n) for i in range(tiles) for j in range(tiles)
              for k in range(tiles)]
        batch = BlockBatch(blocks, ng=ng, device=dev)
        halo = tiling_halo(batch, tl, (N, N, N))
        pool = batch.empty_pool()
        for b, (o, blk) in enumerate(zip(tl, blocks)):
            batch.view(pool, b).copy_(U[:, o[0]:o[0] + n + 2 * ng, o[1]:o[1] + n + 2 * ng,
                                        o[2]:o[2] + n + 2 * ng])
        batch.halo_(pool, halo)

        def to_grid():
            for b, o in enumerate(tl):
                U[:, ng + o[0]:ng + o[0] + n, ng + o[1]:ng + o[1] + n, ng + o[2]:ng + o[2] + n] = \
                    batch.view(pool, b)[:, ng:ng + n, ng:ng + n, ng:ng + n]
            BSSNEvolution(grid, p, order=4, bc="periodic").periodic_sync_(U)
            return U

        if scheme == "batched":
            def advance(k, t):
                for i in range(k):
                    batch.step_(pool, t + i * dt, dt, p, 4, 0.1, "halo", halo)
        else:
            def advance(k, t):
                for i in range(k):
                    batch.int8_step_(pool, t + i * dt, dt, p, halo, ko_sigma=0.1)
        state = to_grid
    else:
        ev = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="periodic", scheme=scheme)

        def advance(k, t):
            for i in range(k):
                ev.step_(U, t + i * dt, dt)
        state = lambda: U
    samples: List[Dict] = [_sample(U, cs, 0, 0.0)]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = 0.0
    done = 0
    U0 = U.clone()
    advance(1, 0.0)                      # warm-up step (code loading), then back to t = 0
    torch.cuda.synchronize(dev)
    U.copy_(U0)
    del U0
    if scheme in ("batched", "int8_batched"):
        for b, o in enumerate(tl):
            batch.view(pool, b).copy_(U[:, o[0]:o[0] + n + 2 * ng, o[1]:o[1] + n + 2 * ng,
                                        o[2]:o[2] + n + 2 * ng])
        batch.halo_(pool, halo)
    w0 = time.perf_counter()
    while done < steps:
        k = min(sample_every, steps - done)
        e0.record()
        advance(k, done * dt)
        e1.record()
        torch.cuda.synchronize(dev)
        ms += e0.elapsed_time(e1)
        done += k
        samples.append(_sample(state(), cs, done, done * dt))
        if verbose:
            s = samples[-1]
            print(f"  {scheme:8s} step {done:4d}  H {s['H']:.6e}  M {s['M']:.6e}  "
                  f"max|alpha| {s['max_alpha']:.5f}  finite {s['finite']}", flush=True)
        if not samples[-1]["finite"]:
            break
    wall = time.perf_counter() - w0
    return dict(scheme=scheme, N=N, steps=done, dt=dt, samples=samples,
                ms_per_step=ms / max(done, 1), mpts_s=npts / (ms / max(done, 1) * 1e-3) / 1e6,
                wall_s=wall)


def compare(ref: Dict, cand: Dict, floor: float = 1e-12) -> Dict:
    n = min(len(ref["samples"]), len(cand["samples"]))
    dev = {}
    for m in ("H", "M", "max_alpha"):
        a = np.array([s[m] for s in cand["samples"][:n]], dtype=np.float64)
        b = np.array([s[m] for s in ref["samples"][:n]], dtype=np.float64)
        dev[m] = float(np.max(np.abs(a - b) / (np.abs(b) + floor))) if n else 0.0
    return dict(dev=dev, same_length=len(ref["samples"]) == len(cand["samples"]),
                finite=all(s["finite"] for s in cand["samples"]))


def stress(N: int = 384, steps: int = 200, sample_every: int = 20, seed: int = 0,
           schemes=("fused", "int8", "batched", "int8_batched"), device="cuda",
           verbose: bool = False) -> Dict:
    out = {}
    for s in schemes:
        out[s] = run_scheme(s, N, steps, sample_every, seed, device, verbose=verbose)
        torch.cuda.empty_cache()
    if "fused" in out:
        for s in schemes:
            if s != "fused":
                out[s]["vs_fused"] = compare(out["fused"], out[s])
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="extreme-mass-ratio BSSN stress run")
    ap.add_argument("--N", type=int, default=384)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--sample-every", type=int, default=20)
    ap.add_argument("--schemes", default="fused,int8,batched,int8_batched")
    ap.add_argument("--out", default=None, help="write the results as JSON here")
    a = ap.parse_args(argv)
    res = stress(a.N, a.steps, a.sample_every, schemes=tuple(a.schemes.split(",")), verbose=True)
    for s, r in res.items():
        print(f"# {s}: {r['ms_per_step']:.2f} ms/step, {r['mpts_s']:.1f} Mpts/s, "
              f"{r['steps']} steps" + (f", vs fused {r['vs_fused']}" if "vs_fused" in r else ""),
              flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f)
    return 0


if __name__ == "__main__":
    sys.exit(main())
