"""Autotuner of the fused stage kernels' schedule on the device it runs on.

    python integer-schemes_amd/autotune.py [--n 256] [--steps 6] [--space full|quick]
                                           [--write] [--out tuning.json]

The stage kernels' schedule is a set of compile-time knobs (bssn_params.h, bssn_brick.h,
bssn_stage.h): the brick of output points one wave owns (BSSN_BRICK_X/Y/Z, 64 points: the
4x4x4 cube stages the fewest halo cells, the flatter bricks longer z runs), the LDS box's
x-plane stride (BSSN_BOX_XPLANE_MOD32: dense, or padded to a residue modulo 32 doubles that
spreads the brick's lanes over the banks), the order in which an XCD's workgroups take their
bricks (BSSN_BRICK_TILE_XY: z columns in square tiles of columns, or 0 for rows of columns), how many fields are staged ahead of the one being
evaluated (BSSN_STAGE_DEPTH for the RK/MSRK stage kernels, BSSN_RHS_DEPTH for the bare RHS),
how many algebra phases ahead of its store a field's write-back operands are loaded
(BSSN_WB_LEAD), and how many waves per SIMD the register allocator targets (BSSN_STAGE_WAVES,
BSSN_RHS_WAVES; two waves force one field in flight). A knob moves loads, LDS addresses and
register allocation, not arithmetic, so every variant computes the same RHS.

For every candidate the tuner
  1. builds a variant library (build.build(out=..., defines=...)) under integer-schemes_amd/_tune/;
  2. in a fresh process with that library (BSSN_HIP_LIB), checks the variant on a 48^3 puncture
     grid against the installed library's RK4 step (normwise 1e-12: the knobs may only move
     FMA formation) and times the RK4 step of the 256^3 equal-mass binary (the benchmark
     workload) and the bare RHS;
and keeps the defaults unless a passing candidate is faster on the step by more than the timing
noise (0.5 %) without slowing the bare RHS by more (choose()). With --write the result goes to
tuning.json next to build.py, whose defines build.build() then compiles by default; the batched
launch shape is tuned at run time on each batch (BlockBatch.autotune,
PatchHierarchy.tune_launch).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
from pathlib import Path

PKG = Path(__file__).resolve().parent
TUNING = PKG / "tuning.json"

SPACES = {
    "quick": [
        {},
        {"BSSN_STAGE_DEPTH": 2},
        {"BSSN_BOX_XPLANE_MOD32": 16},
    ],
    "full": [
        {},
        {"BSSN_STAGE_DEPTH": 2},
        {"BSSN_WB_LEAD": 1},
        {"BSSN_WB_LEAD": 3},
        {"BSSN_STAGE_DEPTH": 2, "BSSN_WB_LEAD": 1},
        {"BSSN_STAGE_DEPTH": 2, "BSSN_WB_LEAD": 3},
        {"BSSN_RHS_DEPTH": 1},
        {"BSSN_STAGE_WAVES": 2},
        {"BSSN_RHS_WAVES": 2},
        {"BSSN_BOX_XPLANE_MOD32": 16},
        {"BSSN_BOX_XPLANE_MOD32": 16, "BSSN_WB_LEAD": 3},
        {"BSSN_BRICK_X": 2, "BSSN_BRICK_Y": 4, "BSSN_BRICK_Z": 8},
        {"BSSN_BRICK_X": 4, "BSSN_BRICK_Y": 2, "BSSN_BRICK_Z": 8},
        {"BSSN_BRICK_X": 2, "BSSN_BRICK_Y": 8, "BSSN_BRICK_Z": 4},
        {"BSSN_BRICK_X": 8, "BSSN_BRICK_Y": 4, "BSSN_BRICK_Z": 2},
        {"BSSN_BRICK_X": 2, "BSSN_BRICK_Y": 2, "BSSN_BRICK_Z": 16},
        {"BSSN_BRICK_TILE_XY": 0},
        {"BSSN_BRICK_TILE_XY": 2},
        {"BSSN_BRICK_TILE_XY": 8},
    ],
}
# the brick shapes alone, and the brick orders alone (with the defaults as the baseline every
# space starts with)
SPACES["bricks"] = [{}] + [d for d in SPACES["full"] if "BSSN_BRICK_X" in d]
SPACES["order"] = [{}] + [d for d in SPACES["full"] if "BSSN_BRICK_TILE_XY" in d]

# Runs in the candidate's process: time the step and the RHS, compare the 48^3 step with the
# installed library's result (written by the parent).
_CHILD = r"""
import json, sys, time
sys.path.insert(0, sys.argv[1])
import numpy as np, torch
from bssn3d_hip import initial_data as idata
from bssn3d_hip.evolve import BSSNEvolution
from bssn3d_hip.rhs import rhs_into
from bssn3d_hip.state import PhysicsParams, scalars
n, steps, ref_path, mode = int(sys.argv[2]), int(sys.argv[3]), sys.argv[4], sys.argv[5]
dev = torch.device("cuda:0")
p = PhysicsParams()
g, U, _ = idata.config_single_puncture(seed=3, N=48, device=dev)
ev = BSSNEvolution(g, p, order=4, ko_sigma=0.1, bc="periodic")
dt = 0.25 * g.dx
for k in range(2):
    ev.step_(U, k * dt, dt)
small = U.cpu().numpy()
if mode == "ref":
    np.save(ref_path, small)
    print(json.dumps({"ok": True}))
    sys.exit(0)
ref = np.load(ref_path)
den = np.maximum(np.abs(ref).reshape(24, -1).max(axis=1), 1e-12)
err = float((np.abs(small - ref).reshape(24, -1).max(axis=1) / den).max())
del U
grid, U, _ = idata.config_binary(seed=0, N=n, device=dev)
ev = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="periodic")
dt = 0.25 * grid.dx
for k in range(2):
    ev.step_(U, k * dt, dt)
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
torch.cuda.synchronize()
e0.record()
for k in range(steps):
    ev.step_(U, (2 + k) * dt, dt)
e1.record()
torch.cuda.synchronize()
step_ms = e0.elapsed_time(e1) / steps
out = torch.empty_like(U)
S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, 0.0)
rhs_into(U, out, S16, 4)
e0.record()
for k in range(steps):
    rhs_into(U, out, S16, 4)
e1.record()
torch.cuda.synchronize()
print(json.dumps({"ok": bool(torch.isfinite(U).all()), "err": err, "step_ms": step_ms,
                  "rhs_ms": e0.elapsed_time(e1) / steps}))
"""


# a candidate must beat the defaults' step time by more than this fraction (timing noise)
MARGIN = 0.005


def key_of(defines: dict) -> str:
    return "_".join(f"{k.replace('BSSN_', '').lower()}{v}" for k, v in sorted(defines.items())) \
        or "default"


def run_child(lib: Path | None, n: int, steps: int, ref: Path, mode: str) -> dict:
    env = dict(os.environ)
    if lib is not None:
        env["BSSN_HIP_LIB"] = str(lib)
    r = subprocess.run([sys.executable, "-c", _CHILD, str(PKG), str(n), str(steps), str(ref),
                        mode], capture_output=True, text=True, env=env, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"candidate run failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def tune(space: str = "full", n: int = 256, steps: int = 6, tol: float = 1e-12,
         verbose: bool = True) -> dict:
    import importlib.util
    spec = importlib.util.spec_from_file_location("bssn_build", PKG / "build.py")
    build = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(build)
    cands = SPACES[space]
    work = PKG / "_tune"
    libs = {}
    for d in cands:
        k = key_of(d)
        out = work / k / "libbssn_hip.so"
        if verbose:
            print(f"[autotune] building {k}", flush=True)
        libs[k] = build.build(out=out, defines=[f"{a}={b}" for a, b in d.items()])
    results = {}
    with tempfile.TemporaryDirectory() as td:
        ref = Path(td) / "ref.npy"
        run_child(None, n, steps, ref, "ref")          # the installed library's 48^3 step
        for d in cands:
            k = key_of(d)
            r = run_child(libs[k], n, steps, ref, "time")
            r["passed"] = bool(r["ok"] and r["err"] <= tol)
            r["defines"] = d
            results[k] = r
            if verbose:
                print(f"[autotune] {k:28s} step {r['step_ms']:8.3f} ms  rhs {r['rhs_ms']:7.3f} ms"
                      f"  err {r['err']:.1e}  {'ok' if r['passed'] else 'REJECTED'}", flush=True)
    import torch
    return choose(results, workload=f"bssn_rk4_step_{n}^3_fp64_bl_binary",
                  device=torch.cuda.get_device_name(0), steps_timed=steps)


def choose(results: dict, margin: float = MARGIN, **info) -> dict:
    """The tuning record for measured candidates: the defaults unless a passing candidate is
    faster on the step by more than ``margin`` (run-to-run noise of the timing) without making
    the bare RHS slower by more than ``margin``; among those, the fastest step."""
    base = results["default"]
    better = {k: r for k, r in results.items()
              if r["passed"] and r["step_ms"] <= base["step_ms"] * (1.0 - margin)
              and r["rhs_ms"] <= base["rhs_ms"] * (1.0 + margin)}
    best = min(better, key=lambda k: better[k]["step_ms"]) if better else "default"
    return {"defines": [f"{a}={b}" for a, b in results[best]["defines"].items()],
            "best": best, "margin": margin, **info, "candidates": results}


def merge(old: dict, new: dict) -> dict:
    """One record from two runs on the same device and workload: every candidate of both (the
    newer measurement where both have one, the newer defaults as the baseline), chosen again."""
    if old.get("workload") != new.get("workload"):
        raise ValueError("records of different workloads")
    cands = dict(old.get("candidates", {}))
    cands.update(new["candidates"])
    info = {k: new[k] for k in ("workload", "device", "steps_timed") if k in new}
    return choose(cands, margin=new.get("margin", MARGIN), **info)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--space", default="full", choices=sorted(SPACES))
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--write", action="store_true", help="install the result as tuning.json")
    ap.add_argument("--out", default=None, help="also write the result here")
    ap.add_argument("--merge", default=None,
                    help="a record of earlier candidates (same device and workload) to choose "
                         "among together with this run's")
    a = ap.parse_args(argv)
    res = tune(a.space, a.n, a.steps)
    if a.merge:
        res = merge(json.loads(Path(a.merge).read_text()), res)
    text = json.dumps(res, indent=1)
    print(text)
    if a.out:
        Path(a.out).write_text(text + "\n")
    if a.write:
        TUNING.write_text(text + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
