"""Long-run stability series on the device (3D/src/bssn3d/longrun_stability.py:76-188).

run(id_name, N, crossings, samples, ko_sigma, scheme, amp, ...) evolves one of the reference's
test data sets on the periodic cube [-1/2, 1/2]^3 (one crossing time = 1 = 1/dt steps at
dt = CFL h, CFL 0.25) in `samples` chunks and records after each chunk

    ncross, steps, H and M (interior RMS of the Hamiltonian / momentum constraints), max|alpha|,
    max|U - Minkowski| and finiteness,

stopping early when a sample is non-finite or max|alpha| exceeds max_alpha_stop. The step keeps
the reference's carried ghost values (the constraint stencils and the maxima read them), so the
series are the reference's. fit_rates(rows) gives the secular linear and exponential rates of
H and M per crossing time (longrun_stability.py:76-107).

Schemes are BSSNSolver's: "fused" (cuda_fused), "twopass" (verbatim) and "int8" (FD order 4).
The reference fixes FD order 6; `order` is a parameter here so the INT8-split scheme (order 4
only) can be run too.
"""
from __future__ import annotations

import argparse
import csv
import math
import sys
from typing import Callable, Dict, List

import numpy as np
import torch

from . import initial_data as idata
from .constraints import ConstraintSolver
from .evolve import BSSNEvolution
from .grid import Grid
from .state import ALPHA, PhysicsParams

CFL = 0.25
IDS = ("gauge", "robust", "minkowski", "hambump")


def id_fn(name: str, amp: float) -> Callable[[Grid, str], torch.Tensor]:
    """Initial data by name (longrun_stability.py:110-121)."""
    if name == "gauge":
        return lambda g, dev: idata.gauge_wave(g, amplitude=amp, wavelength=1.0, device=dev)
    if name == "robust":
        return lambda g, dev: idata.robust_stability(g, amp=amp, seed=1, device=dev)
    if name == "minkowski":
        return lambda g, dev: idata.minkowski(g, device=dev)
    if name == "hambump":
        return lambda g, dev: idata.hamiltonian_bump(g, amplitude=amp, device=dev)
    raise ValueError(f"unknown id {name!r} (gauge|robust|minkowski|hambump)")


def chunking(N: int, crossings: float, samples: int, cfl: float = CFL):
    """(dt, total steps, steps per sample) on the unit cube with N interior points per axis."""
    dt = cfl * (1.0 / N)
    total = round(crossings / dt)
    return dt, total, max(1, round(total / samples))


def sample_row(U: torch.Tensor, t: float, dt: float, cs: ConstraintSolver,
               mink: torch.Tensor) -> Dict:
    H, M = cs.l2(U)
    return dict(ncross=float(t), steps=round(t / dt), H=float(H), M=float(M),
                max_alpha=float(U[ALPHA].abs().max()),
                max_dev=float((U - mink).abs().max()),
                finite=bool(torch.isfinite(U).all()))


def run(id_name: str = "gauge", N: int = 64, crossings: float = 20.0, samples: int = 40,
        ko_sigma: float = 0.4, scheme: str = "fused", amp: float = 0.01, out: str | None = None,
        max_alpha_stop: float = 10.0, cahd_c: float | None = None, order: int = 6,
        integrator: str = "rk4", device="cuda", verbose: bool = False) -> List[Dict]:
    """The sample rows of one long run (longrun_stability.py:124-188)."""
    g = Grid.from_domain(N, order=order)
    params = PhysicsParams() if cahd_c is None else PhysicsParams(cahd_c=cahd_c)
    ev = BSSNEvolution(g, params, order=order, ko_sigma=ko_sigma, bc="periodic",
                       scheme=scheme, integrator=integrator, ghosts="carried")
    cs = ConstraintSolver(g, order=order)
    dt, total, chunk = chunking(N, crossings, samples)
    mink = idata.minkowski(g, device=device)
    U = id_fn(id_name, amp)(g, device)
    t = 0.0
    rows = [sample_row(U, t, dt, cs, mink)]
    if verbose:
        print(f"# {id_name}  N={N}  scheme={ev.solver.scheme}  order={order}  ko={ko_sigma}  "
              f"cahd_c={params.cahd_c}  dt={dt:.5e}  total_steps={total}  chunk={chunk}",
              flush=True)
    for _ in range(samples):
        U = ev.evolve(U, dt, chunk, t0=t).data
        t += chunk * dt
        rows.append(sample_row(U, t, dt, cs, mink))
        r = rows[-1]
        if verbose:
            print(f"  t={r['ncross']:8.3f}cross  steps={r['steps']:7d}  ||H||={r['H']:.4e}  "
                  f"||M||={r['M']:.4e}  max|a|={r['max_alpha']:.5f}  "
                  f"max-dev={r['max_dev']:.3e}  finite={r['finite']}", flush=True)
        if not r["finite"] or r["max_alpha"] > max_alpha_stop:
            if verbose:
                print(f"# stopped at t={r['ncross']:.3f} crossing times: finite={r['finite']}, "
                      f"max|alpha|={r['max_alpha']:.3e}", flush=True)
            break
    if out:
        with open(out, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=list(rows[0].keys()))
            w.writeheader()
            w.writerows(rows)
    return rows


def fit_rates(rows: List[Dict]) -> Dict[str, float]:
    """Least-squares rates per crossing time (longrun_stability.py:76-107): *_lin = slope of
    ||C|| against t, *_exp = slope of ln ||C|| (e-folds per crossing), M_lin_vs_exp_resid =
    RMS residual of the linear fit of ||M|| over that of the exponential fit (< 1: the drift is
    closer to linear than exponential). Needs 3 or more samples."""
    if len(rows) < 3:
        raise ValueError("fit_rates needs at least 3 samples")
    t = np.array([r["ncross"] for r in rows], dtype=np.float64)
    cols = {c: np.array([r[c] for r in rows], dtype=np.float64) for c in ("H", "M")}
    res = {}
    resid = {}
    for c, y in cols.items():
        p = np.polyfit(t, y, 1)
        res[c + "_lin"] = float(p[0])
        resid[c + "_lin"] = math.sqrt(float(np.mean((np.polyval(p, t) - y) ** 2)))
        q = np.polyfit(t, np.log(np.maximum(y, 1e-300)), 1)
        res[c + "_exp"] = float(q[0])
        resid[c + "_exp"] = math.sqrt(float(np.mean((np.exp(np.polyval(q, t)) - y) ** 2)))
    res["M_lin_vs_exp_resid"] = resid["M_lin"] / max(resid["M_exp"], 1e-300)
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="BSSN long-run stability series")
    ap.add_argument("--id", default="gauge", choices=IDS)
    ap.add_argument("--N", type=int, default=64)
    ap.add_argument("--crossings", type=float, default=20.0)
    ap.add_argument("--samples", type=int, default=40)
    ap.add_argument("--ko", type=float, default=0.4)
    ap.add_argument("--amp", type=float, default=0.01)
    ap.add_argument("--scheme", default="fused")
    ap.add_argument("--order", type=int, default=6)
    ap.add_argument("--cahd-c", type=float, default=None)
    ap.add_argument("--max-alpha", type=float, default=10.0)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    rows = run(a.id, N=a.N, crossings=a.crossings, samples=a.samples, ko_sigma=a.ko,
               scheme=a.scheme, amp=a.amp, out=a.out, max_alpha_stop=a.max_alpha,
               cahd_c=a.cahd_c, order=a.order, verbose=True)
    if len(rows) >= 3:
        r = fit_rates(rows)
        print("# rates/crossing: " + "  ".join(f"{k}={v:+.4e}" for k, v in r.items()))
    return 0


if __name__ == "__main__":
    sys.exit(main())
