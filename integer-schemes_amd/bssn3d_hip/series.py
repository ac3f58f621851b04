"""Constraint-norm series of long evolutions on the device, and their secular rates.

The reference watches long runs of small periodic test problems for secular constraint growth
(3D/src/bssn3d/longrun_stability.py:76-188: evolve, sample ||H||, ||M||, max|alpha| and the
distance from flat space at regular intervals, stop at a blow-up, fit linear and exponential
rates per crossing time). This module does the same on the device with its own layout: a
``SeriesSpec`` says what to evolve and how often to look; ``record`` returns a ``Series`` whose
samples are columns (numpy arrays), so comparisons and fits are array expressions.

Conventions kept from the reference, because they decide the numbers:
  * the problem lives on the periodic unit cube with n interior points per axis; one crossing
    time is 1, dt = 0.25 h (CFL 0.25);
  * the samples are taken every round(total / nsample) steps, total = round(span / dt);
  * the step keeps the reference's carried ghost values (``ghosts="carried"``): its constraint
    norms and maxima read the ghost points, so the series equal the reference's only then;
  * a run stops after the first sample that is non-finite or has max|alpha| above alpha_cap.
The reference fixes FD order 6; the order is a field of the spec so that the INT8-split scheme
(FD order 4 only) can be recorded as well.
"""
from __future__ import annotations

import argparse
import math
import sys
from dataclasses import dataclass, replace
from typing import Callable, Dict, Optional, Tuple

import numpy as np
import torch

from . import initial_data as idata
from .constraints import ConstraintSolver
from .evolve import BSSNEvolution
from .grid import Grid
from .state import ALPHA, PhysicsParams

CFL = 0.25

# Initial data of the long-run problems by name: (grid, amplitude, device) -> (24, S, S, S).
# The reference's choices (longrun_stability.py:110-121): a gauge wave of wavelength 1,
# Minkowski with uniform noise drawn with seed 1, flat space, a Hamiltonian-violating bump.
INITIAL_DATA: Dict[str, Callable[[Grid, float, str], torch.Tensor]] = {
    "gauge": lambda g, a, dev: idata.gauge_wave(g, amplitude=a, wavelength=1.0, device=dev),
    "robust": lambda g, a, dev: idata.robust_stability(g, amp=a, seed=1, device=dev),
    "minkowski": lambda g, a, dev: idata.minkowski(g, device=dev),
    "hambump": lambda g, a, dev: idata.hamiltonian_bump(g, amplitude=a, device=dev),
}

COLUMNS = ("t", "step", "H", "M", "alpha_max", "dev_max", "finite")


@dataclass(frozen=True)
class SeriesSpec:
    """One long run: initial data ``data`` of amplitude ``amplitude`` on n^3 interior points,
    evolved for ``span`` crossing times and sampled ``nsample`` times after the start."""
    data: str
    n: int
    span: float
    nsample: int
    ko_sigma: float
    amplitude: float
    order: int = 6
    cahd_c: Optional[float] = None
    scheme: str = "fused"
    integrator: str = "rk4"
    alpha_cap: float = 10.0

    def cadence(self) -> Tuple[float, int, int]:
        """(dt, total steps, steps between samples)."""
        dt = CFL / self.n
        total = round(self.span / dt)
        return dt, total, max(1, round(total / self.nsample))

    def params(self) -> PhysicsParams:
        return PhysicsParams() if self.cahd_c is None else PhysicsParams(cahd_c=self.cahd_c)


@dataclass
class Series:
    """Samples of a long run as columns: time in crossing times, step count, interior RMS of
    the Hamiltonian and momentum constraints, max|alpha|, max|U - flat space|, finiteness."""
    t: np.ndarray
    step: np.ndarray
    H: np.ndarray
    M: np.ndarray
    alpha_max: np.ndarray
    dev_max: np.ndarray
    finite: np.ndarray

    @classmethod
    def empty(cls) -> "Series":
        return cls(*(np.zeros(0, dtype=bool if c == "finite" else
                              (np.int64 if c == "step" else np.float64)) for c in COLUMNS))

    def __len__(self) -> int:
        return int(self.t.size)

    def append(self, **sample) -> None:
        for c in COLUMNS:
            setattr(self, c, np.append(getattr(self, c), sample[c]))

    def column(self, name: str) -> np.ndarray:
        return getattr(self, name)

    @property
    def all_finite(self) -> bool:
        return bool(self.finite.all())

    def save(self, path: str) -> None:
        """Whitespace-separated text, one sample per line, a commented header."""
        table = np.column_stack([getattr(self, c).astype(np.float64) for c in COLUMNS])
        np.savetxt(path, table, header=" ".join(COLUMNS), fmt="%.17g")


def _observe(U: torch.Tensor, t: float, step: int, cs: ConstraintSolver,
             flat: torch.Tensor) -> dict:
    H, M = cs.l2(U)
    return dict(t=float(t), step=int(step), H=float(H), M=float(M),
                alpha_max=float(U[ALPHA].abs().max()),
                dev_max=float((U - flat).abs().max()),
                finite=bool(torch.isfinite(U).all()))


def record(spec: SeriesSpec, device="cuda", log: Optional[Callable[[str], None]] = None) -> Series:
    """Evolve ``spec`` on the device and return its samples (the first at t = 0)."""
    if spec.data not in INITIAL_DATA:
        raise ValueError(f"unknown initial data {spec.data!r}: one of {sorted(INITIAL_DATA)}")
    grid = Grid.from_domain(spec.n, order=spec.order)
    params = spec.params()
    evo = BSSNEvolution(grid, params, order=spec.order, ko_sigma=spec.ko_sigma, bc="periodic",
                        scheme=spec.scheme, integrator=spec.integrator, ghosts="carried")
    cs = ConstraintSolver(grid, order=spec.order)
    flat = idata.minkowski(grid, device=device)
    dt, total, every = spec.cadence()
    U = INITIAL_DATA[spec.data](grid, spec.amplitude, device)
    out = Series.empty()
    out.append(**_observe(U, 0.0, 0, cs, flat))
    if log:
        log(f"{spec.data}: n={spec.n} scheme={evo.solver.scheme} order={spec.order} "
            f"ko={spec.ko_sigma} cahd_c={params.cahd_c} dt={dt:.5e} steps={total} every={every}")
    t, step = 0.0, 0
    for _ in range(spec.nsample):
        U = evo.evolve(U, dt, every, t0=t).data
        t += every * dt
        step += every
        s = _observe(U, t, step, cs, flat)
        out.append(**s)
        if log:
            log(f"  t {s['t']:8.3f}  step {s['step']:7d}  H {s['H']:.4e}  M {s['M']:.4e}  "
                f"alpha_max {s['alpha_max']:.5f}  dev_max {s['dev_max']:.3e}  "
                f"finite {s['finite']}")
        if not s["finite"] or s["alpha_max"] > spec.alpha_cap:
            if log:
                log(f"  blow-up at t = {s['t']:.3f}: finite {s['finite']}, "
                    f"alpha_max {s['alpha_max']:.3e}")
            break
    return out


def secular_rates(series: Series) -> Dict[str, float]:
    """Least-squares rates per crossing time of ||H|| and ||M||: ``*_slope`` of the norm against
    t, ``*_efold`` of its logarithm (e-folds per crossing), and ``M_line_vs_exp``, the RMS
    residual of M's straight-line fit over that of its exponential fit (below 1: the drift is
    closer to linear than to exponential). Needs three samples or more."""
    if len(series) < 3:
        raise ValueError("secular_rates needs at least 3 samples")
    t = np.asarray(series.t, dtype=np.float64)
    A = np.column_stack([t, np.ones_like(t)])
    out, resid = {}, {}
    for name in ("H", "M"):
        y = np.asarray(series.column(name), dtype=np.float64)
        (a, b), *_ = np.linalg.lstsq(A, y, rcond=None)
        out[name + "_slope"] = float(a)
        resid[name + "_line"] = math.sqrt(float(np.mean((A @ np.array([a, b]) - y) ** 2)))
        (c, d), *_ = np.linalg.lstsq(A, np.log(np.maximum(y, 1e-300)), rcond=None)
        out[name + "_efold"] = float(c)
        resid[name + "_exp"] = math.sqrt(float(np.mean((np.exp(A @ np.array([c, d])) - y) ** 2)))
    out["M_line_vs_exp"] = resid["M_line"] / max(resid["M_exp"], 1e-300)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="record a constraint-norm series of a long run")
    ap.add_argument("data", nargs="?", default="gauge", choices=sorted(INITIAL_DATA))
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--span", type=float, default=20.0, help="crossing times")
    ap.add_argument("--nsample", type=int, default=40)
    ap.add_argument("--ko", type=float, default=0.4)
    ap.add_argument("--amplitude", type=float, default=0.01)
    ap.add_argument("--scheme", default="fused")
    ap.add_argument("--order", type=int, default=6)
    ap.add_argument("--cahd-c", type=float, default=None)
    ap.add_argument("--alpha-cap", type=float, default=10.0)
    ap.add_argument("--save", default=None, help="write the samples to this text file")
    a = ap.parse_args(argv)
    spec = SeriesSpec(a.data, a.n, a.span, a.nsample, a.ko, a.amplitude, order=a.order,
                      cahd_c=a.cahd_c, scheme=a.scheme, alpha_cap=a.alpha_cap)
    s = record(spec, log=lambda m: print(m, flush=True))
    if a.save:
        s.save(a.save)
    if len(s) >= 3:
        print("rates per crossing: " + "  ".join(f"{k} {v:+.4e}"
                                                  for k, v in secular_rates(s).items()))
    return 0


def with_scheme(spec: SeriesSpec, scheme: str, order: Optional[int] = None) -> SeriesSpec:
    """The same run with another scheme (and FD order)."""
    return replace(spec, scheme=scheme, order=spec.order if order is None else order)


if __name__ == "__main__":
    sys.exit(main())
