"""Tracking gate: does a reassociating scheme follow the reference scheme over long runs?

A scheme that reorders the arithmetic (the fused kernel) or takes its derivatives from int8
digits (the INT8-split form) cannot be held to bit parity with the verbatim data flow. The
reference accepts such a scheme when its long-run constraint series track the verbatim ones
(3D/src/bssn3d/apples_gate.py:38-112): on a suite of short periodic problems, every sampled
quantity of the candidate stays within a tolerance of the reference's in the floored relative
measure |a - b| / (|b| + floor), and both runs end the same way (same number of samples, the
same finiteness at every sample).

Here both series are ``series.Series`` columns, so a case's deviation is one array expression
per quantity; the verdict of a gate is a ``GateReport`` that can print itself. The series come
from ``recorder`` (``series.record`` by default), which is also how tests feed synthetic series.
The reference role defaults to the two-pass scheme (materialised bundle, then the algebra: the
verbatim data flow) at the reference's FD order 6; the INT8-split form is gated at order 4.
"""
from __future__ import annotations

import argparse
import sys
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Sequence

import numpy as np

from .series import Series, SeriesSpec, record

QUANTITIES = ("H", "M", "alpha_max", "dev_max")


@dataclass(frozen=True)
class GateCase:
    """One problem of the suite (a long run, as a SeriesSpec without scheme and order)."""
    data: str
    n: int
    span: float
    nsample: int
    ko_sigma: float
    amplitude: float
    purpose: str

    def spec(self, scheme: str, order: int) -> SeriesSpec:
        return SeriesSpec(self.data, self.n, self.span, self.nsample, self.ko_sigma,
                          self.amplitude, order=order, scheme=scheme)


# The reference's suite (apples_gate.py:47-52): three problems on 24^3 points, two crossing
# times, eight samples each.
SUITE: Sequence[GateCase] = (
    GateCase("robust", 24, 2.0, 8, 0.1, 1e-8, "flat space plus noise: H damped, stays bounded"),
    GateCase("gauge", 24, 2.0, 8, 0.1, 1e-2, "gauge wave: bounded momentum-constraint drift"),
    GateCase("hambump", 24, 2.0, 8, 0.1, 1e-3, "coherent H bump: damped by the CAHD term"),
)


def deviation(cand: Series, ref: Series, floor: float) -> Dict[str, float]:
    """Largest floored relative deviation of each quantity over the samples both runs hold."""
    n = min(len(cand), len(ref))
    out = {}
    for q in QUANTITIES:
        a = np.asarray(cand.column(q)[:n], dtype=np.float64)
        b = np.asarray(ref.column(q)[:n], dtype=np.float64)
        out[q] = float(np.max(np.abs(a - b) / (np.abs(b) + floor))) if n else 0.0
    return out


def same_ending(cand: Series, ref: Series) -> bool:
    """Both runs stopped after the same number of samples with the same finiteness throughout."""
    return len(cand) == len(ref) and bool(np.array_equal(cand.finite, ref.finite))


@dataclass
class CaseResult:
    case: GateCase
    dev: Dict[str, float]
    samples: tuple          # (reference, candidate)
    same_ending: bool

    @property
    def worst(self) -> float:
        return max(self.dev.values())

    def passed(self, tol: float) -> bool:
        return self.same_ending and self.worst <= tol


@dataclass
class GateReport:
    candidate: str
    reference: str
    order: int
    tol: float
    cases: List[CaseResult] = field(default_factory=list)

    @property
    def passed(self) -> bool:
        return all(c.passed(self.tol) for c in self.cases)

    @property
    def worst(self) -> float:
        return max((c.worst for c in self.cases), default=0.0)

    def table(self) -> str:
        head = (f"tracking gate: {self.candidate} against {self.reference}, FD order "
                f"{self.order}, tolerance {self.tol:.1e}")
        cols = "  ".join(f"{q:>10s}" for q in QUANTITIES)
        rows = [head, f"  {'problem':10s}  {cols}  ending     result"]
        for c in self.cases:
            vals = "  ".join(f"{c.dev[q]:10.2e}" for q in QUANTITIES)
            end = "same" if c.same_ending else f"{c.samples[0]} vs {c.samples[1]}"
            rows.append(f"  {c.case.data:10s}  {vals}  {end:9s}  "
                        f"{'pass' if c.passed(self.tol) else 'FAIL'}")
        rows.append(f"  overall: {'pass' if self.passed else 'FAIL'} (worst {self.worst:.2e})")
        return "\n".join(rows)


def run_gate(candidate: str, reference: str = "twopass", order: int = 6, tol: float = 1e-6,
             floor: float = 1e-12, suite: Sequence[GateCase] = SUITE,
             recorder: Callable[..., Series] = record, **record_kw) -> GateReport:
    """Record every case of the suite with both schemes and compare them."""
    rep = GateReport(candidate, reference, order, tol)
    for case in suite:
        ref = recorder(case.spec(reference, order), **record_kw)
        cand = recorder(case.spec(candidate, order), **record_kw)
        rep.cases.append(CaseResult(case, deviation(cand, ref, floor), (len(ref), len(cand)),
                                    same_ending(cand, ref)))
    return rep


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="tracking gate of a BSSN scheme")
    ap.add_argument("candidate", nargs="?", default="fused")
    ap.add_argument("--reference", default="twopass")
    ap.add_argument("--order", type=int, default=6)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--first", action="store_true", help="run the first problem only")
    a = ap.parse_args(argv)
    rep = run_gate(a.candidate, a.reference, a.order, a.tol,
                   suite=SUITE[:1] if a.first else SUITE)
    print(rep.table())
    return 0 if rep.passed else 1


if __name__ == "__main__":
    sys.exit(main())
