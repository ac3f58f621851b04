"""Apples-with-apples acceptance gate for a reassociating scheme (3D/src/bssn3d/apples_gate.py).

A scheme whose arithmetic is reordered (or whose derivatives come from int8 digits) cannot be
held to bit parity; instead its long-run constraint trajectories must TRACK those of the
reference scheme. For each configuration of the suite both schemes run the same long-run series
(longrun.run) and, sample by sample, the candidate's H, M, max|alpha| and max-deviation must
stay within `tol` of the reference's in the floored relative measure
|a - b| / (|b| + floor); both runs must also reach the same verdict (same number of samples,
i.e. the same early-stop point, and the same finiteness at every sample).

The reference scheme defaults to "twopass" (the reference's verbatim data flow: materialised
derivative bundle, then the algebra) and the candidate to "fused" -- the reference's own
default gate, verbatim against cuda_fused. `order` (6 in the reference) is a parameter so the
INT8-split scheme (FD order 4 only) can be gated: gate("int8", order=4).
"""
from __future__ import annotations

import argparse
import sys
from dataclasses import dataclass
from typing import Dict, List

from .longrun import run


@dataclass
class Config:
    id_name: str
    N: int
    crossings: float
    samples: int
    ko_sigma: float
    amp: float
    note: str


# the reference's suite (apples_gate.py:47-52): small, 2 crossing times each
GATE_SUITE: List[Config] = [
    Config("robust", 24, 2.0, 8, 0.1, 1e-8, "CAHD damps H on Minkowski + noise (bounded)"),
    Config("gauge", 24, 2.0, 8, 0.1, 1e-2, "gauge-wave momentum drift, bounded"),
    Config("hambump", 24, 2.0, 8, 0.1, 1e-3, "CAHD decay of a coherent H bump"),
]

METRICS = ("H", "M", "max_alpha", "max_dev")


def _rel(a: float, b: float, floor: float) -> float:
    """|a - b| / (|b| + floor): relative, floored so that near-zero values do not blow it up."""
    return abs(a - b) / (abs(b) + floor)


def compare_rows(ref: List[Dict], cand: List[Dict], floor: float) -> Dict:
    """Per-metric maximum deviation over the common samples, and whether the verdicts match."""
    n = min(len(ref), len(cand))
    dev = {m: 0.0 for m in METRICS}
    for i in range(n):
        for m in METRICS:
            dev[m] = max(dev[m], _rel(cand[i][m], ref[i][m], floor))
    verdict = len(ref) == len(cand) and all(ref[i]["finite"] == cand[i]["finite"]
                                            for i in range(n))
    return dict(dev=dev, n_ref=len(ref), n_cand=len(cand), verdict_match=verdict)


def compare_one(cfg: Config, scheme: str, floor: float, reference: str = "twopass",
                order: int = 6, verbose: bool = True, **kw) -> Dict:
    """Run the reference scheme and the candidate on one configuration and compare."""
    if verbose:
        print(f"\n=== apples gate :: {cfg.id_name} (N={cfg.N}, {cfg.crossings} crossings) -- "
              f"{cfg.note} ===", flush=True)
    common = dict(N=cfg.N, crossings=cfg.crossings, samples=cfg.samples,
                  ko_sigma=cfg.ko_sigma, amp=cfg.amp, order=order, verbose=verbose, **kw)
    ref = run(cfg.id_name, scheme=reference, **common)
    cand = run(cfg.id_name, scheme=scheme, **common)
    r = compare_rows(ref, cand, floor)
    r["id"] = cfg.id_name
    return r


def gate(scheme: str, tol: float = 1e-6, floor: float = 1e-12, suite: List[Config] = GATE_SUITE,
         quick: bool = False, reference: str = "twopass", order: int = 6,
         verbose: bool = True, results: List[Dict] | None = None, **kw) -> bool:
    """True when the candidate tracks the reference scheme on every configuration (worst
    deviation <= tol and matching verdicts). `results`, if given, receives the per-config
    comparisons."""
    if quick:
        suite = suite[:1]
    res = [compare_one(c, scheme, floor, reference, order, verbose, **kw) for c in suite]
    if results is not None:
        results.extend(res)
    overall = True
    lines = [f"APPLES GATE :: candidate {scheme!r} vs {reference!r}  order {order}  "
             f"tol {tol:.1e}",
             f"{'config':>10} | {'maxdev H':>10} {'maxdev M':>10} {'max|a|':>10} "
             f"{'max-dev':>10} | verdict"]
    for r in res:
        d = r["dev"]
        ok = max(d.values()) <= tol and r["verdict_match"]
        overall &= ok
        v = "ok" if r["verdict_match"] else f"MISMATCH({r['n_ref']}!={r['n_cand']})"
        lines.append(f"{r['id']:>10} | {d['H']:>10.2e} {d['M']:>10.2e} "
                     f"{d['max_alpha']:>10.2e} {d['max_dev']:>10.2e} | {v} "
                     f"{'PASS' if ok else 'FAIL'}")
    lines.append(f">> {'PASS' if overall else 'FAIL'}")
    if verbose:
        print("\n" + "\n".join(lines), flush=True)
    return overall


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="apples-with-apples gate for a BSSN scheme")
    ap.add_argument("--scheme", default="fused")
    ap.add_argument("--reference", default="twopass")
    ap.add_argument("--order", type=int, default=6)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--quick", action="store_true")
    a = ap.parse_args(argv)
    ok = gate(a.scheme, tol=a.tol, quick=a.quick, reference=a.reference, order=a.order)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
