"""GPU: carried-ghost steps, the long-run series and the apples gate.

Bars:
  * the carried-ghost RK4 and MSRK steps (periodic = 2) against the oracle's periodic steps over
    the WHOLE padded array, ghosts included: normwise <= 1e-12 per field; their interiors equal
    the default (wrapped, synced) step's interiors bit for bit;
  * the same for the composed two-pass step with carried ghosts;
  * the device long-run series (fused) against the oracle's series sample by sample (relative
    1e-9) and against the reference's calibrated CAHD decay rates (3D/tests/validation/
    test_bssn_cahd_rate.py:16-18) to their printed digits;
  * the apples gate (3D/src/bssn3d/apples_gate.py) on the reference's suite: fused against the
    two-pass (verbatim) scheme at FD order 6, and the INT8-split scheme at FD order 4.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import longrun as OL
import oracle as O
from conftest import gpu_available, normwise

from bssn3d_hip import apples_gate as ag
from bssn3d_hip import initial_data as idata
from bssn3d_hip.evolve import BSSNEvolution
from bssn3d_hip.grid import Grid
from bssn3d_hip.longrun import chunking, fit_rates, run
from bssn3d_hip.state import PhysicsParams, scalars

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DEV = "cuda:0"
CAHD_GOLDEN = [(0.0, 0.001, 0.0005), (0.03, -2.14, 0.005), (0.06, -3.05, 0.005)]


def say(msg):
    print(f"[case] {msg}", flush=True)


def _state(N, order):
    """The single-puncture workload data with 4 ghosts (every field carries structure, so the
    per-field normwise bar is well conditioned), on the cube [-6, 6)^3."""
    g, U, _ = idata.config_single_puncture(seed=5, N=N, ng=4)
    return g, U


@pytest.mark.parametrize("integrator,order,nsteps", [("rk4", 6, 3), ("rk4", 4, 2),
                                                     ("rk4_2_2", 6, 4)])
def test_carried_step_matches_oracle(integrator, order, nsteps):
    say(f"carried ghosts {integrator} order {order} x{nsteps}")
    g, U0 = _state(16, order)
    p = PhysicsParams()
    dt = 0.25 * g.dx
    ev = BSSNEvolution(g, p, order=order, ko_sigma=0.1, bc="periodic", integrator=integrator,
                       ghosts="carried")
    got = ev.evolve(U0.to(DEV), dt, nsteps, t0=0.2).data.cpu().numpy()
    S16 = scalars(p, g.dx, g.dy, g.dz, dt, 0.0)
    want = O.evolve(U0.numpy(), S16, dt, nsteps, t0=0.2, integrator=integrator, order=order,
                    ko_order=order + 2, ko_sigma=0.1, ng=g.ng, periodic=True)
    e = normwise(got, want)
    print(f"whole array: max {e.max():.2e}")
    assert e.max() <= 1e-12
    syn = BSSNEvolution(g, p, order=order, ko_sigma=0.1, bc="periodic", integrator=integrator)
    s = syn.evolve(U0.to(DEV), dt, nsteps, t0=0.2).data.cpu().numpy()
    I = (slice(None),) + g.interior
    assert np.array_equal(got[I], s[I])
    assert not np.array_equal(got, s)              # the carried ghosts are not the synced ones


def test_carried_step_devtime_graph():
    say("carried ghosts, device time, graph replay")
    g, U0 = _state(16, 6)
    dt = 0.25 * g.dx
    ev = BSSNEvolution(g, PhysicsParams(), order=6, ko_sigma=0.1, bc="periodic",
                       ghosts="carried")
    a = ev.evolve(U0.to(DEV), dt, 6, t0=0.1).data
    b = ev.evolve(U0.to(DEV), dt, 6, t0=0.1, graph_chunk=3).data
    e = normwise(b.cpu().numpy(), a.cpu().numpy())
    assert e.max() <= 1e-12          # the graph forms the SSL factor with the device's exp


def test_composed_carried_matches_oracle():
    say("composed two-pass step, carried ghosts")
    g, U0 = _state(16, 6)
    p = PhysicsParams()
    dt = 0.25 * g.dx
    ev = BSSNEvolution(g, p, order=6, ko_sigma=0.1, bc="periodic", scheme="twopass",
                       ghosts="carried")
    got = ev.evolve(U0.to(DEV), dt, 2, t0=0.0).data.cpu().numpy()
    S16 = scalars(p, g.dx, g.dy, g.dz, dt, 0.0)
    want = O.evolve(U0.numpy(), S16, dt, 2, order=6, ko_order=8, ko_sigma=0.1, ng=g.ng,
                    periodic=True)
    e = normwise(got, want)
    print(f"whole array: max {e.max():.2e}")
    assert e.max() <= 1e-12


def test_longrun_rows_match_oracle_series():
    say("long-run rows vs oracle (hambump N=12)")
    rows = run("hambump", N=12, crossings=1.0, samples=4, ko_sigma=0.02, amp=0.005,
               cahd_c=0.06, device=DEV)
    g = Grid.from_domain(12, order=6)
    p = PhysicsParams(cahd_c=0.06)
    dt, total, chunk = chunking(12, 1.0, 4)
    S16 = scalars(p, g.dx, g.dy, g.dz, dt, 0.0)
    want, _ = OL.series(OL.hambump(g.shape, g.ng, g.dx, 0.005), S16, dt, g.dx, g.ng, total,
                        chunk, 4, 0.02)
    assert len(rows) == len(want) == 5
    for a, b in zip(rows, want):
        assert a["steps"] == b["steps"] and a["finite"] and b["finite"]
        for m in ("H", "M", "max_alpha", "max_dev"):
            assert ag._rel(a[m], b[m], 1e-12) <= 1e-9, (m, a[m], b[m])


@pytest.mark.parametrize("cahd_c,golden,tol", CAHD_GOLDEN)
def test_longrun_cahd_rates_match_reference(cahd_c, golden, tol):
    say(f"CAHD rate cahd_c={cahd_c}")
    rows = run("hambump", N=12, crossings=2.0, samples=8, ko_sigma=0.02, amp=0.005,
               cahd_c=cahd_c, device=DEV)
    rate = fit_rates(rows)["H_exp"]
    print(f"H_exp {rate:+.4f} (recorded {golden:+.3f})")
    assert len(rows) == 9 and abs(rate - golden) <= tol


def test_longrun_reference_assertions():
    """The reference's long-run assertions (test_bssn_constraint_growth.py, test_bssn_longrun.py)
    on the device series."""
    say("robust / gauge long runs N=12")
    rows = run("robust", N=12, crossings=2.0, samples=8, ko_sigma=0.1, amp=1e-8, device=DEV)
    r = fit_rates(rows)
    assert all(x["finite"] for x in rows) and r["H_exp"] < 0.2 and rows[-1]["H"] <= rows[0]["H"]
    assert max(x["max_dev"] for x in rows) < 1e-3
    rows = run("gauge", N=12, crossings=2.0, samples=8, ko_sigma=0.1, amp=0.01, device=DEV)
    r = fit_rates(rows)
    assert all(x["finite"] for x in rows) and r["M_lin"] > 0.0 and rows[-1]["M"] < 0.5
    assert r["M_lin_vs_exp_resid"] < 3.0 and max(x["max_alpha"] for x in rows) < 2.0


@pytest.mark.parametrize("scheme,order", [("fused", 6), ("int8", 4)])
def test_apples_gate(scheme, order):
    say(f"apples gate {scheme} order {order}")
    res = []
    ok = ag.gate(scheme, order=order, results=res, device=DEV)
    for r in res:
        print(r["id"], {k: f"{v:.2e}" for k, v in r["dev"].items()}, r["verdict_match"])
    assert ok
