"""GPU: carried-ghost steps, the long-run series and the apples gate.

Bars:
  * the carried-ghost RK4 and MSRK steps (periodic = 2) against the oracle's periodic steps over
    the WHOLE padded array, ghosts included: normwise <= 1e-12 per field; their interiors equal
    the default (wrapped, synced) step's interiors bit for bit;
  * the same for the composed two-pass step with carried ghosts;
  * the device long-run series (fused) against the oracle's series sample by sample (relative
    1e-9) and against the reference's calibrated CAHD decay rates (3D/tests/validation/
    test_bssn_cahd_rate.py:16-18) to their printed digits;
  * the tracking gate (the reference's apples gate, 3D/src/bssn3d/apples_gate.py) on the
    reference's suite: fused against the two-pass (verbatim) scheme at FD order 6, and the
    INT8-split scheme at FD order 4.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import longrun as OL
import oracle as O
from conftest import gpu_available, normwise

from bssn3d_hip import tracking_gate as tg
from bssn3d_hip import initial_data as idata
from bssn3d_hip.evolve import BSSNEvolution
from bssn3d_hip.grid import Grid
from bssn3d_hip.series import SeriesSpec, record, secular_rates
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
    say("long-run series vs oracle (hambump N=12)")
    spec = SeriesSpec("hambump", 12, 1.0, 4, 0.02, 0.005, cahd_c=0.06)
    got = record(spec, device=DEV)
    g = Grid.from_domain(12, order=6)
    p = PhysicsParams(cahd_c=0.06)
    dt, total, chunk = spec.cadence()
    S16 = scalars(p, g.dx, g.dy, g.dz, dt, 0.0)
    want, _ = OL.series(OL.hambump(g.shape, g.ng, g.dx, 0.005), S16, dt, g.dx, g.ng, total,
                        chunk, 4, 0.02)
    assert len(got) == len(want) == 5 and got.all_finite
    assert got.step.tolist() == [w["steps"] for w in want]
    for mine, theirs in (("H", "H"), ("M", "M"), ("alpha_max", "max_alpha"),
                         ("dev_max", "max_dev")):
        b = np.array([w[theirs] for w in want])
        rel = np.abs(got.column(mine) - b) / (np.abs(b) + 1e-12)
        assert rel.max() <= 1e-9, (mine, rel)


@pytest.mark.parametrize("cahd_c,golden,tol", CAHD_GOLDEN)
def test_longrun_cahd_rates_match_reference(cahd_c, golden, tol):
    say(f"CAHD rate cahd_c={cahd_c}")
    s = record(SeriesSpec("hambump", 12, 2.0, 8, 0.02, 0.005, cahd_c=cahd_c), device=DEV)
    rate = secular_rates(s)["H_efold"]
    print(f"H e-folds {rate:+.4f} (recorded {golden:+.3f})")
    assert len(s) == 9 and abs(rate - golden) <= tol


def test_longrun_reference_assertions():
    """The reference's long-run assertions (test_bssn_constraint_growth.py, test_bssn_longrun.py)
    on the device series."""
    say("robust / gauge long runs N=12")
    s = record(SeriesSpec("robust", 12, 2.0, 8, 0.1, 1e-8), device=DEV)
    r = secular_rates(s)
    assert s.all_finite and r["H_efold"] < 0.2 and s.H[-1] <= s.H[0]
    assert s.dev_max.max() < 1e-3
    s = record(SeriesSpec("gauge", 12, 2.0, 8, 0.1, 0.01), device=DEV)
    r = secular_rates(s)
    assert s.all_finite and r["M_slope"] > 0.0 and s.M[-1] < 0.5
    assert r["M_line_vs_exp"] < 3.0 and s.alpha_max.max() < 2.0


@pytest.mark.parametrize("scheme,order", [("fused", 6), ("int8", 4)])
def test_tracking_gate(scheme, order):
    say(f"tracking gate {scheme} order {order}")
    rep = tg.run_gate(scheme, order=order, device=DEV)
    print(rep.table())
    assert rep.passed
