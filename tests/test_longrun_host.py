"""CPU: the long-run series and the apples gate -- oracle pins and host logic.

* The oracle's long-run series (oracle/longrun.py) against the calibrated Hamiltonian decay
  rates the reference records for its CAHD certificate (3D/tests/validation/
  test_bssn_cahd_rate.py:16-18: hambump, N=12, 2 crossings, 8 samples, KO 0.02, amp 0.005,
  cahd_c = 0 / 0.03 / 0.06 -> H_exp = +0.001 / -2.14 / -3.05 e-folds per crossing). The same
  series with the ghosts synced before sampling gives -2.21 / -4.41 and fails the pin: the
  reference's constraint norms read the ghost values its step carries.
* The reference's qualitative long-run assertions on robust noise and the gauge wave
  (3D/tests/validation/test_bssn_constraint_growth.py, 3D/tests/integration/
  test_bssn_longrun.py), on the oracle at N=12.
* fit_rates on closed forms, the initial data, the chunking, and the gate's comparison logic
  with synthetic trajectories (3D/tests/unit/test_apples_gate.py).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import longrun as OL
import oracle as O

from bssn3d_hip import apples_gate as ag
from bssn3d_hip import initial_data as idata
from bssn3d_hip.grid import Grid
from bssn3d_hip.longrun import chunking, fit_rates
from bssn3d_hip.state import PhysicsParams, scalars

# (cahd_c, recorded H_exp, half of the last printed digit)
CAHD_GOLDEN = [(0.0, 0.001, 0.0005), (0.03, -2.14, 0.005), (0.06, -3.05, 0.005)]


def oracle_run(U0, N, crossings, samples, ko_sigma, cahd_c=None, order=6, sync_before=False):
    g = Grid.from_domain(N, order=order)
    p = PhysicsParams() if cahd_c is None else PhysicsParams(cahd_c=cahd_c)
    dt, total, chunk = chunking(N, crossings, samples)
    S16 = scalars(p, g.dx, g.dy, g.dz, dt, 0.0)
    rows, U = OL.series(U0, S16, dt, g.dx, g.ng, total, chunk, samples, ko_sigma, order,
                        chi_floor=p.chi_floor, alpha_floor=p.alpha_floor)
    if sync_before:
        # the same run sampled with the ghosts synced first (not the reference's convention)
        rows2, Ut = [], np.array(U0)
        t = 0.0
        for r in range(samples + 1):
            if r:
                Ut = O.evolve(Ut, S16, dt, chunk, t0=t, order=order, ko_order=order + 2,
                              ko_sigma=ko_sigma, ng=g.ng, periodic=True)
                t += chunk * dt
            H, M = O.constraint_l2(O.periodic_sync(Ut, g.ng), (g.dx,) * 3, order, g.ng)
            rows2.append(dict(ncross=t, H=H, M=M))
        return rows2
    return rows


def test_chunking_matches_reference_suite():
    assert chunking(24, 2.0, 8) == (1.0 / 96, 192, 24)
    assert chunking(12, 2.0, 8) == (1.0 / 48, 96, 12)
    dt, total, chunk = chunking(64, 20.0, 40)
    assert total == 5120 and chunk == 128


@pytest.mark.parametrize("cahd_c,golden,tol", CAHD_GOLDEN)
def test_oracle_cahd_rates_match_reference_calibration(cahd_c, golden, tol):
    g = Grid.from_domain(12, order=6)
    U0 = OL.hambump(g.shape, g.ng, g.dx, 0.005)
    rows = oracle_run(U0, 12, 2.0, 8, 0.02, cahd_c)
    assert len(rows) == 9 and all(r["finite"] for r in rows)
    rate = fit_rates(rows)["H_exp"]
    print(f"cahd_c {cahd_c}: H_exp {rate:+.4f} (recorded {golden:+.3f})")
    assert abs(rate - golden) <= tol


def test_synced_ghost_sampling_fails_the_pin():
    g = Grid.from_domain(12, order=6)
    U0 = OL.hambump(g.shape, g.ng, g.dx, 0.005)
    rate = fit_rates(oracle_run(U0, 12, 2.0, 8, 0.02, 0.06, sync_before=True))["H_exp"]
    assert abs(rate - (-3.05)) > 0.5


def test_oracle_robust_noise_bounded_and_h_damped():
    """test_bssn_constraint_growth.py:35-44 and test_bssn_longrun.py:86-98 on the oracle."""
    g = Grid.from_domain(12, order=6)
    U0 = idata.robust_stability(g, amp=1e-8, seed=1).numpy()
    rows = oracle_run(U0, 12, 2.0, 8, 0.1)
    r = fit_rates(rows)
    assert all(x["finite"] for x in rows)
    assert r["H_exp"] < 0.2 and rows[-1]["H"] <= rows[0]["H"]
    assert max(x["max_dev"] for x in rows) < 1e-3
    assert max(x["max_alpha"] for x in rows) < 2.0


def test_oracle_gauge_wave_momentum_drift_subexponential():
    """test_bssn_constraint_growth.py:47-60 and test_bssn_longrun.py:74-83 on the oracle."""
    g = Grid.from_domain(12, order=6)
    U0 = idata.gauge_wave(g, amplitude=0.01, wavelength=1.0).numpy()
    rows = oracle_run(U0, 12, 2.0, 8, 0.1)
    r = fit_rates(rows)
    assert all(x["finite"] for x in rows)
    assert r["M_lin"] > 0.0 and rows[-1]["M"] < 0.5
    assert r["M_lin_vs_exp_resid"] < 3.0
    assert max(x["max_alpha"] for x in rows) < 2.0
    assert max(x["H"] for x in rows) < 1.0 and max(x["M"] for x in rows) < 2.0


def test_initial_data_hambump_and_robust():
    g = Grid.from_domain(12, order=6)
    a = idata.hamiltonian_bump(g, amplitude=0.005).numpy()
    b = OL.hambump(g.shape, g.ng, g.dx, 0.005)
    assert np.max(np.abs(a - b)) <= 1e-17
    # chi at the origin-adjacent cell: cos(2 pi x) factors at x = lo + i h
    i = g.ng + 6                                     # x = 0
    assert a[1, i, i, i] == 1.0 + 0.005
    assert np.all(a[np.r_[0, 3, 6, 8]] == 1.0) and np.all(a[np.r_[2, 4, 5, 7, 9:24]] == 0.0)
    r1 = idata.robust_stability(g, amp=1e-8, seed=1).numpy()
    r2 = idata.robust_stability(g, amp=1e-8, seed=1).numpy()
    mink = OL.minkowski(g.shape)
    d = r1 - mink
    assert np.array_equal(r1, r2)
    assert np.all(np.abs(d) <= 1e-8) and np.all(d != 0.0)
    assert abs(float(np.mean(d))) < 1e-10                      # centred noise
    assert not np.array_equal(r1, idata.robust_stability(g, amp=1e-8, seed=2).numpy())


def test_fit_rates_closed_forms():
    t = np.linspace(0.0, 2.0, 9)
    lin = [dict(ncross=x, H=1.0 + 0.5 * x, M=2.0 + 3.0 * x) for x in t]
    r = fit_rates(lin)
    assert r["H_lin"] == pytest.approx(0.5, rel=1e-12)
    assert r["M_lin"] == pytest.approx(3.0, rel=1e-12)
    assert r["M_lin_vs_exp_resid"] < 1e-9                     # a line fits a line exactly
    ex = [dict(ncross=x, H=1e-3 * math.exp(-2.5 * x), M=1e-4 * math.exp(0.7 * x)) for x in t]
    r = fit_rates(ex)
    assert r["H_exp"] == pytest.approx(-2.5, rel=1e-12)
    assert r["M_exp"] == pytest.approx(0.7, rel=1e-12)
    assert r["M_lin_vs_exp_resid"] > 1e6                      # exponential fits exactly
    with pytest.raises(ValueError):
        fit_rates(lin[:2])


# --- gate logic on synthetic trajectories (3D/tests/unit/test_apples_gate.py) ---------------

def _rows(H, M=1e-6, a=1.0, dev=1e-7, finite=True, n=4):
    return [dict(H=H, M=M, max_alpha=a, max_dev=dev, finite=finite) for _ in range(n)]


def _patch(monkeypatch, ref_rows, cand_rows):
    def fake_run(id_name, **kw):
        return ref_rows if kw.get("scheme") == "twopass" else cand_rows
    monkeypatch.setattr(ag, "run", fake_run)


def test_rel_is_floored():
    assert ag._rel(0.0, 0.0, 1e-12) == 0.0
    assert ag._rel(1.0 + 1e-9, 1.0, 1e-12) == pytest.approx(1e-9, rel=1e-3)


def test_identical_trajectories_pass(monkeypatch):
    rows = _rows(7.1e-6)
    _patch(monkeypatch, rows, rows)
    assert ag.gate("fused", tol=1e-6, quick=True, verbose=False) is True


def test_roundoff_drift_within_tol_passes(monkeypatch):
    _patch(monkeypatch, _rows(7.1e-6), _rows(7.1e-6 * (1 + 1e-9)))
    assert ag.gate("fused", tol=1e-6, quick=True, verbose=False) is True


def test_drift_beyond_tol_fails(monkeypatch):
    _patch(monkeypatch, _rows(7.1e-6), _rows(7.1e-6 * (1 + 1e-3)))
    assert ag.gate("fused", tol=1e-6, quick=True, verbose=False) is False


def test_blowup_point_mismatch_fails(monkeypatch):
    _patch(monkeypatch, _rows(7.1e-6, n=8), _rows(7.1e-6, n=4))
    assert ag.gate("fused", tol=1e-6, quick=True, verbose=False) is False


def test_finiteness_disagreement_fails(monkeypatch):
    ref = _rows(7.1e-6, n=4)
    cand = _rows(7.1e-6, n=4)
    cand[-1] = dict(cand[-1], finite=False)
    _patch(monkeypatch, ref, cand)
    assert ag.gate("fused", tol=1e-6, quick=True, verbose=False) is False


def test_gate_runs_the_whole_suite_and_reports(monkeypatch):
    seen = []

    def fake_run(id_name, **kw):
        seen.append((id_name, kw["scheme"], kw["N"], kw["order"]))
        return _rows(1e-5)
    monkeypatch.setattr(ag, "run", fake_run)
    res = []
    assert ag.gate("int8", order=4, verbose=False, results=res) is True
    assert [r["id"] for r in res] == ["robust", "gauge", "hambump"]
    assert seen == [(c.id_name, s, 24, 4) for c in ag.GATE_SUITE for s in ("twopass", "int8")]
