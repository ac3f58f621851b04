"""CPU: the long-run series and the tracking gate -- oracle pins and host logic.

* The oracle's long-run series (oracle/longrun.py) against the calibrated Hamiltonian decay
  rates the reference records for its CAHD certificate (3D/tests/validation/
  test_bssn_cahd_rate.py:16-18: hambump, N=12, 2 crossings, 8 samples, KO 0.02, amp 0.005,
  cahd_c = 0 / 0.03 / 0.06 -> H e-folding rate = +0.001 / -2.14 / -3.05 per crossing). The
  same series with the ghosts synced before sampling gives -2.21 / -4.41 and fails the pin: the
  reference's constraint norms read the ghost values its step carries.
* The reference's qualitative long-run assertions on robust noise and the gauge wave
  (3D/tests/validation/test_bssn_constraint_growth.py, 3D/tests/integration/
  test_bssn_longrun.py), on the oracle at N=12.
* secular_rates on closed forms, the initial data, the sampling cadence, and the gate's
  comparison logic on synthetic series (the cases of 3D/tests/unit/test_apples_gate.py).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import longrun as OL
import oracle as O

from bssn3d_hip import initial_data as idata
from bssn3d_hip import tracking_gate as tg
from bssn3d_hip.grid import Grid
from bssn3d_hip.series import Series, SeriesSpec, secular_rates
from bssn3d_hip.state import PhysicsParams, scalars


def as_series(rows) -> Series:
    """The oracle's sample rows as a Series."""
    s = Series.empty()
    for r in rows:
        s.append(t=r["ncross"], step=r.get("steps", 0), H=r["H"], M=r["M"],
                 alpha_max=r.get("max_alpha", 1.0), dev_max=r.get("max_dev", 0.0),
                 finite=r.get("finite", True))
    return s


def cadence(N, crossings, samples):
    return SeriesSpec("hambump", N, crossings, samples, 0.0, 0.0).cadence()


# (cahd_c, recorded H_exp, half of the last printed digit)
CAHD_GOLDEN = [(0.0, 0.001, 0.0005), (0.03, -2.14, 0.005), (0.06, -3.05, 0.005)]


def oracle_run(U0, N, crossings, samples, ko_sigma, cahd_c=None, order=6, sync_before=False):
    g = Grid.from_domain(N, order=order)
    p = PhysicsParams() if cahd_c is None else PhysicsParams(cahd_c=cahd_c)
    dt, total, chunk = cadence(N, crossings, samples)
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
        return as_series(rows2)
    return as_series(rows)


def test_cadence_matches_reference_suite():
    assert cadence(24, 2.0, 8) == (1.0 / 96, 192, 24)
    assert cadence(12, 2.0, 8) == (1.0 / 48, 96, 12)
    dt, total, chunk = cadence(64, 20.0, 40)
    assert total == 5120 and chunk == 128


@pytest.mark.parametrize("cahd_c,golden,tol", CAHD_GOLDEN)
def test_oracle_cahd_rates_match_reference_calibration(cahd_c, golden, tol):
    g = Grid.from_domain(12, order=6)
    U0 = OL.hambump(g.shape, g.ng, g.dx, 0.005)
    rows = oracle_run(U0, 12, 2.0, 8, 0.02, cahd_c)
    assert len(rows) == 9 and rows.all_finite
    rate = secular_rates(rows)["H_efold"]
    print(f"cahd_c {cahd_c}: H e-folds {rate:+.4f} (recorded {golden:+.3f})")
    assert abs(rate - golden) <= tol


def test_synced_ghost_sampling_fails_the_pin():
    g = Grid.from_domain(12, order=6)
    U0 = OL.hambump(g.shape, g.ng, g.dx, 0.005)
    rate = secular_rates(oracle_run(U0, 12, 2.0, 8, 0.02, 0.06, sync_before=True))["H_efold"]
    assert abs(rate - (-3.05)) > 0.5


def test_oracle_robust_noise_bounded_and_h_damped():
    """test_bssn_constraint_growth.py:35-44 and test_bssn_longrun.py:86-98 on the oracle."""
    g = Grid.from_domain(12, order=6)
    U0 = idata.robust_stability(g, amp=1e-8, seed=1).numpy()
    rows = oracle_run(U0, 12, 2.0, 8, 0.1)
    r = secular_rates(rows)
    assert rows.all_finite
    assert r["H_efold"] < 0.2 and rows.H[-1] <= rows.H[0]
    assert rows.dev_max.max() < 1e-3
    assert rows.alpha_max.max() < 2.0


def test_oracle_gauge_wave_momentum_drift_subexponential():
    """test_bssn_constraint_growth.py:47-60 and test_bssn_longrun.py:74-83 on the oracle."""
    g = Grid.from_domain(12, order=6)
    U0 = idata.gauge_wave(g, amplitude=0.01, wavelength=1.0).numpy()
    rows = oracle_run(U0, 12, 2.0, 8, 0.1)
    r = secular_rates(rows)
    assert rows.all_finite
    assert r["M_slope"] > 0.0 and rows.M[-1] < 0.5
    assert r["M_line_vs_exp"] < 3.0
    assert rows.alpha_max.max() < 2.0
    assert rows.H.max() < 1.0 and rows.M.max() < 2.0


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


def test_secular_rates_closed_forms():
    t = np.linspace(0.0, 2.0, 9)
    lin = as_series([dict(ncross=x, H=1.0 + 0.5 * x, M=2.0 + 3.0 * x) for x in t])
    r = secular_rates(lin)
    assert r["H_slope"] == pytest.approx(0.5, rel=1e-12)
    assert r["M_slope"] == pytest.approx(3.0, rel=1e-12)
    assert r["M_line_vs_exp"] < 1e-9                          # a line fits a line exactly
    ex = as_series([dict(ncross=x, H=1e-3 * math.exp(-2.5 * x), M=1e-4 * math.exp(0.7 * x))
                    for x in t])
    r = secular_rates(ex)
    assert r["H_efold"] == pytest.approx(-2.5, rel=1e-12)
    assert r["M_efold"] == pytest.approx(0.7, rel=1e-12)
    assert r["M_line_vs_exp"] > 1e6                           # exponential fits exactly
    short = as_series([dict(ncross=x, H=1.0, M=1.0) for x in t[:2]])
    with pytest.raises(ValueError):
        secular_rates(short)


def test_series_columns_and_save(tmp_path):
    s = as_series([dict(ncross=0.5 * i, steps=12 * i, H=1e-3 / (i + 1), M=2e-4, finite=i < 3)
                   for i in range(4)])
    assert len(s) == 4 and s.step.tolist() == [0, 12, 24, 36] and not s.all_finite
    s.save(str(tmp_path / "s.txt"))
    back = np.loadtxt(tmp_path / "s.txt")
    assert back.shape == (4, 7) and np.array_equal(back[:, 2], s.H)


# --- gate logic on synthetic series (the cases of 3D/tests/unit/test_apples_gate.py) ---------

def _series(H, M=1e-6, a=1.0, dev=1e-7, finite=True, n=4):
    return as_series([dict(ncross=0.25 * i, H=H, M=M, max_alpha=a, max_dev=dev, finite=finite)
                      for i in range(n)])


def _recorder(ref, cand, seen=None):
    def rec(spec):
        if seen is not None:
            seen.append((spec.data, spec.scheme, spec.n, spec.order))
        return ref if spec.scheme == "twopass" else cand
    return rec


def test_deviation_is_floored():
    z = _series(0.0, M=0.0, a=0.0, dev=0.0)
    assert max(tg.deviation(z, z, 1e-12).values()) == 0.0
    a, b = _series(1.0 + 1e-9), _series(1.0)
    assert tg.deviation(a, b, 1e-12)["H"] == pytest.approx(1e-9, rel=1e-3)


def test_identical_series_pass():
    s = _series(7.1e-6)
    assert tg.run_gate("fused", suite=tg.SUITE[:1], recorder=_recorder(s, s)).passed


def test_roundoff_drift_within_tol_passes():
    rec = _recorder(_series(7.1e-6), _series(7.1e-6 * (1 + 1e-9)))
    assert tg.run_gate("fused", tol=1e-6, suite=tg.SUITE[:1], recorder=rec).passed


def test_drift_beyond_tol_fails():
    rec = _recorder(_series(7.1e-6), _series(7.1e-6 * (1 + 1e-3)))
    assert not tg.run_gate("fused", tol=1e-6, suite=tg.SUITE[:1], recorder=rec).passed


def test_blowup_point_mismatch_fails():
    rep = tg.run_gate("fused", suite=tg.SUITE[:1],
                      recorder=_recorder(_series(7.1e-6, n=8), _series(7.1e-6, n=4)))
    assert not rep.passed and rep.cases[0].samples == (8, 4)
    assert "8 vs 4" in rep.table()


def test_finiteness_disagreement_fails():
    ref = _series(7.1e-6, n=4)
    cand = _series(7.1e-6, n=4)
    cand.finite[-1] = False
    assert not tg.run_gate("fused", suite=tg.SUITE[:1], recorder=_recorder(ref, cand)).passed


def test_gate_runs_the_whole_suite_and_reports():
    seen = []
    s = _series(1e-5)
    rep = tg.run_gate("int8", order=4, recorder=_recorder(s, s, seen))
    assert rep.passed and [c.case.data for c in rep.cases] == ["robust", "gauge", "hambump"]
    assert seen == [(c.data, sch, 24, 4) for c in tg.SUITE for sch in ("twopass", "int8")]
    assert "overall: pass" in rep.table()
