"""CPU: pin the oracle (oracle/bssn_oracle.c) before it is trusted as the checker.

Pins, each chosen so that a plausible mistake fails one of them:
  * the algebra and constraints against the reference's compiled Dendro CSE on the reference
    oracle's own seeded inputs (3D/tests/integration/test_bssn_oracle.py:24-31, tol 1e-12),
  * a padded grid through the reference CSE on the oracle's own FD bundle (orders 4 and 6),
  * Minkowski is static (3D/tests/unit/test_bssn_rhs.py:62-65),
  * stencil coefficients (3D/tests/unit/test_derivatives.py:29-54), exactness on polynomials,
    sine accuracy on every axis (:59-84), edge padding against numpy's pad(mode="edge"),
  * mixed = d1_j(d1_i) (3D/src/bssn3d/derivative_bundle.py:41-46),
  * the KO symbol -sin^6(theta/2) (the dissipative sign),
  * periodic sync against the reference's slab copies written out in numpy,
  * enforce invariants det = 1, tr = 0 (3D/tests/integration/test_bssn_apples.py:40-48),
  * MSRK consistency (2D/src/mcs2d/msrk.py:79-88) and RK4 / MSRK 4th-order agreement
    (3D/tests/integration/test_msrk_evolve.py:45-60),
  * constraint convergence on the exact gauge wave.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from conftest import normwise


def test_algebra_matches_reference_cse(golden_points):
    g = golden_points
    out = O.algebra(g["F"], g["D"], g["sc13"])
    ref = g["rhs"]
    rel = np.abs(out - ref) / np.maximum(np.maximum(np.abs(out), np.abs(ref)), 1e-12)
    assert rel.max() < 1e-12, f"max rel {rel.max():.3e}"
    # the seeds 0..4 of the reference's own bit-compare are in the set
    assert list(g["seeds"][:5]) == [0, 1, 2, 3, 4]


def test_golden_inputs_follow_reference_recipe(golden_points):
    """The committed inputs are the reference oracle's random_inputs(seed) draws."""
    g = golden_points
    for seed in (0, 3, 17):
        F, D = O.random_inputs(seed)
        np.testing.assert_array_equal(g["F"][:, seed], F)
        np.testing.assert_array_equal(g["D"][:, seed], D)


def test_constraints_match_reference_cse(golden_points):
    g = golden_points
    out = O.constraints_point(g["F"], g["D"])
    ref = g["hm"]
    rel = np.abs(out - ref) / np.maximum(np.maximum(np.abs(out), np.abs(ref)), 1e-12)
    assert rel.max() < 1e-12, f"max rel {rel.max():.3e}"


@pytest.mark.parametrize("order", [4, 6])
def test_grid_rhs_matches_reference_cse(golden_grid, order):
    g = golden_grid
    out = O.rhs(g["F"], g["S16"], order)
    err = normwise(out, g[f"rhs{order}"])
    assert err.max() < 1e-13, err


def test_algebra_discriminates():
    """A single-term perturbation of the inputs moves the outputs: the comparison above is
    not vacuous (every derivative slot reaches some output)."""
    F, D = O.random_inputs(11)
    base = O.algebra(F, D, O.ORACLE_SCALARS13)
    for d in range(138):
        D2 = D.copy()
        D2[d] += 1e-3
        assert np.abs(O.algebra(F, D2, O.ORACLE_SCALARS13) - base).max() > 1e-9, d


def _grid_state(n=10, ng=3, seed=0, amp=0.05):
    rng = np.random.default_rng(seed)
    S = n + 2 * ng
    x = (np.arange(S) - ng) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    F = np.zeros((24, S, S, S))
    for f in range(24):
        k = rng.integers(1, 3, size=3)
        F[f] = amp * np.sin(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z) + rng.uniform(0, 6))
    for f in (0, 1, 3, 6, 8):
        F[f] += 1.0
    return F, 1.0 / n


def test_minkowski_static():
    S = 12
    F = np.zeros((24, S, S, S))
    for f in (0, 1, 3, 6, 8):
        F[f] = 1.0
    S16 = [2.0, 0.06, 0.02, 0.1, 0.6, 20.0, 0.0, 1, 1, 1, 1, 1, 0, 0.1, 0.1, 0.1]
    for order in (4, 6):
        assert np.abs(O.rhs(F, S16, order, ko_sigma=0.3)).max() < 1e-12


def test_stencil_coefficients():
    np.testing.assert_allclose(O.stencil(0, 4), np.array([1, -8, 0, 8, -1]) / 12.0, rtol=0, atol=0)
    np.testing.assert_allclose(O.stencil(1, 4), np.array([-1, 16, -30, 16, -1]) / 12.0, rtol=0, atol=0)
    np.testing.assert_allclose(O.stencil(0, 6), np.array([-1, 9, -45, 0, 45, -9, 1]) / 60.0, rtol=0, atol=0)
    np.testing.assert_allclose(O.stencil(1, 6), np.array([2, -27, 270, -490, 270, -27, 2]) / 180.0, rtol=0, atol=0)
    np.testing.assert_allclose(O.stencil(2, 6), np.array([1, -6, 15, -20, 15, -6, 1]) / 64.0, rtol=0, atol=0)
    np.testing.assert_allclose(O.stencil(2, 8), np.array([-1, 8, -28, 56, -70, 56, -28, 8, -1]) / 256.0, rtol=0, atol=0)
    # order 8 (common/src/mcs_common/derivatives.py:36, 42)
    np.testing.assert_allclose(O.stencil(0, 8), np.array([3, -32, 168, -672, 0, 672, -168, 32, -3]) / 840.0, rtol=0, atol=0)
    np.testing.assert_allclose(O.stencil(1, 8), np.array([-9, 128, -1008, 8064, -14350, 8064, -1008, 128, -9]) / 5040.0, rtol=0, atol=0)
    for order in (4, 6, 8):
        c1, c2 = O.stencil(0, order), O.stencil(1, order)
        k = np.arange(order + 1) - order // 2
        # moment conditions: sum c1 k^m = [m == 1], sum c2 k^m = 2 [m == 2], m <= order
        for m in range(order + 1):
            assert abs(np.sum(c1 * k ** m) - (1.0 if m == 1 else 0.0)) < 1e-12
            assert abs(np.sum(c2 * k ** m) - (2.0 if m == 2 else 0.0)) < 1e-11


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("order", [4, 6, 8])
def test_polynomial_exactness(axis, order):
    rng = np.random.default_rng(axis + 10 * order)
    S = 16
    x = np.arange(S) * 0.1
    shape = [1, 1, 1]
    shape[axis] = S
    c = rng.uniform(-1, 1, size=order + 1)
    p = np.polynomial.Polynomial(c)
    u = np.broadcast_to(p(x).reshape(shape), (S, S, S)).copy()
    d1 = O.apply_axis(u, axis, 0, order, 1 / 0.1)
    d2 = O.apply_axis(u, axis, 1, order, 1 / 0.01)
    r = order // 2
    sl = [slice(None)] * 3
    sl[axis] = slice(r, S - r)
    e1 = np.broadcast_to(p.deriv(1)(x).reshape(shape), (S, S, S))
    e2 = np.broadcast_to(p.deriv(2)(x).reshape(shape), (S, S, S))
    assert np.abs(d1[tuple(sl)] - e1[tuple(sl)]).max() < 1e-9
    assert np.abs(d2[tuple(sl)] - e2[tuple(sl)]).max() < 1e-7


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sine_accuracy(axis):
    N = 32
    x = np.linspace(0.0, 2 * np.pi, N, endpoint=False)
    dx = x[1] - x[0]
    shape = [1, 1, 1]
    shape[axis] = N
    f = np.broadcast_to(np.sin(x).reshape(shape), (N, N, N)).copy()
    d1 = O.apply_axis(f, axis, 0, 6, 1 / dx)
    d2 = O.apply_axis(f, axis, 1, 6, 1 / dx ** 2)
    sl = [slice(None)] * 3
    sl[axis] = slice(3, -3)
    cos = np.broadcast_to(np.cos(x).reshape(shape), (N, N, N))
    assert np.abs(d1 - cos)[tuple(sl)].max() < 1e-4
    assert np.abs(d2 + f)[tuple(sl)].max() < 1e-3


@pytest.mark.parametrize("order", [4, 6, 8])
def test_edge_padding_against_numpy(order):
    rng = np.random.default_rng(order)
    u = rng.normal(size=(7, 9, 8))
    c = O.stencil(0, order)
    r = order // 2
    for axis in range(3):
        pad = [(0, 0)] * 3
        pad[axis] = (r, r)
        up = np.pad(u, pad, mode="edge")
        n = u.shape[axis]
        ref = sum(c[k] * 2.0 * np.take(up, np.arange(k, k + n), axis=axis) for k in range(order + 1))
        got = O.apply_axis(u, axis, 0, order, 2.0)
        assert np.abs(got - ref).max() < 1e-12


def test_default_ko_pairing():
    """ko_order 0 pairs FD order 4 with KO 6, 6 with KO 8 and 8 with KO 8 (the reference's widest
    KO stencil, derivatives.py:45-49); an order-10 KO request is refused."""
    rng = np.random.default_rng(8)
    F = 0.01 * rng.normal(size=(24, 11, 12, 13))
    for f in (0, 1, 3, 6, 8):
        F[f] += 1.0
    S16 = [2.0, 0.06, 0.02, 0.1, 0.6, 20.0, 0.4, 1.1, 0.9, 1.2, 0.8, 1.0, 0.5, 0.1, 0.11, 0.09]
    for order, ko in ((4, 6), (6, 8), (8, 8)):
        np.testing.assert_array_equal(O.rhs(F, S16, order, ko_sigma=0.1),
                                      O.rhs(F, S16, order, ko_order=ko, ko_sigma=0.1))
        other = 4 if ko != 4 else 6
        assert not np.array_equal(O.rhs(F, S16, order, ko_sigma=0.1),
                                  O.rhs(F, S16, order, ko_order=other, ko_sigma=0.1))
    with pytest.raises(ValueError):
        O.rhs(F, S16, 8, ko_order=10, ko_sigma=0.1)


def test_mixed_is_nested_first_derivatives():
    F, h = _grid_state()
    h3 = [h, 1.1 * h, 0.9 * h]
    D = O.deriv_bundle(F, h3, 4)
    for f in ("alpha", "gt4", "beta2"):
        u = F[O.IDX[f]]
        for i, j in ((0, 1), (0, 2), (1, 2)):
            nested = O.apply_axis(O.apply_axis(u, i, 0, 4, 1 / h3[i]), j, 0, 4, 1 / h3[j])
            np.testing.assert_array_equal(D[O.deriv_index("d2", f, i, j)], nested)


def test_ko_symbol():
    """KO6 on cos(theta i) is -sigma/h sin^6(theta/2) cos(theta i): dissipative."""
    N = 32
    i = np.arange(N)
    for m in (1, 5, 16):
        th = 2 * np.pi * m / N
        u = np.broadcast_to(np.cos(th * i).reshape(N, 1, 1), (N, 4, 4)).copy()
        k = O.apply_axis(u, 0, 2, 6, 0.3 / 0.5)
        expect = -(0.3 / 0.5) * np.sin(th / 2) ** 6 * u
        assert np.abs(k - expect)[3:-3].max() < 1e-13


def test_periodic_sync_matches_slab_copies():
    rng = np.random.default_rng(5)
    F = rng.normal(size=(24, 9, 11, 10))
    ng = 2
    ref = F.copy()
    ref[:, :ng] = ref[:, -2 * ng:-ng]
    ref[:, -ng:] = ref[:, ng:2 * ng]
    ref[:, :, :ng] = ref[:, :, -2 * ng:-ng]
    ref[:, :, -ng:] = ref[:, :, ng:2 * ng]
    ref[:, :, :, :ng] = ref[:, :, :, -2 * ng:-ng]
    ref[:, :, :, -ng:] = ref[:, :, :, ng:2 * ng]
    np.testing.assert_array_equal(O.periodic_sync(F, ng), ref)


def test_enforce_invariants():
    F, _ = _grid_state(amp=0.05)
    E = O.enforce(F)
    g = E[3:9]
    det = g[0] * (g[3] * g[5] - g[4] ** 2) - g[1] * (g[1] * g[5] - g[4] * g[2]) + g[2] * (g[1] * g[4] - g[3] * g[2])
    assert np.abs(det - 1).max() < 1e-12
    a = E[12:18]
    iu = [g[3] * g[5] - g[4] ** 2, g[2] * g[4] - g[1] * g[5], g[1] * g[4] - g[2] * g[3],
          g[0] * g[5] - g[2] ** 2, g[1] * g[2] - g[0] * g[4], g[0] * g[3] - g[1] ** 2]
    tr = iu[0] * a[0] + iu[3] * a[3] + iu[5] * a[5] + 2 * (iu[1] * a[1] + iu[2] * a[2] + iu[4] * a[4])
    assert np.abs(tr).max() < 1e-12
    low = F.copy()
    low[1] = -1.0
    low[0] = 1e-9
    E = O.enforce(low)
    assert E[1].min() == 1e-4 and E[0].min() == 1e-4


def test_msrk_coefficients():
    """rk4_2_2: sum(b) = 1, c2 = a20 + a21 = -99/50, c3 = a30 + a31 + a32 = 101/100."""
    F_ = Fraction
    b = [F_(-191, 882), F_(48241, 59994), F_(193750, 4351347), F_(100000, 271791)]
    assert sum(b) == 1
    assert F_(1309, 15500) + F_(-31999, 15500) == F_(-99, 50)
    assert F_(-241289, 5880000) + F_(22846301, 16170000) + F_(-936169, 2587200) == F_(101, 100)


def _intercept(radius, hi=4.0, tol=1e-7):
    """Largest B with spectral radius <= 1 at z = iB, by bisection along the imaginary axis."""
    lo = 0.0
    while radius(1j * hi) <= 1.0 + 1e-9:
        hi *= 2.0
    while hi - lo > tol:
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if radius(1j * mid) <= 1.0 + 1e-9 else (lo, mid)
    return 0.5 * (lo + hi)


def test_msrk_stability_intercept_matches_paper():
    """The oracle's rk4_2_2 coefficients, applied to y' = lambda y (z = lambda dt), give the
    two-step recurrence y_{n+1} = P(z) y_n + Q(z) y_{n-1}; the companion matrix [[P, Q], [1, 0]]
    has spectral radius <= 1 on the imaginary axis up to 2.46201, the published intercept the
    reference checks (2D/tests/integration/test_msrk.py:18-28, msrk.py:112-157). Classic RK4's
    sqrt(8) is the control."""
    a20, a21, a30, a31, a32, b0, b1, b2, b3 = O.msrk_coeffs()

    def step(z, yn, ynm1):
        g0, g1 = z * ynm1, z * yn
        g2 = z * (yn + a20 * g0 + a21 * g1)
        g3 = z * (yn + a30 * g0 + a31 * g1 + a32 * g2)
        return yn + b0 * g0 + b1 * g1 + b2 * g2 + b3 * g3

    def radius(z):
        C = np.array([[step(z, 1.0, 0.0), step(z, 0.0, 1.0)], [1.0, 0.0]], dtype=complex)
        return float(np.max(np.abs(np.linalg.eigvals(C))))
    rk4 = lambda z: abs(1 + z + z * z / 2 + z ** 3 / 6 + z ** 4 / 24)
    assert abs(_intercept(rk4) - np.sqrt(8.0)) < 1e-5
    assert abs(_intercept(radius) - 2.46201) < 1e-4
    assert abs(sum((b0, b1, b2, b3)) - 1.0) < 1e-15


def _gauge_wave_np(n, amp=0.01, ng=3):
    h = 1.0 / n
    x = -0.5 + (np.arange(n + 2 * ng) - ng) * h
    X = np.broadcast_to(x.reshape(-1, 1, 1), (n + 2 * ng,) * 3)
    k = 2 * np.pi
    H = 1 - amp * np.sin(k * X)
    dH = -amp * k * np.cos(k * X)
    sq = np.sqrt(H)
    Kxx = dH / (2 * sq)
    Kt = Kxx / H
    chi = H ** (-1 / 3)
    F = np.zeros((24,) + X.shape)
    F[0], F[1], F[2] = sq, chi, Kt
    F[3], F[6], F[8] = H ** (2 / 3), chi, chi
    F[12] = chi * (Kxx - H * Kt / 3)
    F[15] = F[17] = chi * (-Kt / 3)
    F[18] = (2 / 3) * H ** (-5 / 3) * dH
    return F, h


def test_rk4_msrk_fourth_order_agreement():
    F, h = _gauge_wave_np(12)
    S16 = [2.0, 0.06, 0.1 * h, h, 0.6, 20.0, 0.0, 1, 1, 1, 1, 1, 0, h, h, h]
    kw = dict(order=4, ko_sigma=0.4, ng=3)

    def gap(cfl, n):
        dt = cfl * h
        a = O.evolve(F, S16, dt, n, integrator="rk4", **kw)
        b = O.evolve(F, S16, dt, n, integrator="rk4_2_2", **kw)
        return np.sqrt(np.mean((a - b) ** 2))

    # the reference's operating point: CFL 0.1 -> 0.05 to the same final time, KO 0.4
    g1, g2 = gap(0.1, 20), gap(0.05, 40)
    assert g1 > 0 and g1 / g2 > 8.0, (g1, g2)


def test_constraint_convergence_gauge_wave():
    errs = []
    for n in (16, 32):
        F, h = _gauge_wave_np(n)
        H, M = O.constraint_l2(F, [h, h, h], 4, 3)
        errs.append((H, M))
    assert np.log2(errs[0][0] / errs[1][0]) > 3.5
    assert np.log2(errs[0][1] / errs[1][1]) > 3.5


def test_sampled_bundle_and_long_double_step():
    """The sampled-derivative wrapper equals the full bundle at the same points bit for bit;
    the long-double builds of the bundle and of the RK4 step agree with the FP64 ones to
    round-off (they are the accuracy references of the full-size GPU checks)."""
    rng = np.random.default_rng(4)
    F = 0.05 * rng.standard_normal((24, 11, 10, 12))
    F[[0, 1, 3, 6, 8]] += 1.0          # alpha, chi and the diagonal of gt near 1
    h3 = [0.1, 0.11, 0.12]
    D = O.deriv_bundle(F, h3, 4).reshape(138, -1)
    idx = rng.integers(0, 11 * 10 * 12, size=40)
    np.testing.assert_array_equal(O.deriv_points(F, h3, idx, 4), D[:, idx])
    Dl = O.deriv_points(F, h3, idx, 4, long_double=True)
    assert np.max(np.abs(Dl - D[:, idx])) <= 1e-12 * np.max(np.abs(D))
    S16 = [2.0, 0.06, 0.02, 0.1, 0.6, 20.0, 0.0, 1, 1, 1, 1, 1, 0, 0.1, 0.11, 0.12]
    a = O.rk4_step(F, S16, 0.0, 0.02, order=4, ko_sigma=0.1, ng=3)
    b = O.rk4_step_ld(F, S16, 0.0, 0.02, order=4, ko_sigma=0.1, ng=3)
    inc = np.abs(a - F).reshape(24, -1).max(axis=1)
    err = np.abs(a - b.astype(np.float64)).reshape(24, -1).max(axis=1)
    assert np.all(err <= 1e-12 * np.maximum(inc, 1e-12)), err / inc


def _uniform_lapse_state(alpha0: float, shape=(9, 9, 9)) -> np.ndarray:
    """A spatially uniform state: chi = 1, gt = delta, alpha = alpha0, everything else 0. All
    derivatives vanish, K, At, beta, B and Gt stay 0, and the lapse obeys the slice-locking ODE
    d(alpha - 1)/dt = -h_ssl (alpha - 1) exp(-t^2 / (2 sig^2)) alone."""
    F = np.zeros((24,) + shape)
    F[0] = alpha0
    F[1] = 1.0
    F[[3, 6, 8]] = 1.0
    return F


def ssl_exact(alpha0: float, t0: float, t1: float, h: float, sig: float) -> float:
    """alpha(t1) of the uniform-state lapse ODE from alpha(t0) = alpha0, in closed form."""
    from math import erf, exp, pi, sqrt
    integral = sig * sqrt(pi / 2.0) * (erf(t1 / (sig * sqrt(2.0))) - erf(t0 / (sig * sqrt(2.0))))
    return 1.0 + (alpha0 - 1.0) * exp(-h * integral)


def test_rk4_stage_times_pinned_by_exact_lapse_solution():
    """The oracle's RK4 stage times (t, t + dt/2, t + dt/2, t + dt) enter through the SSL
    factor. With sig_ssl = 0.4 the factor changes by ~20 % over a step, so a stage evaluated at
    a wrong time misses the closed-form solution by ~1e-5; RK4 itself misses it by O(dt^5)."""
    h, sig, t0, dt, a0 = 0.6, 0.4, 0.35, 0.02, 1.3
    F = _uniform_lapse_state(a0)
    S16 = [2.0, 0.06, dt, 0.1, h, sig, t0, 1, 1, 1, 1, 1, 0, 0.1, 0.1, 0.1]
    U = O.rk4_step(F, S16, t0, dt, order=4, ko_sigma=0.1, ng=3)
    exact = ssl_exact(a0, t0, t0 + dt, h, sig)
    assert np.all(U[0] == U[0, 4, 4, 4])
    err = abs(U[0, 4, 4, 4] - exact) / (a0 - 1.0)
    assert err < 1e-9, err
    # the same step with every stage at the step's start time misses by far more
    wrong = O.stage_rhs(F, S16, t0, order=4, ko_sigma=0.1, ng=3)[0, 4, 4, 4]
    assert abs((a0 + dt * wrong) - exact) / (a0 - 1.0) > 1e-5
    # the other fields stay where they were (to round-off)
    assert np.max(np.abs(U[2])) < 1e-13 and np.max(np.abs(U[9:12])) < 1e-13
    assert np.max(np.abs(U[1] - 1.0)) < 1e-13
