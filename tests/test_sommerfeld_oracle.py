"""CPU: pins of the radiative-boundary restatement (oracle/sommerfeld.py) before it is trusted
as the checker of the GPU boundary: exactness on linear data where no tangential stencil is
clamped, the face priority at edges and corners (z over y over x, the reference MCS solver's
write order), untouched interiors, and first-order convergence of the condition on an
outgoing radial profile u = G(r)."""
from __future__ import annotations

import numpy as np

import sommerfeld as SF

NG = 3


def grid(S, h, xmin):
    ax = [xmin[a] + (np.arange(S[a]) - NG) * h[a] for a in range(3)]
    return np.meshgrid(*ax, indexing="ij")


def faces(S):
    I = np.meshgrid(*[np.arange(n) for n in S], indexing="ij")
    g = [(I[a] < NG) | (I[a] >= S[a] - NG) for a in range(3)]
    gz = g[2]
    gy = g[1] & ~gz
    gx = g[0] & ~gz & ~gy
    return I, gx, gy, gz


def test_linear_data_exact_and_interior_untouched():
    S, h, xmin = (11, 12, 13), (0.1, 0.2, 0.15), (-0.4, -0.9, -0.6)
    X, Y, Z = grid(S, h, xmin)
    u = 1.5 + 0.3 * X - 0.7 * Y + 0.2 * Z
    d = SF.apply(np.full(S, 7.0), u, NG, xmin, h, 4)
    I, gx, gy, gz = faces(S)
    inner = [(I[a] >= 2) & (I[a] < S[a] - 2) for a in range(3)]
    ok = (gz & inner[0] & inner[1]) | (gy & inner[0] & inner[2]) | (gx & inner[1] & inner[2])
    want = -(0.3 * X - 0.7 * Y + 0.2 * Z) / np.sqrt(X * X + Y * Y + Z * Z)
    assert ok.sum() > 500
    assert np.abs(d[ok] - want[ok]).max() < 1e-13
    assert np.all(d[~(gx | gy | gz)] == 7.0)


def test_face_priority_at_edges_and_corners():
    S, h, xmin = (10, 10, 10), (0.2, 0.2, 0.2), (-0.9, -0.7, -0.8)
    rng = np.random.default_rng(0)
    u = rng.standard_normal(S)
    d = SF.apply(np.zeros(S), u, NG, xmin, h, 4)
    rb = SF.bc_rhs(u, NG, xmin, h, 4)
    I, gx, gy, gz = faces(S)
    for mask, a in ((gz, 2), (gy, 1), (gx, 0)):
        low = mask & (I[a] < NG)
        high = mask & (I[a] >= S[a] - NG)
        assert np.array_equal(d[low], rb[(a, 0)][low])
        assert np.array_equal(d[high], rb[(a, 1)][high])


def test_outgoing_profile_first_order():
    """u = G(r) = exp(-(r - 2)^2): the condition's exact value is -G'(r); the one-sided normal
    difference makes the error O(h)."""
    errs = []
    for n in (16, 32):
        h = 2.0 / n
        S = (n + 2 * NG,) * 3
        xmin = (1.0, -1.0, -1.0)        # the box x in [1, 3): r stays away from 0
        X, Y, Z = grid(S, (h,) * 3, xmin)
        R = np.sqrt(X * X + Y * Y + Z * Z)
        u = np.exp(-(R - 2.0) ** 2)
        d = SF.apply(np.zeros(S), u, NG, xmin, (h,) * 3, 4)
        want = 2.0 * (R - 2.0) * u       # -G'(r)
        I, gx, gy, gz = faces(S)
        inner = [(I[a] >= 2) & (I[a] < S[a] - 2) for a in range(3)]
        ok = gx & inner[1] & inner[2]
        errs.append(np.abs(d[ok] - want[ok]).max())
    assert errs[0] / errs[1] > 1.7
