"""CPU: pins of the INT8-split restatement (oracle/int8_split.py) before it is trusted as the
bit-exact checker of the HIP INT8 bundle.

  * digits: int8 range (|q| <= 64), exact reconstruction of the truncated value, remainder
    bound 2^(-7 k) relative to the tile's range, zero and constant tiles;
  * banded matrices: each row is the stencil at its output, so B q equals a direct integer
    convolution (np.convolve) of the digit planes;
  * exactness on polynomials: integer-valued quadratics are split without remainder, so the
    INT8 bundle equals the closed-form derivatives exactly;
  * agreement with the FP64 oracle bundle (itself pinned by test_oracle.py) at 7 digits, and
    the error falling ~2^7 per added digit.
"""
from __future__ import annotations

import numpy as np
import pytest

import int8_split as I8
import oracle as O
from conftest import normwise


def test_digits_range_and_reconstruction():
    rng = np.random.default_rng(0)
    for scale in (1e-9, 1.0, 3e5):
        box = scale * (0.7 + rng.standard_normal((32, 32)))
        for k in (1, 4, 7):
            q, e = I8.split(box, k)
            assert np.abs(q).max() <= 64
            c = I8.offset(box)
            rec = sum(np.ldexp(q[s].astype(float), -7 * s) for s in range(k))
            err = np.abs(np.ldexp(box - c, e) - rec).max()
            assert err <= 0.5 * 2.0 ** (-7 * (k - 1)) + 1e-12
    q, e = I8.split(np.zeros((32, 32)), 7)
    assert e == 0 and not q.any()
    q, e = I8.split(np.full((32, 32), -2.5), 7)
    assert not q.any()


def test_band_matrices_are_the_stencils():
    rng = np.random.default_rng(1)
    q = rng.integers(-64, 65, size=(32, 32))
    for B, c in ((I8.B1, I8.C1), (I8.B2, I8.C2)):
        P = B @ q
        for col in range(32):
            conv = np.convolve(q[:, col], c[::-1], mode="valid")   # sum_t c[t] q[r + t]
            assert np.array_equal(P[:, col], conv)
    assert I8.B1.sum(axis=1).tolist() == [0] * 28 and I8.B2.sum(axis=1).tolist() == [0] * 28


def test_polynomial_data_is_exact():
    """u = integer quadratic: digits carry it exactly, so derivatives are the closed form
    (up to the final scale multiply, which is exact for h = 1/4, 1/8)."""
    S = (9, 11, 13)
    x, y, z = np.meshgrid(*[np.arange(n, dtype=float) for n in S], indexing="ij")
    F = np.zeros((24,) + S)
    F[0] = 3 * x * x - 2 * x * y + 5 * z * z + 7 * y - 1          # alpha
    F[2] = 4 * x + 9 * y - 6 * z                                 # K (first derivatives only)
    h = (0.25, 0.125, 0.25)
    D = I8.bundle(F, h, 7)
    inner = (slice(2, -2),) * 3
    hx, hy, hz = h
    assert np.all(D[0][inner] == ((6 * x - 2 * y) / hx)[inner])
    assert np.all(D[1][inner] == ((-2 * x + 7) / hy)[inner])
    assert np.all(D[2][inner] == ((10 * z) / hz)[inner])
    assert np.all(D[6][inner] == 4 / hx) and np.all(D[7][inner] == 9 / hy)
    assert np.all(D[72 + 0][inner] == 6 / hx ** 2)               # d_xx alpha
    assert np.all(D[72 + 1][inner] == -2 / (hx * hy))            # d_xy alpha
    assert np.all(D[72 + 5][inner] == 10 / hz ** 2)              # d_zz alpha
    assert np.all(D[72 + 2][inner] == 0) and np.all(D[72 + 4][inner] == 0)


def test_against_fp64_oracle_and_digit_convergence():
    from bssn3d_hip import initial_data as idata
    grid, st, _ = idata.config_single_puncture(seed=0, N=16)
    F = st.numpy()
    h = (grid.dx,) * 3
    D = O.deriv_bundle(F, h, order=4)
    errs = {k: normwise(I8.bundle(F, h, k), D).max() for k in (2, 3, 4, 7)}
    assert errs[7] <= 5e-11
    assert errs[2] > 50 * errs[3] > 50 * 50 * errs[4]
