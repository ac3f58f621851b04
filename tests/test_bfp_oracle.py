"""CPU: the block-floating-point storage restated in oracle/bfp.py, pinned.

  * quantize / block_exponent / pack48 on closed forms: representable values exact, error at
    most half a unit of the last mantissa bit, saturation, round half to even, all-zero arrays,
    48-bit limbs of negative mantissas (common/src/mcs_common/bfp_compress.py:37-90,
    bfp48.py:7-43);
  * the reference's BSSN storage-compression gate (3D/tests/integration/
    test_bssn_compression_accuracy.py) on polarized Gowdy data, N=16, FD order 6: fp64 storage
    exact, BFP48 below 1e-13, errors increasing 48 < 32 < 24, BFP32 constraint perturbations
    under its bars and BFP24 >100x BFP32 -- and the values its docstrings record (BFP32 RHS
    ~7e-10, constraint perturbation ~2.6e-9 / 7.4e-8), which the restatement reproduces;
  * the reference emulator's model (mixed derivatives quantized twice) and the device
    form (mixed taken from the state, quantized once) agree on fp64 storage bit for bit and
    differ only in the mixed arrays, at the quantization level;
  * the Gowdy constraints converge at the FD order (an exact solution).
"""
from __future__ import annotations

import numpy as np
import pytest

import bfp as B
import oracle as O

from bssn3d_hip import initial_data as idata
from bssn3d_hip.grid import Grid
from bssn3d_hip.rhs import bfp_exponents
from bssn3d_hip.state import PhysicsParams, scalars


def test_quantize_closed_forms():
    a = np.array([0.5, -0.25, 0.375, 0.0, 1.0 - 2.0 ** -20])
    e = B.block_exponent(a, 24)
    assert e == 23                    # (1 - 2^-20) 2^23 = 2^23 - 8 <= 2^23 - 1 < (1 - 2^-20) 2^24
    assert np.array_equal(B.quantize(a, 24), a)            # all representable in 23 bits
    rng = np.random.default_rng(0)
    x = rng.standard_normal(10000) * 3.0
    for k in (24, 32, 48):
        e = B.block_exponent(x, k)
        m = B.mantissas(x, k, e)
        assert np.max(np.abs(m)) <= 2.0 ** (k - 1) - 1 and np.max(np.abs(m)) > 2.0 ** (k - 2)
        assert np.max(np.abs(B.quantize(x, k) - x)) <= 2.0 ** (-e - 1)
    # saturation with a given exponent, round half to even
    assert np.array_equal(B.mantissas(np.array([10.0, -10.0]), 4, 0), [7.0, -8.0])
    # non-finite values with a given exponent: +-inf saturate, NaN stores as -2^(k-1) (an integer
    # mantissa cannot hold it); the float round trip keeps NaN as the reference's emulation does
    v = np.array([np.nan, -np.inf, np.inf, -1e300, 1e300])
    for k in (24, 32, 48):
        lim = 2.0 ** (k - 1)
        assert np.array_equal(B.mantissas(v, k, 3), [-lim, -lim, lim - 1, -lim, lim - 1])
    q = B.quantize(np.array([np.nan, 0.5]), 24, 3)
    assert np.isnan(q[0]) and q[1] == 0.5
    assert np.array_equal(B.mantissas(np.array([0.5, 1.5, 2.5, -0.5]), 8, 0), [0, 2, 2, -0])
    assert B.block_exponent(np.zeros(5), 32) == 0 and np.all(B.quantize(np.zeros(5), 32) == 0)
    # the product's exponent helper is the same formula
    amax = np.abs(rng.standard_normal(138)) * 10.0 ** rng.integers(-8, 8, 138)
    amax[5] = 0.0
    want = [B.block_exponent(np.array([v]), 32) for v in amax]
    assert np.array_equal(bfp_exponents(amax, 32), want)
    with pytest.raises(ValueError):
        bfp_exponents(np.array([1.0, np.nan]), 32)


def test_pack48_limbs():
    m = np.array([0, 1, -1, 2 ** 47 - 1, -2 ** 47, 123456789012, -98765432109])
    q = B.pack48(m.astype(np.float64))
    assert q.dtype == np.uint16 and q.shape == (7, 3)
    assert np.array_equal(q[2], [0xFFFF, 0xFFFF, 0xFFFF])
    assert np.array_equal(B.unpack48(q), m)


@pytest.fixture(scope="module")
def gowdy16():
    g = Grid.from_domain(16)
    F = idata.gowdy(g, 1.0).numpy()
    h3 = (g.dx, g.dy, g.dz)
    S16 = scalars(PhysicsParams(), g.dx, g.dy, g.dz, 0.25 * g.dx, 0.0)
    D = O.deriv_bundle(F, h3, 6)
    return g, F, h3, S16, D


def _rhs_err(gowdy16, k, model):
    g, F, h3, S16, D = gowdy16
    I = (slice(None),) + (slice(g.ng, -g.ng),) * 3
    a = B.rhs(F, S16, D)[I]
    Dq = B.bundle_ref_model(F, h3, 6, k) if model == "ref" else B.bundle(F, h3, 6, k)[0]
    b = B.rhs(F, S16, Dq)[I]
    return float(np.sqrt(np.mean((b - a) ** 2)) / np.sqrt(np.mean(a ** 2)))


def test_reference_gate_rhs(gowdy16):
    assert _rhs_err(gowdy16, 53, "ref") == 0.0
    e48, e32, e24 = (_rhs_err(gowdy16, k, "ref") for k in (48, 32, 24))
    assert e48 < 1e-13 and e48 < e32 < e24
    assert 6.5e-10 < e32 < 8.0e-10                         # recorded: ~7e-10
    assert _rhs_err(gowdy16, 32, "device") == pytest.approx(e32, rel=1e-6)


def test_reference_gate_constraints(gowdy16):
    g, F, h3, S16, D = gowdy16
    H0, M0 = B.constraint_l2(F, D, g.ng)
    assert (H0, M0) == pytest.approx(O.constraint_l2(F, h3, 6, g.ng), rel=1e-12)

    def pert(k):
        H, M = B.constraint_l2(F, B.bundle_ref_model(F, h3, 6, k), g.ng)
        return abs(H - H0) / H0, abs(M - M0) / M0
    dH32, dM32 = pert(32)
    dH24, dM24 = pert(24)
    assert dH32 < 1e-7 and dM32 < 1e-6 and dH24 > 100 * dH32
    assert dH32 == pytest.approx(2.6e-9, rel=0.05) and dM32 == pytest.approx(7.4e-8, rel=0.05)


def test_models_differ_only_in_mixed_arrays():
    grid, st, _ = idata.config_single_puncture(seed=2, N=12)
    F = st.numpy()
    h3 = (grid.dx, grid.dy, grid.dz)
    assert np.array_equal(B.bundle_ref_model(F, h3, 4, 53), O.deriv_bundle(F, h3, 4))
    R = B.bundle_ref_model(F, h3, 4, 32)
    Dd = B.bundle(F, h3, 4, 32)[0]
    mixed = np.array([72 + 6 * s + p for s in range(11) for p in (1, 2, 4)])
    other = np.setdiff1d(np.arange(138), mixed)
    assert np.array_equal(R[other], Dd[other])
    D = O.deriv_bundle(F, h3, 4)
    for i in mixed:
        scale = np.max(np.abs(D[i]))
        assert np.max(np.abs(R[i] - Dd[i])) <= 64 * 2.0 ** -31 * max(scale, 1e-300) + 1e-12 * scale


def test_gowdy_constraints_converge():
    Hs = []
    for N in (16, 32):
        g = Grid.from_domain(N)
        F = idata.gowdy(g, 1.0).numpy()
        Hs.append(O.constraint_l2(F, (g.dx, g.dy, g.dz), 6, g.ng))
    assert Hs[0][0] / Hs[1][0] > 2 ** 6 * 0.6 and Hs[0][1] / Hs[1][1] > 2 ** 6 * 0.6
