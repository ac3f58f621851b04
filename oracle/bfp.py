"""Block-floating-point storage of the derivative bundle, restated in numpy.

TEST INFRASTRUCTURE ONLY: the device BFP bundle and algebra are checked against this. Nothing
in the product imports it.

  * block_exponent / quantize: common/src/mcs_common/bfp_compress.py:37-90 (one power-of-two
    exponent per array, floor(log2((2^(k-1) - 1) / max|a|)), 0 for an all-zero array;
    m = clip(round-half-even(a 2^e), -2^(k-1), 2^(k-1) - 1); value m 2^-e);
  * pack48 / unpack48: common/src/mcs_common/bfp48.py:7-43 (a 48-bit mantissa as three
    16-bit two's-complement limbs, low first);
  * bundle: the oracle's 138-array bundle (oracle.deriv_bundle) with every array stored at k
    bits -- the device form, whose mixed derivatives are taken from the state directly;
  * bundle_ref_model: the reference's emulator (QuantizingDeriv wrapping every derivative
    output, bfp_compress.py:49-62, fed by derivative_bundle.py:31-48): first and diagonal
    second derivatives quantized, mixed = Q(d1_j(Q(d1_i f))) -- the inner first derivative is
    stored too.
"""
from __future__ import annotations

import numpy as np

import oracle as O


def block_exponent(a: np.ndarray, k: int) -> int:
    mx = float(np.max(np.abs(a)))
    if mx == 0.0:
        return 0
    return int(np.floor(np.log2((2.0 ** (k - 1) - 1.0) / mx)))


def mantissas(a: np.ndarray, k: int, e: int) -> np.ndarray:
    """The stored integer mantissas. Parity unpinned for NaN: the reference's emulation keeps
    floats (a NaN stays NaN, bfp_compress.py:88-90); an integer mantissa cannot hold one, so the
    device stores it as the negative saturation value -2^(k-1), like -inf."""
    lim = 2.0 ** (k - 1)
    m = np.clip(np.rint(np.ldexp(a, e)), -lim, lim - 1.0)
    return np.where(np.isnan(m), -lim, m)


def quantize(a: np.ndarray, k: int, e: int | None = None) -> np.ndarray:
    """The reference's float round trip (bfp_compress.py:70-90): NaN stays NaN."""
    e = block_exponent(a, k) if e is None else e
    a = np.asarray(a, dtype=np.float64)
    return np.where(np.isnan(a), np.nan, np.ldexp(mantissas(a, k, e), -e))


def pack48(m: np.ndarray) -> np.ndarray:
    mi = m.astype(np.int64)
    return np.stack([(mi & 0xFFFF), (mi >> 16) & 0xFFFF, (mi >> 32) & 0xFFFF], -1).astype(np.uint16)


def unpack48(q: np.ndarray) -> np.ndarray:
    c = q[..., 0].astype(np.int64) | (q[..., 1].astype(np.int64) << 16) | \
        (q[..., 2].astype(np.int64) << 32)
    return np.where(c >= (1 << 47), c - (1 << 48), c)


def bundle(F: np.ndarray, h3, order: int, k: int):
    """(Dq, exps, m): the bundle stored at k bits per array, its exponents and mantissas."""
    D = O.deriv_bundle(F, h3, order)
    exps = np.array([block_exponent(D[i], k) for i in range(138)], dtype=np.int64)
    m = np.stack([mantissas(D[i], k, int(exps[i])) for i in range(138)])
    Dq = np.stack([np.ldexp(m[i], -int(exps[i])) for i in range(138)])
    return Dq, exps, m


# second-derivative slots: alpha, chi, gt0..5, beta0..2; pairs 00, 01, 02, 11, 12, 22
_S2 = [0, 1] + list(range(3, 9)) + [9, 10, 11]
_PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def bundle_ref_model(F: np.ndarray, h3, order: int, k: int) -> np.ndarray:
    """The reference emulator's stored bundle (mixed derivatives quantized twice)."""
    D = np.zeros((138,) + F.shape[1:])
    q = (lambda a: a) if k >= 53 else (lambda a: quantize(a, k))
    first = {}
    for f in range(24):
        for a in range(3):
            first[(f, a)] = q(O.apply_axis(F[f], a, 0, order, 1.0 / h3[a]))
            D[3 * f + a] = first[(f, a)]
    for s, f in enumerate(_S2):
        for p, (i, j) in enumerate(_PAIRS):
            if i == j:
                D[72 + 6 * s + p] = q(O.apply_axis(F[f], i, 1, order, 1.0 / (h3[i] * h3[i])))
            else:
                D[72 + 6 * s + p] = q(O.apply_axis(first[(f, i)], j, 0, order, 1.0 / h3[j]))
    return D


def rhs(F: np.ndarray, S16, Dq: np.ndarray) -> np.ndarray:
    """The algebra on a stored bundle (no Kreiss-Oliger), every padded point."""
    return O.algebra(F, Dq, O.s16_to_13(S16)).reshape(F.shape)


def constraint_l2(F: np.ndarray, Dq: np.ndarray, ng: int):
    HM = O.constraints_point(F, Dq).reshape((4,) + F.shape[1:])
    I = (slice(None),) + (slice(ng, -ng),) * 3
    H = HM[0][I[1:]]
    M = HM[1:][I]
    return float(np.sqrt(np.mean(H ** 2))), float(np.sqrt(np.mean(np.sum(M ** 2, axis=0))))
