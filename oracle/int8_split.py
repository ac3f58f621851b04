"""INT8-split (Ozaki scheme I) derivative bundle, restated in numpy with exact int64 products.

TEST INFRASTRUCTURE ONLY: tests compare the HIP INT8 bundle with this bit for bit. Nothing in
the product imports it.

The reference's precision form is Ozaki scheme II (residue number system + CRT) and is used only
for its Maxwell-Chern-Simons solver (2D/src/mcs2d/schemes/ozaki.py:257-562,
3D/src/mcs3d/schemes/ozaki.py:162-282); this form keeps two of its ideas -- a power-of-two scale
per block (ozaki.py:316-323) and the stencil as a banded integer matrix applied by an integer
GEMM (ozaki.py:286-306, 352-377) -- and splits values into int8 digits instead of residues.

Algorithm (FD order 4, integer stencils c1 = [1, -8, 0, 8, -1] / 12 h, c2 = [-1, 16, -30, 16, -1]
/ 12 h^2, mixed = c1 (x) c1 / 144 h_i h_j, as in common/src/mcs_common/derivatives.py:33-49):
  * Planes. Derivatives come from 2D slabs of one field: at each x the (y, z) plane gives
    d_y, d_yy, d_z, d_zz, d_yz; at each y the (x, z) plane gives d_x, d_xx, d_xz; at each z the
    (x, y) plane gives d_xy. A plane is cut into tiles of 28 x 28 outputs; a tile reads the
    32 x 32 box around it, coordinates clamped to the grid (the reference's edge padding;
    clamping is separable, so d1_j(d1_i f) of the padded arrays equals the 2D stencil on the
    clamped box).
  * Offset. c = 0.5 max(box) + 0.5 min(box) is subtracted from the box (float64): the stencils
    annihilate constants exactly, so the digits only carry the variation within the tile.
  * Split. With m = max |box - c| = f * 2^E (0.5 <= f < 1), sigma = 2^(6 - E) (sigma = 1 if
    m = 0); r_0 = (box - c) sigma, q_s = rint(r_s), r_{s+1} = (r_s - q_s) * 128 for s < k. Every
    step after the offset is exact in float64 and |q_s| <= 64, so the digits are int8 and
    (box - c) sigma = sum_s q_s 2^(-7 s) + r_k 2^(-7 k).
  * Contract. Per digit plane, the banded integer matrices B1, B2 (28 x 32, row r holding the
    taps at columns r .. r+4) give  B q (rows), q B^T (columns) and B1 q B1^T (mixed), exactly
    in integers (int32 on the device, int64 here).
  * Recombine. acc = 0; acc += P_s * 2^(-7 s) for s = 0 .. k-1 (each term exact, sums rounded in
    float64 in that order); D = (acc * 2^-(6 - E)) * scale, scale = 1/(12 h), 1/(12 h^2),
    1/(144 h_i h_j).
"""
from __future__ import annotations

import numpy as np

TILE = 28
BOX = 32
R = 2
C1 = np.array([1, -8, 0, 8, -1], dtype=np.int64)
C2 = np.array([-1, 16, -30, 16, -1], dtype=np.int64)
S2_FIELDS = (0, 1, 3, 4, 5, 6, 7, 8, 9, 10, 11)   # alpha, chi, gt0..5, beta0..2
PAIRS = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def band(c: np.ndarray) -> np.ndarray:
    """28 x 32 banded matrix: output r reads box positions r .. r + 4."""
    B = np.zeros((TILE, BOX), dtype=np.int64)
    for r in range(TILE):
        B[r, r:r + 2 * R + 1] = c
    return B


B1, B2 = band(C1), band(C2)


def offset(box: np.ndarray) -> float:
    """Tile offset: the midrange 0.5 max + 0.5 min (float64)."""
    return 0.5 * float(np.max(box)) + 0.5 * float(np.min(box))


def split(box: np.ndarray, k: int):
    """(digits (k, ...) int64, e) with (box - c) * 2^e = sum_s digits[s] 2^(-7 s) + remainder,
    c = offset(box)."""
    box = box - offset(box)
    m = float(np.max(np.abs(box)))
    e = 0 if m == 0.0 else 6 - int(np.frexp(m)[1])
    r = np.ldexp(box, e)
    q = np.empty((k,) + box.shape, dtype=np.int64)
    for s in range(k):
        d = np.rint(r)
        q[s] = d.astype(np.int64)
        r = (r - d) * 128.0
    return q, e


def recombine(P: np.ndarray, e: int, scale: float) -> np.ndarray:
    acc = np.zeros(P.shape[1:], dtype=np.float64)
    for s in range(P.shape[0]):
        acc = acc + np.ldexp(P[s].astype(np.float64), -7 * s)
    return np.ldexp(acc, -e) * scale


def _boxes(n1: int, n2: int):
    for o1 in range(0, n1, TILE):
        for o2 in range(0, n2, TILE):
            i1 = np.clip(np.arange(o1 - R, o1 - R + BOX), 0, n1 - 1)
            i2 = np.clip(np.arange(o2 - R, o2 - R + BOX), 0, n2 - 1)
            yield o1, o2, i1, i2


def plane_derivs(plane: np.ndarray, h1: float, h2: float, k: int, want: set) -> dict:
    """Derivatives of one 2D plane (axes 1, 2). want: subset of
    {"d1a", "d2a", "d1b", "d2b", "mix"}. Returns arrays of the plane's shape."""
    n1, n2 = plane.shape
    out = {w: np.zeros((n1, n2)) for w in want}
    for o1, o2, i1, i2 in _boxes(n1, n2):
        q, e = split(plane[np.ix_(i1, i2)], k)
        m1, m2 = min(TILE, n1 - o1), min(TILE, n2 - o2)
        rows = slice(o1, o1 + m1)
        cols = slice(o2, o2 + m2)
        if "d1a" in want or "mix" in want:
            P1 = np.einsum("rk,skc->src", B1, q)          # rows: outputs along axis 1
        if "d1a" in want:
            out["d1a"][rows, cols] = recombine(P1, e, 1.0 / (12.0 * h1))[:m1, R:R + m2]
        if "d2a" in want:
            P = np.einsum("rk,skc->src", B2, q)
            out["d2a"][rows, cols] = recombine(P, e, 1.0 / (12.0 * (h1 * h1)))[:m1, R:R + m2]
        if "d1b" in want:
            P = np.einsum("srk,ck->src", q, B1)           # columns: outputs along axis 2
            out["d1b"][rows, cols] = recombine(P, e, 1.0 / (12.0 * h2))[R:R + m1, :m2]
        if "d2b" in want:
            P = np.einsum("srk,ck->src", q, B2)
            out["d2b"][rows, cols] = recombine(P, e, 1.0 / (12.0 * (h2 * h2)))[R:R + m1, :m2]
        if "mix" in want:
            P = np.einsum("srk,ck->src", P1, B1)
            out["mix"][rows, cols] = recombine(P, e, 1.0 / (144.0 * (h1 * h2)))[:m1, :m2]
    return out


def bundle(F: np.ndarray, h3, k: int = 7) -> np.ndarray:
    """The 138-array derivative bundle (d = 3 f + axis; d = 72 + 6 slot + pair) of a
    (24, Sx, Sy, Sz) state, every padded point."""
    _, Sx, Sy, Sz = F.shape
    hx, hy, hz = (float(v) for v in h3)
    D = np.zeros((138, Sx, Sy, Sz))
    for f in range(24):
        sl = S2_FIELDS.index(f) if f in S2_FIELDS else -1
        two = sl >= 0
        for x in range(Sx):                               # (y, z) planes
            w = {"d1a", "d1b"} | ({"d2a", "d2b", "mix"} if two else set())
            o = plane_derivs(F[f, x], hy, hz, k, w)
            D[3 * f + 1, x] = o["d1a"]
            D[3 * f + 2, x] = o["d1b"]
            if two:
                D[72 + 6 * sl + 3, x] = o["d2a"]
                D[72 + 6 * sl + 5, x] = o["d2b"]
                D[72 + 6 * sl + 4, x] = o["mix"]
        for y in range(Sy):                               # (x, z) planes
            w = {"d1a"} | ({"d2a", "mix"} if two else set())
            o = plane_derivs(F[f, :, y, :], hx, hz, k, w)
            D[3 * f + 0, :, y, :] = o["d1a"]
            if two:
                D[72 + 6 * sl + 0, :, y, :] = o["d2a"]
                D[72 + 6 * sl + 2, :, y, :] = o["mix"]
        if two:
            for z in range(Sz):                           # (x, y) planes
                o = plane_derivs(F[f, :, :, z], hx, hy, k, {"mix"})
                D[72 + 6 * sl + 1, :, :, z] = o["mix"]
    return D
