"""Write the golden vectors of the BSSN algebra and constraints from the reference itself.

Run in the container that holds the reference tree, after `make -C oracle ref` has compiled the
reference's vendored Dendro-GR CSE blocks into oracle/_ref/libbssn_ref.so:

    python tests/golden/make_golden.py

Outputs (committed; data only):
  algebra_points.npz   seeds 0..31 drawn with the reference oracle's recipe
                       (3D/src/bssn3d/oracle.py:46-85, the seeds 0..4 of
                       3D/tests/integration/test_bssn_oracle.py:24 included): F (24, n),
                       D (138, n), sc13 (the oracle's fixed scalars), and the reference CSE's
                       24 right-hand sides `rhs` (24, n) and constraints `hm` (4, n).
  grid_rhs.npz         one small padded grid (a perturbed Brill-Lindquist puncture, 8^3
                       interior + 3 ghosts, h = 0.25): the state, the scalar pack, and the
                       reference CSE applied to the FD bundle at every padded point, for
                       FD orders 4 and 6.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "integer-schemes_amd"))

import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent


def points(nseeds: int = 32):
    Fs, Ds = [], []
    for seed in range(nseeds):
        F, D = O.random_inputs(seed)
        Fs.append(F)
        Ds.append(D)
    F = np.stack(Fs, axis=1)
    D = np.stack(Ds, axis=1)
    sc = O.ORACLE_SCALARS13.copy()
    rhs = O.ref_algebra(F, D, sc)
    hm = O.ref_constraints_point(F, D)
    np.savez_compressed(OUT / "algebra_points.npz", F=F, D=D, sc13=sc, rhs=rhs, hm=hm,
                        seeds=np.arange(nseeds))


def grid():
    import torch
    from bssn3d_hip import initial_data as idata
    g, state, pos = idata.config_single_puncture(seed=7, N=8)
    F = state.numpy()
    S16 = np.array([2.0, 0.06, 0.25 * g.dx, g.dx, 0.6, 20.0, 0.3, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0,
                    g.dx, g.dy, g.dz])
    h3 = [g.dx, g.dy, g.dz]
    res = {"F": F, "S16": S16, "pos": np.array(pos)}
    for order in (4, 6):
        D = O.deriv_bundle(F, h3, order)
        out = O.ref_algebra(F.reshape(24, -1), D.reshape(138, -1), O.s16_to_13(S16))
        res[f"rhs{order}"] = out.reshape(F.shape)
    np.savez_compressed(OUT / "grid_rhs.npz", **res)
    del torch


if __name__ == "__main__":
    if not O.have_ref():
        raise SystemExit("oracle/_ref/libbssn_ref.so missing: run `make -C oracle ref` first")
    points()
    grid()
    for p in sorted(OUT.glob("*.npz")):
        print(p.name, p.stat().st_size)
