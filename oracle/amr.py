"""Node-centred block hierarchy transfers and the shared-dt hierarchy RK4 step, in numpy.

TEST INFRASTRUCTURE ONLY: the device prolongation, indicator, ghost sync and hierarchy step
are checked against this. Nothing in the product imports it.

Follows the reference's node-centred 3D AMR and its N-level step:
  * weights: Lagrange product at +1/2 over offsets -2..3 (3D/src/multipatch/amr/kernels.py:
    34-45);
  * 1D refinement: edge-pad the coarse array by (2, 3), fine[2k] = coarse[k], fine[2k+1] =
    0.0 + w_0 c[k-2] + ... + w_5 c[k+3] summed left to right (kernels.py:48-63); prolongation
    = refinement along x, then y, then z of the parent's padded block, then the child window
    starting at ng + corner BS (kernels.py:66-86);
  * restriction: injection of the child's even interior nodes (kernels.py:111-140);
  * indicator: max |diff| over the interior plus one ring, all axes (kernels.py:143-158), per
    field;
  * sync: root level = stitch the root interiors, wrap-pad by ng, re-extract (the periodic root
    of 2D/src/mcs2d/amr/kernels.py:253-310 in 3D); each finer level in turn: every ghost
    point from a prolongation of the parent (3D/src/multipatch/amr/sync.py:84-97), then the
    face slabs (interior extent on the other axes) from same-level neighbours found by bbox
    adjacency without wrap (sync.py:100-122);
  * step: per stage, RHS + KO of each block's padded array (oracle.rhs) at its interior,
    interior update from the synced start state, sync; RK4 weights and order of
    3D/src/bssn3d/evolve.py:125-132 with enforcement of the final interiors
    (2D/src/mcs2d/amr/evolve.py:745-771 for the stage pattern).
Hierarchy = {key: (level, bbox)} with bbox the level-L node index of the first interior node.
"""
from __future__ import annotations

import numpy as np

import oracle as O

OFFS = (-2, -1, 0, 1, 2, 3)


def lagrange_weights(offsets, target):
    w = []
    for j, xj in enumerate(offsets):
        p = 1.0
        for m, xm in enumerate(offsets):
            if m != j:
                p *= (target - xm) / (xj - xm)
        w.append(p)
    return np.array(w, dtype=np.float64)


W_MID = lagrange_weights(OFFS, 0.5)


def refine_1d(a: np.ndarray, axis: int) -> np.ndarray:
    N = a.shape[axis]
    pw = [(0, 0)] * a.ndim
    pw[axis] = (2, 3)
    p = np.pad(a, pw, mode="edge")
    mids = 0.0
    for w, off in zip(W_MID, OFFS):
        sl = [slice(None)] * a.ndim
        sl[axis] = slice(2 + off, 2 + off + N)
        mids = mids + w * p[tuple(sl)]
    shape = list(a.shape)
    shape[axis] = 2 * N
    out = np.empty(shape)
    ev = [slice(None)] * a.ndim
    od = [slice(None)] * a.ndim
    ev[axis] = slice(0, None, 2)
    od[axis] = slice(1, None, 2)
    out[tuple(ev)] = a
    out[tuple(od)] = mids
    return out


def prolongate(parent: np.ndarray, corner, BS: int, ng: int) -> np.ndarray:
    """The child's whole padded block (24, W, W, W) from its parent's."""
    r = parent
    for ax in (1, 2, 3):
        r = refine_1d(r, ax)
    W = BS + 2 * ng
    s = [ng + corner[a] * BS for a in range(3)]
    return r[:, s[0]:s[0] + W, s[1]:s[1] + W, s[2]:s[2] + W].copy()


def restrict_into_parent(parent: np.ndarray, child: np.ndarray, corner, BS: int,
                         ng: int) -> np.ndarray:
    half = BS // 2
    out = parent.copy()
    rc = child[:, ng:ng + BS:2, ng:ng + BS:2, ng:ng + BS:2]
    i = [ng + corner[a] * half for a in range(3)]
    out[:, i[0]:i[0] + half, i[1]:i[1] + half, i[2]:i[2] + half] = rc
    return out


def indicator(block: np.ndarray, ng: int) -> np.ndarray:
    """Per field: max |u(p + e_a) - u(p)| over the interior plus one ring (NaN if any)."""
    S = block.shape[1:]
    core = block[:, ng - 1:S[0] - ng + 1, ng - 1:S[1] - ng + 1, ng - 1:S[2] - ng + 1]
    out = np.zeros(block.shape[0])
    for ax in (1, 2, 3):
        d = np.abs(np.diff(core, axis=ax)).reshape(block.shape[0], -1)
        out = np.maximum(out, d.max(axis=1))       # np.maximum propagates NaN
    return out


# --- hierarchy -------------------------------------------------------------------------------

def parent_of(key, hier, BS):
    L, cb = hier[key]
    f0 = [cb[a] // 2 for a in range(3)]
    for k, (lv, pb) in hier.items():
        if lv == L - 1 and all(pb[a] <= f0[a] < pb[a] + BS for a in range(3)):
            return k, tuple((f0[a] - pb[a]) // (BS // 2) for a in range(3))
    raise ValueError(f"block {key} has no parent")


def sync_all(blocks: dict, hier: dict, nb_root, BS: int, ng: int) -> dict:
    out = {k: v.copy() for k, v in blocks.items()}
    W = BS + 2 * ng
    N = [n * BS for n in nb_root]
    nf = next(iter(blocks.values())).shape[0]
    g = np.zeros((nf, *N))
    roots = [k for k, (lv, _) in hier.items() if lv == 0]
    for k in roots:
        b = hier[k][1]
        g[:, b[0]:b[0] + BS, b[1]:b[1] + BS, b[2]:b[2] + BS] = \
            out[k][:, ng:ng + BS, ng:ng + BS, ng:ng + BS]
    gp = np.pad(g, ((0, 0), (ng, ng), (ng, ng), (ng, ng)), mode="wrap")
    for k in roots:
        b = hier[k][1]
        out[k] = gp[:, b[0]:b[0] + W, b[1]:b[1] + W, b[2]:b[2] + W].copy()
    nlev = max(lv for lv, _ in hier.values()) + 1
    halo = np.ones((W, W, W), bool)
    halo[ng:ng + BS, ng:ng + BS, ng:ng + BS] = False
    for L in range(1, nlev):
        keys = [k for k, (lv, _) in hier.items() if lv == L]
        for k in keys:
            pk, corner = parent_of(k, hier, BS)
            pro = prolongate(out[pk], corner, BS, ng)
            out[k][:, halo] = pro[:, halo]
        pos = {hier[k][1]: k for k in keys}
        I = slice(ng, ng + BS)
        for k in keys:
            b = hier[k][1]
            for ax in range(3):
                for side in (0, 1):
                    nbp = list(b)
                    nbp[ax] += BS if side else -BS
                    nb = pos.get(tuple(nbp))
                    if nb is None:
                        continue
                    dst, src = [slice(None), I, I, I], [slice(None), I, I, I]
                    if side == 0:
                        dst[1 + ax], src[1 + ax] = slice(0, ng), slice(BS, BS + ng)
                    else:
                        dst[1 + ax], src[1 + ax] = slice(ng + BS, 2 * ng + BS), slice(ng, 2 * ng)
                    out[k][tuple(dst)] = out[nb][tuple(src)]
    return out


def rk4_step(blocks: dict, hier: dict, nb_root, BS: int, ng: int, h0: float, S16, t: float,
             dt: float, order: int = 4, ko_sigma: float = 0.1, chi_floor: float = 1e-4,
             alpha_floor: float = 1e-4) -> dict:
    """One shared-dt RK4 step of a synced hierarchy; returns the synced new hierarchy."""
    I = (slice(None),) + (slice(ng, ng + BS),) * 3
    ko_order = 0                         # the default pairing, min(order + 2, 8)

    def rhs(Y, tt):
        k = {}
        for key, arr in Y.items():
            h = h0 / 2 ** hier[key][0]
            s = np.array(S16, dtype=np.float64)
            s[2], s[6] = dt, tt
            s[3] = h
            s[13] = s[14] = s[15] = h
            k[key] = O.rhs(arr, s, order=order, ko_order=ko_order, ko_sigma=ko_sigma)[I]
        return k

    def stage(k, c):
        Y = {}
        for key, arr in blocks.items():
            y = arr.copy()
            y[I] = arr[I] + c * k[key]
            Y[key] = y
        return sync_all(Y, hier, nb_root, BS, ng)

    k1 = rhs(blocks, t)
    k2 = rhs(stage(k1, 0.5 * dt), t + 0.5 * dt)
    k3 = rhs(stage(k2, 0.5 * dt), t + 0.5 * dt)
    k4 = rhs(stage(k3, dt), t + dt)
    new = {}
    for key, arr in blocks.items():
        y = arr.copy()
        y[I] = arr[I] + (dt / 6.0) * (((k1[key] + 2.0 * k2[key]) + 2.0 * k3[key]) + k4[key])
        y[I] = O.enforce(np.ascontiguousarray(y[I]), chi_floor, alpha_floor)
        new[key] = y
    return sync_all(new, hier, nb_root, BS, ng)
