"""Host emulation of the device ghost-program entries (test helper): applies bssn_halo_copy and
bssn_prolong_point lists to a numpy pool with the kernels' arithmetic (per point: x sums,
then the y sum over them, then the z sum, each left to right from 0.0, no contraction; even
axes copy; parent reads clamped to the parent's padded block)."""
from __future__ import annotations

import numpy as np

NF = 24
W_MID = np.array([3.0, -25.0, 150.0, 150.0, -25.0, 3.0]) / 256.0


def apply_copies(pool: np.ndarray, e: np.ndarray) -> None:
    for f in range(NF):
        pool[e["dst"] + f * e["dst_stride"]] = pool[e["src"] + f * e["src_stride"]]


def _taps(fine, S):
    k = fine >> 1
    odd = (fine & 1) == 1
    idx = np.where(odd[:, None], k[:, None] + np.arange(-2, 4)[None, :], k[:, None])
    return odd, np.clip(idx, 0, S[:, None] - 1)


def apply_prolongs(pool: np.ndarray, e: np.ndarray) -> None:
    if len(e) == 0:
        return
    out = np.empty((NF, len(e)))
    pS = e["pS"].astype(np.int64)
    fine = e["fine"].astype(np.int64)
    sx, sy = pS[:, 1] * pS[:, 2], pS[:, 2]
    P = sx * pS[:, 0]
    tx = _taps(fine[:, 0], pS[:, 0])
    ty = _taps(fine[:, 1], pS[:, 1])
    tz = _taps(fine[:, 2], pS[:, 2])
    for par in range(8):
        ox, oy, oz = bool(par & 4), bool(par & 2), bool(par & 1)
        m = (tx[0] == ox) & (ty[0] == oy) & (tz[0] == oz)
        if not m.any():
            continue
        ix, iy, iz = tx[1][m], ty[1][m], tz[1][m]
        base = e["parent"][m]
        for f in range(NF):
            c0 = base + f * P[m]
            vz = np.zeros(m.sum())
            for kz in range(6 if oz else 1):
                vy = np.zeros(m.sum())
                for ky in range(6 if oy else 1):
                    row = c0 + iy[:, ky] * sy[m] + iz[:, kz]
                    if ox:
                        vx = np.zeros(m.sum())
                        for kx in range(6):
                            vx = vx + W_MID[kx] * pool[row + ix[:, kx] * sx[m]]
                    else:
                        vx = pool[row + ix[:, 0] * sx[m]]
                    vy = vy + W_MID[ky] * vx if oy else vx
                vz = vz + W_MID[kz] * vy if oz else vy
            out[f, m] = vz
    for f in range(NF):
        pool[e["dst"] + f * e["dst_stride"]] = out[f]


def job_points(job) -> np.ndarray:
    """The padded-block points a whole-block job writes: ghosts outside the skipped face slabs
    (a face slab = ghost along one axis, interior along the other two), and the interior if
    bit 6 is set."""
    BS, ng = int(job["BS"]), int(job["ng"])
    W = BS + 2 * ng
    c = np.stack(np.meshgrid(*[np.arange(W)] * 3, indexing="ij"), -1).reshape(-1, 3)
    inside = (c >= ng) & (c < ng + BS)
    keep = ~np.all(inside, axis=1) if not (job["skip"] & 64) else np.ones(len(c), bool)
    for ax in range(3):
        slab = ~inside[:, ax] & np.all(np.delete(inside, ax, axis=1), axis=1)
        lo = c[:, ax] < ng
        keep &= ~(slab & lo & bool(job["skip"] & (1 << (2 * ax))))
        keep &= ~(slab & ~lo & bool(job["skip"] & (1 << (2 * ax + 1))))
    return c[keep]


def expand_jobs(jobs) -> np.ndarray:
    """Whole-block jobs as the equivalent point entries."""
    from bssn3d_hip.amr import prolong_entries
    out = [prolong_entries(job_points(j), int(j["dst"]), int(j["parent"]), tuple(j["corner"]),
                           int(j["BS"]), int(j["ng"])) for j in jobs]
    from bssn3d_hip.amr import PROLONG_DTYPE
    return np.concatenate(out) if out else np.empty(0, dtype=PROLONG_DTYPE)


def run_program(pool: np.ndarray, prolongs, jobs, copies, pe, be, ce) -> None:
    pb = bb = cb = 0
    for p_end, b_end, c_end in zip(pe, be, ce):
        apply_prolongs(pool, prolongs[pb:p_end])
        apply_prolongs(pool, expand_jobs(jobs[bb:b_end]))
        apply_copies(pool, copies[cb:c_end])
        pb, bb, cb = p_end, b_end, c_end
