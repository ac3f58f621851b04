"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (this repository's C restatement of the reference's verbatim BSSN
path) and, when it was built, ``oracle/_ref/libbssn_ref.so`` (the reference's own vendored
Dendro-GR CSE blocks compiled where they lie, see oracle/Makefile). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and only as the
checker; the product never does.

Conventions: states are numpy float64 arrays of shape (24, Sx, Sy, Sz), C order; the scalar
pack S16 is the reference's (3D/src/bssn3d/fused_rhs_cuda.py:52-55).
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libbssn_ref.so"

VAR_NAMES = ["alpha", "chi", "K", "gt0", "gt1", "gt2", "gt3", "gt4", "gt5",
             "beta0", "beta1", "beta2", "At0", "At1", "At2", "At3", "At4", "At5",
             "Gt0", "Gt1", "Gt2", "B0", "B1", "B2"]
IDX = {n: i for i, n in enumerate(VAR_NAMES)}
S2_FIELDS = ["alpha", "chi", "gt0", "gt1", "gt2", "gt3", "gt4", "gt5", "beta0", "beta1", "beta2"]
PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]

_P = ctypes.POINTER(ctypes.c_double)
_PL = ctypes.POINTER(ctypes.c_long)
_lib = None
_ref = None


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and _ref/ when the reference tree is present) with make."""
    import subprocess
    out = subprocess.run(["make", "-C", str(HERE), "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
    return _lib


PERF_PATH = HERE / "liboracle_perf.so"


def use_build(kind: str) -> None:
    """Select the build the functions of this module call: "check" (liboracle.so: -O2 with no
    FMA contraction, the build every parity test uses) or "perf" (liboracle_perf.so: the same
    source at -O3 with AVX2/FMA, used only to time the CPU baseline)."""
    global _lib
    path = {"check": LIB_PATH, "perf": PERF_PATH}[kind]
    if not path.exists():
        build()
    _lib = ctypes.CDLL(str(path))


def set_threads(n: int) -> int:
    """OpenMP team size of the current build's parallel loops; returns the size in effect."""
    lib().orc_set_threads(int(n))
    return int(lib().orc_get_threads())


def have_ref() -> bool:
    return REF_PATH.exists()


def ref():
    global _ref
    if _ref is None:
        if not REF_PATH.exists():
            raise FileNotFoundError(f"{REF_PATH} not built (make -C oracle ref needs the reference tree)")
        _ref = ctypes.CDLL(str(REF_PATH))
    return _ref


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def _c(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def deriv_index(kind: str, field: str, i: int, j: int = -1) -> int:
    """Slot of a derivative in the 138-array bundle: 'd1' (field, axis i) or 'd2' (i <= j)."""
    if kind == "d1":
        return 3 * IDX[field] + i
    a, b = min(i, j), max(i, j)
    return 72 + 6 * S2_FIELDS.index(field) + PAIRS.index((a, b))


# --- the reference oracle's seeded single-point inputs (3D/src/bssn3d/oracle.py:46-85) -------

GRAD1_ORDER = sorted({(a, f) for a in range(3) for f in VAR_NAMES})
GRAD2_ORDER = sorted({(i, j, f) for (i, j) in PAIRS for f in S2_FIELDS})
ORACLE_SCALARS13 = np.array([2.13, 1.10, 0.90, 1.20, 0.80, 1.00, 0.50,
                             0.06, 0.013, 0.027, 0.6, 20.0, 1.7])


def random_inputs(seed: int):
    """(F[24], D[138]) drawn exactly as the reference oracle draws them: numpy
    default_rng(seed), fields in state order, then first derivatives in sorted (axis, name)
    order, then second derivatives in sorted (i, j, name) order."""
    rng = np.random.default_rng(seed)
    F = np.zeros(24)
    for name in VAR_NAMES:
        if name in ("gt0", "gt3", "gt5"):
            v = 1.0 + 0.1 * rng.uniform(-1, 1)
        elif name in ("gt1", "gt2", "gt4"):
            v = 0.05 * rng.uniform(-1, 1)
        elif name == "chi":
            v = 1.0 + 0.2 * rng.uniform(0, 1)
        elif name == "alpha":
            v = 1.0 + 0.1 * rng.uniform(-1, 1)
        else:
            v = 0.1 * rng.uniform(-1, 1)
        F[IDX[name]] = v
    D = np.zeros(138)
    for a, f in GRAD1_ORDER:
        D[deriv_index("d1", f, a)] = 0.05 * rng.uniform(-1, 1)
    for i, j, f in GRAD2_ORDER:
        D[deriv_index("d2", f, i, j)] = 0.05 * rng.uniform(-1, 1)
    return F, D


def s16_to_13(S16) -> np.ndarray:
    S = np.asarray(S16, dtype=np.float64)
    return np.array([S[0], S[7], S[8], S[9], S[10], S[11], S[12], S[1], S[2], S[3], S[4], S[5], S[6]])


# --- pointwise algebra ------------------------------------------------------------------------

def algebra(F: np.ndarray, D: np.ndarray, sc13) -> np.ndarray:
    """Oracle algebra over n points: F (24, n), D (138, n) -> (24, n)."""
    F = _c(F).reshape(24, -1)
    n = F.shape[1]
    D = _c(D).reshape(138, n)
    out = np.zeros((24, n))
    lib().orc_algebra(ctypes.c_long(n), _d(F), _d(D), _d(_c(sc13)), _d(out))
    return out


def ref_algebra(F: np.ndarray, D: np.ndarray, sc13) -> np.ndarray:
    """The reference's compiled Dendro CSE over n points (same layout)."""
    F = _c(F).reshape(24, -1)
    n = F.shape[1]
    D = _c(D).reshape(138, n)
    out = np.zeros((24, n))
    ref().ref_bssn_rhs(ctypes.c_long(n), _d(F), _d(D), _d(_c(sc13)), _d(out))
    return out


def constraints_point(F: np.ndarray, D: np.ndarray) -> np.ndarray:
    """Oracle constraints (H, M0, M1, M2) from given fields and derivatives at n points."""
    F = _c(F).reshape(24, -1)
    n = F.shape[1]
    D = _c(D).reshape(138, n)
    out = np.zeros((4, n))
    lib().orc_constraints_points(ctypes.c_long(n), _d(F), _d(D), _d(out))
    return out


def ref_constraints_point(F: np.ndarray, D: np.ndarray) -> np.ndarray:
    F = _c(F).reshape(24, -1)
    n = F.shape[1]
    D = _c(D).reshape(138, n)
    out = np.zeros((4, n))
    ref().ref_bssn_constraints(ctypes.c_long(n), _d(F), _d(D), _d(out))
    return out


# --- grid operators ---------------------------------------------------------------------------

def stencil(kind: int, order: int) -> np.ndarray:
    c = np.zeros(9)
    n = lib().orc_stencil(kind, order, _d(c))
    if n < 0:
        raise ValueError(f"unsupported order {order}")
    return c[:n]


def apply_axis(u: np.ndarray, axis: int, kind: int, order: int, factor: float) -> np.ndarray:
    u = _c(u)
    out = np.zeros_like(u)
    Sx, Sy, Sz = u.shape
    lib().orc_apply_axis(_d(u), _d(out), Sx, Sy, Sz, axis, kind, order, ctypes.c_double(factor))
    return out


def deriv_bundle(F: np.ndarray, h3, order: int) -> np.ndarray:
    F = _c(F)
    _, Sx, Sy, Sz = F.shape
    D = np.zeros((138, Sx, Sy, Sz))
    lib().orc_deriv_bundle(_d(F), _d(D), Sx, Sy, Sz, _d(_c(h3)), order)
    return D


def rhs(F: np.ndarray, S16, order: int = 4, ko_order: int = 0, ko_sigma: float = 0.0) -> np.ndarray:
    F = _c(F)
    _, Sx, Sy, Sz = F.shape
    out = np.zeros_like(F)
    rc = lib().orc_rhs(_d(F), _d(out), Sx, Sy, Sz, _d(_c(S16)), order, ko_order,
                       ctypes.c_double(ko_sigma))
    if rc:
        raise ValueError("orc_rhs failed")
    return out


def rhs_points(F: np.ndarray, S16, idx, order: int = 4, ko_order: int = 0,
               ko_sigma: float = 0.0) -> np.ndarray:
    """RHS (+KO) at flat padded indices idx only -> (24, len(idx))."""
    F = _c(F)
    _, Sx, Sy, Sz = F.shape
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = np.zeros((24, idx.size))
    rc = lib().orc_rhs_points(_d(F), Sx, Sy, Sz, _d(_c(S16)), order, ko_order,
                              ctypes.c_double(ko_sigma), ctypes.c_long(idx.size),
                              idx.ctypes.data_as(_PL), _d(out))
    if rc:
        raise ValueError("orc_rhs_points failed")
    return out


LD_PATH = HERE / "liboracle_ld.so"
_lib_ld = None


def lib_ld():
    """The long-double build of the same restatement (x87, 64-bit significand)."""
    global _lib_ld
    if _lib_ld is None:
        if not LD_PATH.exists():
            build()
        _lib_ld = ctypes.CDLL(str(LD_PATH))
    return _lib_ld


def rhs_points_ld(F: np.ndarray, S16, idx, order: int = 4, ko_sigma: float = 0.0) -> np.ndarray:
    """rhs_points evaluated in long double from the same (double) inputs: the accuracy
    reference for FP64 evaluations of the RHS. Returns float64 (24, len(idx)) rounded from the
    long-double result and the rounding residual, as (hi, lo)."""
    _PLD = ctypes.POINTER(ctypes.c_longdouble)
    Fl = np.ascontiguousarray(np.asarray(F, dtype=np.float64).astype(np.longdouble))
    Sl = np.ascontiguousarray(np.asarray(S16, dtype=np.float64).astype(np.longdouble))
    _, Sx, Sy, Sz = Fl.shape
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = np.zeros((24, idx.size), dtype=np.longdouble)
    rc = lib_ld().orc_rhs_points(Fl.ctypes.data_as(_PLD), Sx, Sy, Sz, Sl.ctypes.data_as(_PLD),
                                 order, 0, ctypes.c_longdouble(ko_sigma),
                                 ctypes.c_long(idx.size), idx.ctypes.data_as(_PL),
                                 out.ctypes.data_as(_PLD))
    if rc:
        raise ValueError("orc_rhs_points (long double) failed")
    hi = out.astype(np.float64)
    return hi, (out - hi).astype(np.float64)


def constraints(F: np.ndarray, h3, order: int = 4) -> np.ndarray:
    F = _c(F)
    _, Sx, Sy, Sz = F.shape
    HM = np.zeros((4, Sx, Sy, Sz))
    lib().orc_constraints(_d(F), _d(HM), Sx, Sy, Sz, _d(_c(h3)), order)
    return HM


def constraint_l2(F: np.ndarray, h3, order: int, ng: int):
    F = _c(F)
    _, Sx, Sy, Sz = F.shape
    out = np.zeros(2)
    lib().orc_constraint_l2(_d(F), Sx, Sy, Sz, _d(_c(h3)), order, ng, _d(out))
    return float(out[0]), float(out[1])


def periodic_sync(F: np.ndarray, ng: int) -> np.ndarray:
    F = _c(F).copy()
    _, Sx, Sy, Sz = F.shape
    lib().orc_periodic_sync(_d(F), Sx, Sy, Sz, ng)
    return F


def enforce(F: np.ndarray, chi_floor: float = 1e-4, alpha_floor: float = 1e-4) -> np.ndarray:
    F = _c(F).copy()
    P = F[0].size
    lib().orc_enforce(_d(F), ctypes.c_long(P), ctypes.c_double(chi_floor),
                      ctypes.c_double(alpha_floor))
    return F


def rk4_step(F: np.ndarray, S16, t: float, dt: float, order: int = 4, ko_order: int = 0,
             ko_sigma: float = 0.0, ng: int = 3, periodic: bool = True,
             chi_floor: float = 1e-4, alpha_floor: float = 1e-4) -> np.ndarray:
    U = _c(F).copy()
    _, Sx, Sy, Sz = U.shape
    work = np.zeros((6,) + U.shape)
    rc = lib().orc_rk4_step(_d(U), Sx, Sy, Sz, _d(_c(S16)), order, ko_order,
                            ctypes.c_double(ko_sigma), ng, int(periodic), ctypes.c_double(t),
                            ctypes.c_double(dt), ctypes.c_double(chi_floor),
                            ctypes.c_double(alpha_floor), _d(work))
    if rc:
        raise ValueError("orc_rk4_step failed")
    return U


def stage_rhs(F: np.ndarray, S16, t: float, order: int = 4, ko_order: int = 0,
              ko_sigma: float = 0.0, ng: int = 3, periodic: bool = True) -> np.ndarray:
    Y = _c(F)
    out = np.zeros_like(Y)
    scr = np.zeros_like(Y)
    _, Sx, Sy, Sz = Y.shape
    rc = lib().orc_stage_rhs(_d(Y), _d(out), _d(scr), Sx, Sy, Sz, _d(_c(S16)), order, ko_order,
                             ctypes.c_double(ko_sigma), ng, int(periodic), ctypes.c_double(t))
    if rc:
        raise ValueError("orc_stage_rhs failed")
    return out


def msrk_step(F: np.ndarray, fprev: np.ndarray, S16, t: float, dt: float, order: int = 4,
              ko_order: int = 0, ko_sigma: float = 0.0, ng: int = 3, periodic: bool = True,
              chi_floor: float = 1e-4, alpha_floor: float = 1e-4):
    """One rk4_2_2 step; returns (y_{n+1}, f(y_n))."""
    U = _c(F).copy()
    fp = _c(fprev).copy()
    _, Sx, Sy, Sz = U.shape
    work = np.zeros((6,) + U.shape)
    rc = lib().orc_msrk_step(_d(U), _d(fp), Sx, Sy, Sz, _d(_c(S16)), order, ko_order,
                             ctypes.c_double(ko_sigma), ng, int(periodic), ctypes.c_double(t),
                             ctypes.c_double(dt), ctypes.c_double(chi_floor),
                             ctypes.c_double(alpha_floor), _d(work))
    if rc:
        raise ValueError("orc_msrk_step failed")
    return U, fp


def msrk_coeffs() -> np.ndarray:
    """a20, a21, a30, a31, a32, b0, b1, b2, b3 of the oracle's rk4_2_2 step."""
    c = np.zeros(9)
    lib().orc_msrk_coeffs(_d(c))
    return c


def evolve(F: np.ndarray, S16, dt: float, nsteps: int, t0: float = 0.0,
           integrator: str = "rk4", **kw) -> np.ndarray:
    """BSSNEvolution.evolve semantics (3D/src/bssn3d/evolve.py:155-194)."""
    U = _c(F).copy()
    if integrator == "rk4":
        t = t0
        for _ in range(nsteps):
            U = rk4_step(U, S16, t, dt, **kw)
            t = t + dt
        return U
    step_kw = {k: v for k, v in kw.items() if k in ("order", "ko_order", "ko_sigma", "ng", "periodic")}
    S = np.array(S16, dtype=np.float64).copy()
    S[2] = dt
    fprev = stage_rhs(U, S, t0, **step_kw)
    U = rk4_step(U, S16, t0, dt, **kw)
    t = t0 + dt
    for _ in range(1, nsteps):
        U, fprev = msrk_step(U, fprev, S16, t, dt, **kw)
        t = t + dt
    return U


def num_threads() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


_PLD = ctypes.POINTER(ctypes.c_longdouble)


def deriv_points(F: np.ndarray, h3, idx, order: int = 4, long_double: bool = False) -> np.ndarray:
    """The 138-derivative bundle (orc_deriv_points: edge-padded stencils, mixed = nested first
    derivatives, derivative_bundle.py:31-48) at flat padded indices idx only -> (138, len(idx))
    float64; long_double evaluates it in the long-double build from the same double inputs
    (returned rounded to float64)."""
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    _, Sx, Sy, Sz = np.shape(F)
    if long_double:
        Fl = np.ascontiguousarray(np.asarray(F, dtype=np.float64).astype(np.longdouble))
        hl = np.ascontiguousarray(np.asarray(h3, dtype=np.float64).astype(np.longdouble))
        out = np.zeros((138, idx.size), dtype=np.longdouble)
        lib_ld().orc_deriv_points(Fl.ctypes.data_as(_PLD), Sx, Sy, Sz, hl.ctypes.data_as(_PLD),
                                  order, ctypes.c_long(idx.size), idx.ctypes.data_as(_PL),
                                  out.ctypes.data_as(_PLD))
        return out.astype(np.float64)
    out = np.zeros((138, idx.size))
    lib().orc_deriv_points(_d(_c(F)), Sx, Sy, Sz, _d(_c(h3)), order, ctypes.c_long(idx.size),
                           idx.ctypes.data_as(_PL), _d(out))
    return out


def rk4_step_ld(F: np.ndarray, S16, t: float, dt: float, order: int = 4, ko_order: int = 0,
                ko_sigma: float = 0.0, ng: int = 3, periodic: bool = True,
                chi_floor: float = 1e-4, alpha_floor: float = 1e-4) -> np.ndarray:
    """rk4_step evaluated in long double from the same double inputs: the accuracy reference
    for FP64 steps. Returns the long-double result (numpy longdouble)."""
    U = np.ascontiguousarray(np.asarray(F, dtype=np.float64).astype(np.longdouble))
    Sl = np.ascontiguousarray(np.asarray(S16, dtype=np.float64).astype(np.longdouble))
    _, Sx, Sy, Sz = U.shape
    work = np.zeros((6,) + U.shape, dtype=np.longdouble)
    L = ctypes.c_longdouble
    rc = lib_ld().orc_rk4_step(U.ctypes.data_as(_PLD), Sx, Sy, Sz, Sl.ctypes.data_as(_PLD),
                               order, ko_order, L(ko_sigma), ng, int(periodic), L(t), L(dt),
                               L(chi_floor), L(alpha_floor), work.ctypes.data_as(_PLD))
    if rc:
        raise ValueError("orc_rk4_step (long double) failed")
    return U
