"""ctypes binding of libbssn_hip.so (the C ABI declared in include/bssn_hip.h).

The library is built in-tree by integer-schemes_amd/build.py. A missing library raises
FileNotFoundError naming the build command, as the reference does for its FFI library
(3D/src/bssn3d/fused_rhs_cuda.py:41-49). There is no fallback path: every operation of this
package runs the HIP kernels or fails.
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

# BSSN_HIP_LIB names another build of the same library (a variant the autotuner times)
LIB_PATH = Path(os.environ.get("BSSN_HIP_LIB") or
                Path(__file__).resolve().parent / "libbssn_hip.so")
BUILD_HINT = "python integer-schemes_amd/build.py"
ABI_VERSION = 1

_c_int, _c_long, _c_double, _c_void_p, _c_size_t = (
    ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t)
_PD = ctypes.POINTER(ctypes.c_double)
_c_ll = ctypes.c_longlong
_PI = ctypes.POINTER(ctypes.c_int)
_PLL = ctypes.POINTER(ctypes.c_longlong)

# name -> (restype, argtypes); mirrors include/bssn_hip.h
SIGNATURES = {
    "bssn_abi_version": (_c_int, []),
    "bssn_last_error": (ctypes.c_char_p, []),
    "bssn_rhs_f64": (_c_int, [_c_void_p, _PD, _c_void_p, _c_int, _c_int, _c_int, _c_int,
                              _c_double, _c_void_p]),
    "bssn_rhs_sdev_f64": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_int, _c_int,
                                   _c_int, _c_double, _c_void_p]),
    "bssn_rk4_step_sdev_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int,
                                        _c_void_p, _c_int, _c_double, _c_int, _c_int, _c_double,
                                        _c_double, _c_void_p]),
    "bssn_deriv_bundle_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_int,
                                       _PD, _c_void_p]),
    "bssn_deriv_bundle_i8_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_int,
                                          _PD, _c_int, _c_void_p]),
    "bssn_algebra_f64": (_c_int, [_c_void_p, _c_void_p, _PD, _c_void_p, _c_long, _c_void_p]),
    "bssn_ko_add_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_int,
                                 _c_double, _PD, _c_int, _c_int, _c_void_p]),
    "bssn_periodic_sync_f64": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_int, _c_void_p]),
    "bssn_enforce_f64": (_c_int, [_c_void_p, _c_long, _c_double, _c_double, _c_void_p]),
    "bssn_step_workspace_bytes": (_c_size_t, [_c_int, _c_int, _c_int, _c_int]),
    "bssn_rk4_step_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _PD, _c_int,
                                   _c_double, _c_int, _c_int, _c_double, _c_double, _c_double,
                                   _c_double, _c_void_p]),
    "bssn_rk4_step_devtime_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _PD,
                                           _c_int, _c_double, _c_int, _c_int, _c_void_p,
                                           _c_double, _c_double, _c_double, _c_void_p]),
    "bssn_sommerfeld_rhs_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_int,
                                         _PD, _c_int, _PD, _c_void_p]),
    "bssn_rk4_step_sommerfeld_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _PD,
                                              _c_int, _c_double, _c_int, _PD, _c_double,
                                              _c_double, _c_double, _c_double, _c_void_p]),
    "bssn_msrk_step_f64": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_int, _c_int,
                                    _c_int, _PD, _c_int, _c_double, _c_int, _c_int, _c_double,
                                    _c_double, _c_double, _c_double, _c_void_p]),
    "bssn_stage_rhs_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _PD, _c_int,
                                    _c_double, _c_int, _c_int, _c_void_p]),
    "bssn_constraints_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int, _c_int, _PD,
                                      _c_void_p]),
    "bssn_l2_partial_doubles": (_c_long, [_c_int, _c_int, _c_int, _c_int]),
    "bssn_constraint_l2_f64": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_int, _c_int,
                                        _c_int, _PD, _c_int, _PD, _c_void_p]),
    "bssn_l2_static_capacity_doubles": (_c_long, []),
    "bssn_constraint_l2_static_f64": (_c_int, [_c_void_p, _c_int, _c_int, _c_int, _c_int, _PD,
                                               _c_int, _PD, _c_void_p]),
    # block batches
    "bssn_batch_plan_bytes": (_c_size_t, [_c_int, _PI, _c_int]),
    "bssn_batch_create": (_c_int, [ctypes.POINTER(_c_void_p), _c_void_p, _c_size_t, _c_int, _PLL,
                                   _PI, _PD, _c_int, _c_ll, _c_void_p]),
    "bssn_batch_destroy": (None, [_c_void_p]),
    "bssn_batch_set_launch": (_c_int, [_c_void_p, _c_int]),
    "bssn_batch_points": (_c_ll, [_c_void_p, _c_int]),
    "bssn_batch_rhs_f64": (_c_int, [_c_void_p, _c_void_p, _PD, _c_void_p, _c_int, _c_double,
                                    _c_void_p]),
    "bssn_batch_bundle_doubles": (_c_ll, [_c_void_p]),
    "bssn_batch_deriv_bundle_i8_f64": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_int,
                                                _c_void_p]),
    "bssn_batch_step_workspace_bytes": (_c_size_t, [_c_void_p, _c_int]),
    "bssn_batch_periodic_sync_f64": (_c_int, [_c_void_p, _c_void_p, _c_void_p]),
    "bssn_batch_halo_f64": (_c_int, [_c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p]),
    "bssn_count_skipped_entries": (_c_ll, [_c_int, _c_void_p, _c_ll, _c_ll]),
    "bssn_batch_rk4_step_f64": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _PD, _c_int, _c_double,
                                         _c_int, _c_void_p, _c_ll, _c_double, _c_double,
                                         _c_double, _c_double, _c_void_p, _c_void_p]),
    "bssn_prolong_f64": (_c_int, [_c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p]),
    "bssn_ghost_fill_f64": (_c_int, [_c_void_p, _c_ll, _c_void_p, _c_void_p]),
    "bssn_prolong_blocks_f64": (_c_int, [_c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p]),
    "bssn_batch_rk4_step_prog_f64": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _PD, _c_int,
                                              _c_double, _c_void_p, _c_double, _c_double,
                                              _c_double, _c_double, _c_void_p, _c_void_p]),
    "bssn_batch_rk4_step_link_f64": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _PD, _c_int,
                                              _c_double, _c_void_p, _c_ll, _c_void_p, _c_void_p,
                                              _c_double, _c_double, _c_double, _c_double,
                                              _c_void_p, _c_void_p]),
    "bssn_batch_indicator_f64": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p]),
    "bssn_bfp_bytes": (ctypes.c_size_t, [_c_int, _c_ll]),
    "bssn_deriv_bundle_absmax_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, _c_int,
                                              _c_int, _PD, _c_void_p]),
    "bssn_deriv_bundle_bfp": (_c_int, [_c_void_p, _c_void_p, ctypes.POINTER(_c_int), _c_int,
                                       _c_int, _c_int, _c_int, _c_int, _PD, _c_void_p]),
    "bssn_algebra_bfp_f64": (_c_int, [_c_void_p, _c_void_p, _c_int, _PD, _c_void_p, _c_ll,
                                      _c_void_p]),
}


class GhostProgram(ctypes.Structure):
    """bssn_ghost_program: phase table (host) and entry arrays (device)."""
    _fields_ = [("nphase", ctypes.c_int),
                ("prolong_end", ctypes.POINTER(ctypes.c_longlong)),
                ("block_end", ctypes.POINTER(ctypes.c_longlong)),
                ("copy_end", ctypes.POINTER(ctypes.c_longlong)),
                ("prolongs", ctypes.c_void_p),
                ("blocks", ctypes.c_void_p),
                ("copies", ctypes.c_void_p),
                ("tile_end", ctypes.POINTER(ctypes.c_longlong)),
                ("tiles", ctypes.c_void_p),
                ("tile_points", ctypes.c_void_p),
                ("n_tile_points", ctypes.c_longlong)]

_lib = None
_lock = threading.Lock()


class BSSNError(RuntimeError):
    """A nonzero return code from the library, with its bssn_last_error() message."""


def load():
    """Load (once) and return the library with typed entry points."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise FileNotFoundError(f"{LIB_PATH} not built. Run: {BUILD_HINT}")
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            if lib.bssn_abi_version() != ABI_VERSION:
                raise RuntimeError(f"libbssn_hip ABI {lib.bssn_abi_version()} != {ABI_VERSION}")
            _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().bssn_last_error().decode(errors="replace")
        raise BSSNError(f"libbssn_hip error {rc}: {msg}")


def dbl(values) -> ctypes.Array:
    """A host double array for the scalar pack / spacing arguments."""
    vals = [float(v) for v in values]
    return (ctypes.c_double * len(vals))(*vals)
