"""BSSN right-hand side on the device, behind the reference's call shapes.

Reference interface mirrored here:
  BSSNSolver(grid, params, order, dt, scheme, precision, ko_order).rhs(state, t, dt) and
  .rhs_dict(state, t, dt)                        (3D/src/bssn3d/rhs.py:63-139)
  device_rhs_fused(state, params, dx, dy, dz, dt, t=0.0)
  cuda_fused_algebra(F, dx, dy, dz, eta, lmbda, lambda_f, cahd_c, dt, ssl_h, ssl_sigma, t)
                                                 (3D/src/bssn3d/fused_rhs_cuda.py:62-80)
States are float64 torch tensors of shape (24, Sx, Sy, Sz) on a HIP device (or any object with
__dlpack__, or a BSSNState holding one). Work is enqueued on torch's current stream.

Schemes:
  "fused"   one fused kernel per evaluation (derivatives in LDS/registers, no bundle in HBM)
  "twopass" the materialised path: a 138-array derivative bundle, then the pointwise algebra
            (the data flow of the reference's "verbatim" scheme)
  "int8"    the two-pass path with the bundle taken from int8 digits on the INT8 matrix cores
            (Ozaki scheme I, 7 digits by default; FD order 4); the algebra stays FP64
  "bfp48" / "bfp40" / "bfp32" / "bfp24"
            the two-pass path with the bundle stored in block floating point of that many
            mantissa bits (one exponent per derivative array; the reference's storage-
            compression form, common/src/mcs_common/bfp_compress.py); arithmetic stays FP64
"""
from __future__ import annotations

import ctypes
from typing import Dict

import numpy as np
import torch

from . import _lib
from .grid import Grid, default_ko_order, ghost_width
from .state import BSSNState, PhysicsParams, VAR_NAMES, pack_S, scalars

SCHEMES = ("fused", "twopass", "int8", "bfp48", "bfp40", "bfp32", "bfp24")
BFP_BITS = {"bfp48": 48, "bfp40": 40, "bfp32": 32, "bfp24": 24}


def as_tensor(state) -> torch.Tensor:
    """Accept a BSSNState, a torch tensor or any __dlpack__ exporter; check layout and device."""
    data = state.data if isinstance(state, BSSNState) else state
    if not isinstance(data, torch.Tensor):
        data = torch.from_dlpack(data)
    if data.dtype != torch.float64:
        raise TypeError(f"state must be float64, got {data.dtype}")
    if data.dim() != 4 or data.shape[0] != 24:
        raise ValueError(f"state must have shape (24, Sx, Sy, Sz), got {tuple(data.shape)}")
    if not data.is_contiguous():
        raise ValueError("state must be C-contiguous")
    if data.device.type != "cuda":
        raise ValueError("state must live on a HIP device (torch 'cuda' device)")
    return data


def stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _ptr(t: torch.Tensor):
    return t.data_ptr()


def rhs_into(F: torch.Tensor, OUT: torch.Tensor, S16, order: int = 4, ko_sigma: float = 0.0) -> None:
    """OUT <- RHS(F) (+ KO when ko_sigma > 0) at every padded point, fused kernel."""
    _, Sx, Sy, Sz = F.shape
    lib = _lib.load()
    _lib.check(lib.bssn_rhs_f64(_ptr(F), _lib.dbl(S16), _ptr(OUT), Sx, Sy, Sz, order,
                                float(ko_sigma), stream_handle(F.device)))


def rhs_twopass_into(F: torch.Tensor, OUT: torch.Tensor, D: torch.Tensor, S16, order: int = 4,
                     ko_sigma: float = 0.0) -> None:
    """OUT <- algebra(F, bundle(F)) (+ KO), with the 138-array bundle materialised in D."""
    _, Sx, Sy, Sz = F.shape
    lib = _lib.load()
    st = stream_handle(F.device)
    h3 = _lib.dbl(S16[13:16])
    _lib.check(lib.bssn_deriv_bundle_f64(_ptr(F), _ptr(D), Sx, Sy, Sz, order, h3, st))
    _lib.check(lib.bssn_algebra_f64(_ptr(F), _ptr(D), _lib.dbl(S16), _ptr(OUT), Sx * Sy * Sz, st))
    if ko_sigma > 0.0:
        _lib.check(lib.bssn_ko_add_f64(_ptr(F), _ptr(OUT), Sx, Sy, Sz, order, float(ko_sigma),
                                       h3, 0, 0, st))


def rhs_int8_into(F: torch.Tensor, OUT: torch.Tensor, D: torch.Tensor, S16, order: int = 4,
                  ko_sigma: float = 0.0, slices: int = 7) -> None:
    """OUT <- algebra(F, INT8-split bundle(F)) (+ FP64 KO), bundle materialised in D."""
    _, Sx, Sy, Sz = F.shape
    lib = _lib.load()
    st = stream_handle(F.device)
    h3 = _lib.dbl(S16[13:16])
    _lib.check(lib.bssn_deriv_bundle_i8_f64(_ptr(F), _ptr(D), Sx, Sy, Sz, order, h3, int(slices),
                                            st))
    _lib.check(lib.bssn_algebra_f64(_ptr(F), _ptr(D), _lib.dbl(S16), _ptr(OUT), Sx * Sy * Sz, st))
    if ko_sigma > 0.0:
        _lib.check(lib.bssn_ko_add_f64(_ptr(F), _ptr(OUT), Sx, Sy, Sz, order, float(ko_sigma),
                                       h3, 0, 0, st))


def bfp_exponents(amax: np.ndarray, mant_bits: int) -> np.ndarray:
    """Per-array exponents floor(log2((2^(k-1) - 1) / max|d|)), 0 for an all-zero array
    (common/src/mcs_common/bfp_compress.py:37-46)."""
    amax = np.asarray(amax, dtype=np.float64)
    if not np.all(np.isfinite(amax)):
        raise ValueError("non-finite derivative: the bundle cannot be stored in BFP")
    full = 2.0 ** (mant_bits - 1) - 1.0
    with np.errstate(divide="ignore"):
        e = np.where(amax == 0.0, 0.0, np.floor(np.log2(full / np.where(amax == 0.0, 1.0, amax))))
    return e.astype(np.int64)


def bfp_buffers(F: torch.Tensor, mant_bits: int):
    """(Q, amax): BFP storage of the bundle of F's grid and the 138 per-array maxima."""
    npts = F[0].numel()
    nbytes = _lib.load().bssn_bfp_bytes(int(mant_bits), npts)
    if nbytes == 0:
        raise ValueError(f"mant_bits {mant_bits} outside 2..48")
    Q = torch.empty(int(nbytes), dtype=torch.uint8, device=F.device)
    amax = torch.empty(138, dtype=torch.float64, device=F.device)
    return Q, amax


def rhs_bfp_into(F: torch.Tensor, OUT: torch.Tensor, Q: torch.Tensor, amax: torch.Tensor, S16,
                 order: int = 4, ko_sigma: float = 0.0, mant_bits: int = 32,
                 exps: np.ndarray | None = None) -> np.ndarray:
    """OUT <- algebra(F, bundle(F) stored at mant_bits per value) (+ FP64 KO): per-array maxima
    (one pass), exponents on the host, the bundle into BFP storage Q, the algebra from Q.
    Given ``exps`` (fixed exponents, e.g. from an earlier call), the maxima pass and the host
    round trip are skipped and out-of-range values saturate (bfp_compress.py:76-83).
    Returns the exponents used."""
    _, Sx, Sy, Sz = F.shape
    lib = _lib.load()
    st = stream_handle(F.device)
    h3 = _lib.dbl(S16[13:16])
    if exps is None:
        _lib.check(lib.bssn_deriv_bundle_absmax_f64(_ptr(F), _ptr(amax), Sx, Sy, Sz, order, h3,
                                                    st))
        exps = bfp_exponents(amax.cpu().numpy(), mant_bits)
    exps = np.asarray(exps, dtype=np.int64)
    if exps.shape != (138,):
        raise ValueError("exps must hold 138 exponents")
    e = (ctypes.c_int * 138)(*[int(v) for v in exps])
    _lib.check(lib.bssn_deriv_bundle_bfp(_ptr(F), _ptr(Q), e, int(mant_bits), Sx, Sy, Sz, order,
                                         h3, st))
    _lib.check(lib.bssn_algebra_bfp_f64(_ptr(F), _ptr(Q), int(mant_bits), _lib.dbl(S16),
                                        _ptr(OUT), Sx * Sy * Sz, st))
    if ko_sigma > 0.0:
        _lib.check(lib.bssn_ko_add_f64(_ptr(F), _ptr(OUT), Sx, Sy, Sz, order, float(ko_sigma),
                                       h3, 0, 0, st))
    return exps


def cuda_fused_algebra(F, dx, dy, dz, eta, lmbda, lambda_f, cahd_c, dt, ssl_h, ssl_sigma, t,
                       order: int = 6) -> Dict[str, torch.Tensor]:
    """The reference's fused-scheme callable (fused_rhs_cuda.py:62-72): a dict of 24 fields in,
    a dict of 24 right-hand sides out. ``order`` defaults to the reference kernel's 6."""
    fields = torch.stack([F[n] for n in VAR_NAMES]).contiguous()
    fields = as_tensor(fields)
    S = pack_S(eta, cahd_c, dt, dx, ssl_h, ssl_sigma, t, lmbda, lambda_f, dy, dz)
    out = torch.empty_like(fields)
    rhs_into(fields, out, S, order)
    return {n: out[k] for k, n in enumerate(VAR_NAMES)}


def device_rhs_fused(state, params: PhysicsParams, dx, dy, dz, dt, t=0.0,
                     order: int = 6) -> Dict[str, torch.Tensor]:
    """fused_rhs_cuda.py:75-80: the fused RHS of a state, as {field: d/dt}."""
    F = as_tensor(state)
    out = torch.empty_like(F)
    rhs_into(F, out, scalars(params, dx, dy, dz, dt, t), order)
    return {n: out[k] for k, n in enumerate(VAR_NAMES)}


class BSSNSolver:
    """Grid + params + FD order -> RHS evaluator (3D/src/bssn3d/rhs.py:63-139).

    ``scheme`` is "fused" (default), "twopass" or "int8"; the reference's scheme names map as
    cuda_fused/fused/fused_fp64/fused_tiled -> "fused" and verbatim/staged/pallas -> "twopass".
    ``precision`` must be "fp64" (the "int8" scheme computes its derivatives from int8 digits
    and everything else in fp64; FD order 4).
    """

    _ALIASES = {"cuda_fused": "fused", "fused_fp64": "fused", "fused_tiled": "fused",
                "verbatim": "twopass", "staged": "twopass", "pallas": "twopass"}

    def __init__(self, grid: Grid, params: PhysicsParams | None = None, order: int = 6,
                 dt: float | None = None, scheme: str = "fused", precision: str = "fp64",
                 ko_order: int | None = None):
        self.grid = grid
        self.params = params or PhysicsParams()
        scheme = self._ALIASES.get(scheme, scheme)
        if scheme not in SCHEMES:
            raise ValueError(f"unknown scheme {scheme!r} (use one of {SCHEMES})")
        if precision != "fp64":
            raise ValueError(f"unknown precision {precision!r} (this library computes in fp64)")
        if order not in (4, 6, 8):
            raise ValueError(f"order {order} not supported on the device (4, 6 or 8)")
        if scheme == "int8" and order != 4:
            raise ValueError("the int8 scheme supports FD order 4 only")
        paired = default_ko_order(order)
        ko_order = paired if ko_order is None else ko_order
        if ko_order != paired:
            raise ValueError(f"KO order must be {paired} with FD order {order} on the device")
        ng = ghost_width(order, ko_order)
        if ng > grid.ng:
            raise ValueError(f"FD order {order} / KO order {ko_order} need ng>={ng} "
                             f"but grid has ng={grid.ng}.")
        self.scheme = scheme
        self.precision = precision
        self.order = order
        self.ko_order = ko_order
        self.dt = dt if dt is not None else 0.25 * grid.dx
        self._D = None
        self._Q = None
        self.last_exponents = None

    def S16(self, t: float = 0.0, dt: float | None = None):
        g, p = self.grid, self.params
        return scalars(p, g.dx, g.dy, g.dz, self.dt if dt is None else dt, t)

    def rhs_tensor(self, state, t: float = 0.0, dt: float | None = None,
                   out: torch.Tensor | None = None) -> torch.Tensor:
        F = as_tensor(state)
        if tuple(F.shape[1:]) != self.grid.shape:
            raise ValueError(f"state shape {tuple(F.shape[1:])} != grid shape {self.grid.shape}")
        out = torch.empty_like(F) if out is None else out
        S = self.S16(t, dt)
        if self.scheme == "fused":
            rhs_into(F, out, S, self.order)
        elif self.scheme in BFP_BITS:
            k = BFP_BITS[self.scheme]
            if self._Q is None or self._Q[0].device != F.device or \
                    self._Q[2] != tuple(F.shape):
                self._Q = bfp_buffers(F, k) + (tuple(F.shape),)
            self.last_exponents = rhs_bfp_into(F, out, self._Q[0], self._Q[1], S, self.order,
                                               mant_bits=k)
        else:
            if self._D is None or self._D.device != F.device:
                self._D = torch.empty((138,) + tuple(F.shape[1:]), dtype=torch.float64,
                                      device=F.device)
            if self.scheme == "int8":
                rhs_int8_into(F, out, self._D, S, self.order)
            else:
                rhs_twopass_into(F, out, self._D, S, self.order)
        return out

    def rhs_dict(self, state, t: float = 0.0, dt: float | None = None) -> Dict[str, torch.Tensor]:
        out = self.rhs_tensor(state, t, dt)
        return {n: out[k] for k, n in enumerate(VAR_NAMES)}

    def rhs(self, state, t: float = 0.0, dt: float | None = None) -> BSSNState:
        return BSSNState(self.rhs_tensor(state, t, dt))
