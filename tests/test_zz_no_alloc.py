"""bssn_constraint_l2_static_f64: the interior constraint norms with no allocation (workspace in
a static device array) against the allocating entry point bssn_constraint_l2_f64.

GPU: two calls in a row on one device -- the 48^3 single puncture (54^3 padded), then a smaller
20^3 grid, so the second call reuses the array the first one filled -- must equal the
allocating entry point bit for bit (same kernels; the final sums run in the host loop's order on
the device), at FD orders 4 and 6.
CPU: a grid beyond the static capacity is rejected before any device work.
"""
from __future__ import annotations

import ctypes

import pytest
import torch

from conftest import gpu_available

from bssn3d_hip import _lib
from bssn3d_hip import initial_data as idata
from bssn3d_hip.constraints import ConstraintSolver
from bssn3d_hip.rhs import stream_handle


def test_static_l2_rejects_oversize_grids():
    lib = _lib.load()
    cap = lib.bssn_l2_static_capacity_doubles()
    assert cap == 10 * 1024 * 1024
    out = (ctypes.c_double * 2)()
    h3 = _lib.dbl([0.1, 0.1, 0.1])
    fake = ctypes.c_void_p(256)
    # 200^3 padded: 4 * 8e6 doubles of constraint arrays alone exceed the capacity
    assert lib.bssn_constraint_l2_static_f64(fake, 200, 200, 200, 4, h3, 3, out, None) == 1
    assert b"static workspace" in lib.bssn_last_error()
    assert lib.bssn_constraint_l2_static_f64(None, 20, 20, 20, 4, h3, 3, out, None) == 1


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")
@pytest.mark.parametrize("order", [4, 6])
def test_static_l2_equals_allocating_entry_point(order):
    lib = _lib.load()
    for N in (48, 20):
        print(f"[case] static l2 order {order} N {N}", flush=True)
        grid, st, _ = idata.config_single_puncture(seed=N, N=N, ng=3)
        F = st.to("cuda:0")
        _, Sx, Sy, Sz = F.shape
        out = (ctypes.c_double * 2)()
        _lib.check(lib.bssn_constraint_l2_static_f64(F.data_ptr(), Sx, Sy, Sz, order,
                                                     _lib.dbl([grid.dx, grid.dy, grid.dz]),
                                                     grid.ng, out, stream_handle(F.device)))
        want = ConstraintSolver(grid, order=order).l2(F)
        assert (out[0], out[1]) == want, (N, (out[0], out[1]), want)
        assert out[0] > 0.0 and out[1] > 0.0
        torch.cuda.synchronize()
