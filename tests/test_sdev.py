"""GPU: the entry points that read the 16-scalar pack from device memory.

The reference's FFI handler takes S as a device buffer, Fused(stream, F, S) -> OUT
(3D/src/bssn3d/cuda/rhs_fused.cu:22-61, S built on the device in fused_rhs_cuda.py:52-55,70-71).
bssn_rhs_sdev_f64 and bssn_rk4_step_sdev_f64 read the pack on the device (no copy, no stream
synchronisation); they must equal the host-pack entry points bit for bit, FD orders 4 and 6,
Kreiss-Oliger on and off, a nonzero time (the SSL factor), wrapped and carried ghosts, and a
captured step replayed with eager work between the replays.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import gpu_available

from bssn3d_hip import _lib
from bssn3d_hip import initial_data as idata
from bssn3d_hip.rhs import stream_handle
from bssn3d_hip.state import PhysicsParams, scalars

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

DEV = "cuda:0"


def say(msg):
    print(f"[case] {msg}", flush=True)


def _state(N=20, seed=3):
    grid, st, _ = idata.config_single_puncture(seed=seed, N=N)
    return grid, st.numpy()


@pytest.mark.parametrize("order,ko", [(4, 0.0), (4, 0.1), (6, 0.0), (6, 0.1), (8, 0.1)])
def test_rhs_device_pack_is_bit_identical(order, ko):
    say(f"bssn_rhs_sdev_f64 vs bssn_rhs_f64, order {order}, ko {ko}")
    grid, F = _state()
    dt = 0.25 * grid.dx
    S16 = scalars(PhysicsParams(), grid.dx, grid.dy, grid.dz, dt, 1.7)
    lib = _lib.load()
    Fd = torch.from_numpy(F).to(DEV)
    a = torch.empty_like(Fd)
    b = torch.full_like(Fd, np.nan)
    Sd = torch.tensor(np.asarray(S16, dtype=np.float64), device=DEV)
    sh = stream_handle(Fd.device)
    Sx, Sy, Sz = F.shape[1:]
    _lib.check(lib.bssn_rhs_f64(Fd.data_ptr(), _lib.dbl(S16), a.data_ptr(), Sx, Sy, Sz, order,
                                ko, sh))
    _lib.check(lib.bssn_rhs_sdev_f64(Fd.data_ptr(), Sd.data_ptr(), b.data_ptr(), Sx, Sy, Sz,
                                     order, ko, sh))
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert bool(torch.isfinite(a).all())


@pytest.mark.parametrize("order,periodic", [(4, 1), (4, 2), (6, 1), (8, 1), (4, 0)])
def test_rk4_step_device_pack_is_bit_identical(order, periodic):
    say(f"bssn_rk4_step_sdev_f64 vs bssn_rk4_step_f64, order {order}, periodic {periodic}")
    grid, F = _state(N=16)
    ng = grid.ng if order == 4 else 4
    if order >= 6:
        pad = np.pad(F, ((0, 0), (1, 1), (1, 1), (1, 1)), mode="edge")
        F = pad
    Sx, Sy, Sz = F.shape[1:]
    dt = 0.25 * grid.dx
    t0 = 0.3
    p = PhysicsParams()
    lib = _lib.load()
    U1 = torch.from_numpy(F).to(DEV)
    U2 = U1.clone()
    nbytes = lib.bssn_step_workspace_bytes(0, Sx, Sy, Sz)
    work = torch.empty(nbytes // 8, dtype=torch.float64, device=DEV)
    sh = stream_handle(U1.device)
    Sd = torch.tensor(np.asarray(scalars(p, grid.dx, grid.dy, grid.dz, dt, t0)), device=DEV)
    for k in range(2):
        S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, t0 + k * dt)
        _lib.check(lib.bssn_rk4_step_f64(U1.data_ptr(), work.data_ptr(), Sx, Sy, Sz,
                                         _lib.dbl(S16), order, 0.1, ng, periodic, t0 + k * dt,
                                         dt, p.chi_floor, p.alpha_floor, sh))
        _lib.check(lib.bssn_rk4_step_sdev_f64(U2.data_ptr(), work.data_ptr(), Sx, Sy, Sz,
                                              Sd.data_ptr(), order, 0.1, ng, periodic,
                                              p.chi_floor, p.alpha_floor, sh))
    torch.cuda.synchronize()
    assert torch.equal(U1, U2)
    # the pack's time advanced by two steps on the device, exactly as the host adds
    t_host = t0
    for _ in range(2):
        t_host = t_host + dt
    assert float(Sd[6]) == t_host


def test_device_pack_step_graph_replays():
    say("captured bssn_rk4_step_sdev_f64 replayed 3 times vs eager host-pack steps")
    grid, F = _state(N=16)
    Sx, Sy, Sz = F.shape[1:]
    dt = 0.25 * grid.dx
    p = PhysicsParams()
    lib = _lib.load()
    Ue = torch.from_numpy(F).to(DEV)
    Ug = Ue.clone()
    nbytes = lib.bssn_step_workspace_bytes(0, Sx, Sy, Sz)
    we = torch.empty(nbytes // 8, dtype=torch.float64, device=DEV)
    wg = torch.empty_like(we)
    Sd = torch.tensor(np.asarray(scalars(p, grid.dx, grid.dy, grid.dz, dt, 0.0)), device=DEV)
    # first launches outside the capture (code-object loading), on copies
    warm, Sw = Ug.clone(), Sd.clone()
    _lib.check(lib.bssn_rk4_step_sdev_f64(warm.data_ptr(), wg.data_ptr(), Sx, Sy, Sz,
                                          Sw.data_ptr(), 4, 0.1, grid.ng, 1, p.chi_floor,
                                          p.alpha_floor, stream_handle(Ug.device)))
    torch.cuda.synchronize()
    del warm, Sw
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            _lib.check(lib.bssn_rk4_step_sdev_f64(Ug.data_ptr(), wg.data_ptr(), Sx, Sy, Sz,
                                                  Sd.data_ptr(), 4, 0.1, grid.ng, 1,
                                                  p.chi_floor, p.alpha_floor,
                                                  stream_handle(Ug.device)))
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    # capture ran nothing: Ug and the pack are unchanged
    assert float(Sd[6]) == 0.0
    t = 0.0
    sh = stream_handle(Ue.device)
    for r in range(3):
        graph.replay()
        torch.cuda.synchronize()
        S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, t)
        _lib.check(lib.bssn_rk4_step_f64(Ue.data_ptr(), we.data_ptr(), Sx, Sy, Sz,
                                         _lib.dbl(S16), 4, 0.1, grid.ng, 1, t, dt, p.chi_floor,
                                         p.alpha_floor, sh))
        torch.cuda.synchronize()
        t = t + dt
        assert torch.equal(Ue, Ug), f"replay {r}"
        assert float(Sd[6]) == t
