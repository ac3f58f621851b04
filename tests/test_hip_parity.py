"""GPU: the HIP library, through its C ABI, against the oracle and the reference's golden vectors.

Bar: per-field normwise relative error max|hip - ref| / max|ref| <= 1e-12 (summation-order and
FMA differences only); ghost-sync copies bit-exact. Cases cover the reference's own sizes and
edge cases: the golden point vectors (reference CSE), a padded golden grid with edge padding,
the 48^3 single-puncture workload, ragged non-cubic extents that do not divide the 4x4x4
brick (the bare RHS, and RK4 / MSRK steps on periodic grids), FD orders 4 and 6, KO on and off, periodic and edge-padded steps, RK4 and MSRK, the
constraints, and the 256^3 binary workload through sampled points and size-independent
properties.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from conftest import gpu_available, normwise

from bssn3d_hip import initial_data as idata
from bssn3d_hip.constraints import ConstraintSolver
from bssn3d_hip.evolve import BSSNEvolution
from bssn3d_hip.grid import Grid
from bssn3d_hip.rhs import BSSNSolver, rhs_into, rhs_twopass_into, stream_handle
from bssn3d_hip import _lib
from bssn3d_hip.state import PhysicsParams, pack_S, scalars

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]

TOL = 1e-12
DEV = "cuda:0"


def say(msg):
    print(f"[case] {msg}", flush=True)


def to_dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def to_host(t: torch.Tensor) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy()


def smooth_state(shape, seed=0, amp=0.05):
    rng = np.random.default_rng(seed)
    S = shape
    axes = [np.arange(n) / n for n in S]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    F = np.zeros((24,) + tuple(S))
    for f in range(24):
        k = rng.integers(1, 3, size=3)
        F[f] = amp * np.sin(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z) + rng.uniform(0, 6))
    for f in (0, 1, 3, 6, 8):
        F[f] += 1.0
    return F


def test_algebra_kernel_matches_reference_cse(golden_points):
    say("algebra kernel vs reference CSE, 32 seeded points")
    g = golden_points
    sc = g["sc13"]
    # sc13 = eta, lambda0..3, lambda_f0..1, C, dt, dx_i, h_ssl, sig, t -> the 16-pack
    S16 = pack_S(sc[0], sc[7], sc[8], sc[9], sc[10], sc[11], sc[12], sc[1:5], sc[5:7],
                 sc[9], sc[9])
    F = to_dev(g["F"])
    D = to_dev(g["D"])
    out = torch.empty_like(F)
    n = F.shape[1]
    _lib.check(_lib.load().bssn_algebra_f64(F.data_ptr(), D.data_ptr(), _lib.dbl(S16),
                                            out.data_ptr(), n, stream_handle(F.device)))
    got, ref = to_host(out), g["rhs"]
    rel = np.abs(got - ref) / np.maximum(np.maximum(np.abs(got), np.abs(ref)), 1e-12)
    assert rel.max() < TOL, rel.max()


@pytest.mark.parametrize("order", [4, 6])
def test_fused_rhs_golden_grid(golden_grid, order):
    say(f"fused RHS vs reference CSE on the golden 14^3 padded grid, order {order}")
    g = golden_grid
    F = to_dev(g["F"])
    out = torch.empty_like(F)
    rhs_into(F, out, list(g["S16"]), order)
    err = normwise(to_host(out), g[f"rhs{order}"])
    assert err.max() < TOL, err


@pytest.mark.parametrize("ko_sigma", [0.0, 0.1])
def test_fused_rhs_single_puncture(ko_sigma):
    say(f"fused RHS, 48^3 single puncture, ko_sigma={ko_sigma}")
    grid, st, _ = idata.config_single_puncture(seed=0, N=48)
    F = st.numpy()
    S16 = scalars(PhysicsParams(), grid.dx, grid.dy, grid.dz, 0.25 * grid.dx, 0.0)
    ref = O.rhs(F, S16, 4, ko_sigma=ko_sigma)
    Fd = to_dev(F)
    out = torch.empty_like(Fd)
    rhs_into(Fd, out, S16, 4, ko_sigma)
    err = normwise(to_host(out), ref)
    assert err.max() < TOL, err


def test_twopass_matches_oracle():
    say("two-pass bundle + algebra + KO, 24^3 single puncture")
    grid, st, _ = idata.config_single_puncture(seed=2, N=24)
    F = st.numpy()
    S16 = scalars(PhysicsParams(), grid.dx, grid.dy, grid.dz, 0.25 * grid.dx, 0.7)
    h3 = [grid.dx, grid.dy, grid.dz]
    Fd = to_dev(F)
    D = torch.empty((138,) + F.shape[1:], dtype=torch.float64, device=DEV)
    out = torch.empty_like(Fd)
    rhs_twopass_into(Fd, out, D, S16, 4, 0.1)
    Dref = O.deriv_bundle(F, h3, 4)
    # A derivative of a nearly constant field (gt_ii = 1 + 1e-4 waves) is a sum of O(1/h)
    # terms cancelling to O(1e-4): its rounding error scales with sum_k |w_k| max|u|, not with
    # the result. Bound each array by 64 eps times that stencil magnitude (or 1e-12 normwise).
    Dg = to_host(D)
    c1 = np.abs(O.stencil(0, 4)).sum()
    c2 = np.abs(O.stencil(1, 4)).sum()
    for d in range(138):
        if d < 72:
            f, a = divmod(d, 3)
            mag = c1 / h3[a]
        else:
            s, q = divmod(d - 72, 6)
            f = O.IDX[O.S2_FIELDS[s]]
            i, j = O.PAIRS[q]
            mag = c2 / h3[i] ** 2 if i == j else c1 * c1 / (h3[i] * h3[j])
        bound = max(TOL * np.abs(Dref[d]).max(), 64 * 2.2e-16 * mag * np.abs(F[f]).max())
        assert np.abs(Dg[d] - Dref[d]).max() <= bound, d
    err = normwise(to_host(out), O.rhs(F, S16, 4, ko_sigma=0.1))
    assert err.max() < TOL, err


@pytest.mark.parametrize("shape,order", [((13, 17, 22), 4), ((9, 11, 21), 6), ((11, 13, 18), 8),
                                         ((5, 6, 7), 4)])
def test_ragged_shapes(shape, order):
    say(f"fused RHS on ragged extents {shape}, order {order}")
    F = smooth_state(shape, seed=sum(shape))
    S16 = [2.0, 0.06, 0.02, 0.1, 0.6, 20.0, 0.4, 1.1, 0.9, 1.2, 0.8, 1.0, 0.5, 0.1, 0.11, 0.09]
    Fd = to_dev(F)
    out = torch.empty_like(Fd)
    rhs_into(Fd, out, S16, order, 0.05)
    err = normwise(to_host(out), O.rhs(F, S16, order, ko_sigma=0.05))
    assert err.max() < TOL, err


@pytest.mark.parametrize("integrator,nsteps", [("rk4", 1), ("rk4", 3), ("rk4_2_2", 4)])
def test_evolution_matches_oracle(integrator, nsteps):
    say(f"{integrator} evolve {nsteps} steps, periodic, puncture 20^3")
    grid, st, _ = idata.config_single_puncture(seed=4, N=20)
    F = st.numpy()
    dt = 0.25 * grid.dx
    p = PhysicsParams()
    evo = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="periodic", integrator=integrator)
    got = to_host(evo.evolve(to_dev(F), dt, nsteps, t0=0.2).data)
    S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, 0.2)
    ref = O.evolve(F, S16, dt, nsteps, t0=0.2, integrator=integrator, order=4, ko_sigma=0.1, ng=3)
    ref = O.periodic_sync(ref, 3)   # the reference state after the sync its next step performs
    err = normwise(got, ref)
    assert err.max() < TOL, err


def periodic_state(interior, ng, seed=0, amp=0.02):
    """Smooth state whose ghost cells are the periodic images of the interior (one sinusoid
    per field, wavenumber 1 over each interior extent); alpha, chi and the diagonal g~ near 1."""
    rng = np.random.default_rng(seed)
    axes = [(np.arange(n + 2 * ng) - ng) / n for n in interior]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    F = np.zeros((24,) + tuple(n + 2 * ng for n in interior))
    for f in range(24):
        s = rng.choice([-1, 1], size=3)
        F[f] = amp * np.sin(2 * np.pi * (s[0] * X + s[1] * Y + s[2] * Z) + rng.uniform(0, 6))
    for f in (0, 1, 3, 6, 8):
        F[f] += 1.0
    return F


@pytest.mark.parametrize("interior,order,integrator,nsteps",
                         [((10, 13, 18), 4, "rk4", 2), ((9, 14, 11), 6, "rk4_2_2", 3),
                          ((7, 9, 6), 4, "rk4_2_2", 3)])
def test_ragged_evolution_matches_oracle(interior, order, integrator, nsteps):
    """RK4 and MSRK steps on non-cubic periodic grids whose extents are not multiples of the
    4x4x4 brick: the stage kernels' partial bricks (lanes past the interior load clamped
    write-back operands and store nothing) against the oracle."""
    say(f"{integrator} {nsteps} steps on ragged periodic interior {interior}, order {order}")
    ng = 3 if order == 4 else 4
    h = (0.05, 0.045, 0.04)
    grid = Grid(Nx=interior[0], Ny=interior[1], Nz=interior[2], dx=h[0], dy=h[1], dz=h[2], ng=ng)
    F = periodic_state(interior, ng, seed=sum(interior))
    dt = 0.25 * min(h)
    p = PhysicsParams()
    evo = BSSNEvolution(grid, p, order=order, ko_sigma=0.1, bc="periodic", integrator=integrator)
    got = to_host(evo.evolve(to_dev(F), dt, nsteps, t0=0.1).data)
    S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, 0.1)
    ref = O.evolve(F, S16, dt, nsteps, t0=0.1, integrator=integrator, order=order,
                   ko_sigma=0.1, ng=ng)
    ref = O.periodic_sync(ref, ng)
    err = normwise(got, ref)
    assert err.max() < TOL, err


def test_graph_evolve_matches_oracle():
    """7 RK4 steps as 2 replays of a captured 3-step graph + 1 device-time step, against the
    oracle (the SSL factor comes from the device's exp: same 1e-12 bar)."""
    say("graph evolve 7 steps (chunk 3), periodic, puncture 20^3")
    grid, st, _ = idata.config_single_puncture(seed=6, N=20)
    F = st.numpy()
    dt = 0.25 * grid.dx
    p = PhysicsParams()
    evo = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="periodic")
    got = to_host(evo.evolve(to_dev(F), dt, 7, t0=0.3, graph_chunk=3).data)
    S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, 0.3)
    ref = O.periodic_sync(O.evolve(F, S16, dt, 7, t0=0.3, integrator="rk4", order=4,
                                   ko_sigma=0.1, ng=3), 3)
    err = normwise(got, ref)
    assert err.max() < TOL, err


def test_graph_replays_with_eager_work_between():
    """A captured 2-step graph replayed 3 times, eager steps of a copy between the replays;
    every replay must agree with the eager result."""
    say("graph replay x3 vs eager, puncture 16^3")
    grid, st, _ = idata.config_single_puncture(seed=7, N=16, device=DEV)
    dt = 0.25 * grid.dx
    evo = BSSNEvolution(grid, PhysicsParams(), order=4, ko_sigma=0.1, bc="periodic")
    U = st.clone()
    V = st.clone()
    tdev = torch.full((1,), 0.1, dtype=torch.float64, device=DEV)
    g = evo.capture_steps(U, tdev, dt, 2)
    t = 0.1
    for r in range(3):
        g.replay()
        torch.cuda.synchronize()
        for _ in range(2):
            evo.step_(V, t, dt)
            t += dt
        err = normwise(to_host(U), to_host(V))
        assert err.max() < TOL, (r, err.max())
    assert abs(float(tdev.item()) - t) < 1e-14


def test_step_edge_padded():
    say("rk4 step with bc='none' (edge-padded stencils, all padded points advanced)")
    grid, st, _ = idata.config_single_puncture(seed=5, N=12)
    F = st.numpy()
    dt = 0.25 * grid.dx
    p = PhysicsParams()
    evo = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="none")
    got = to_host(evo.step(to_dev(F), 0.0, dt).data)
    S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, 0.0)
    ref = O.rk4_step(F, S16, 0.0, dt, order=4, ko_sigma=0.1, ng=3, periodic=False)
    assert normwise(got, ref).max() < TOL


def test_order6_step():
    say("rk4 step, order 6 with KO 8, ng 4, 16^3 single puncture")
    grid, st, _ = idata.config_single_puncture(seed=3, N=16, ng=4)
    F = st.numpy()
    dt = 0.25 * grid.dx
    p = PhysicsParams()
    evo = BSSNEvolution(grid, p, order=6, ko_sigma=0.1, bc="periodic")
    got = to_host(evo.step(to_dev(F), 0.0, dt).data)
    ref = O.periodic_sync(O.rk4_step(F, scalars(p, grid.dx, grid.dy, grid.dz, dt, 0.0), 0.0, dt,
                                     order=6, ko_sigma=0.1, ng=4), 4)
    assert normwise(got, ref).max() < TOL


def test_order8_step():
    say("rk4 step, order 8 with KO 8, ng 4, 16^3 single puncture")
    grid, st, _ = idata.config_single_puncture(seed=4, N=16, ng=4)
    F = st.numpy()
    dt = 0.25 * grid.dx
    p = PhysicsParams()
    evo = BSSNEvolution(grid, p, order=8, ko_sigma=0.1, bc="periodic")
    got = to_host(evo.step(to_dev(F), 0.0, dt).data)
    ref = O.periodic_sync(O.rk4_step(F, scalars(p, grid.dx, grid.dy, grid.dz, dt, 0.0), 0.0, dt,
                                     order=8, ko_sigma=0.1, ng=4), 4)
    assert normwise(got, ref).max() < TOL


def test_order8_constraints():
    say("constraints and interior norms, order 8, 20^3 single puncture, ng 4")
    grid, st, _ = idata.config_single_puncture(seed=7, N=20, ng=4)
    F = st.numpy()
    cs = ConstraintSolver(grid, order=8)
    got = cs.constraints(to_dev(F))
    HM = np.stack([to_host(got[k]) for k in ("H", "M0", "M1", "M2")])
    h3 = [grid.dx, grid.dy, grid.dz]
    assert normwise(HM, O.constraints(F, h3, 8)).max() < TOL
    H, M = cs.l2(to_dev(F))
    Hr, Mr = O.constraint_l2(F, h3, 8, grid.ng)
    assert abs(H - Hr) <= 1e-12 * Hr and abs(M - Mr) <= 1e-12 * Mr


def test_sync_and_enforce_kernels():
    say("periodic sync (bit-exact) and enforce kernels")
    rng = np.random.default_rng(9)
    F = smooth_state((11, 12, 13), seed=9) + 0.01 * rng.normal(size=(24, 11, 12, 13))
    d = to_dev(F)
    lib = _lib.load()
    _lib.check(lib.bssn_periodic_sync_f64(d.data_ptr(), 11, 12, 13, 3, stream_handle(d.device)))
    np.testing.assert_array_equal(to_host(d), O.periodic_sync(F, 3))
    d = to_dev(F)
    _lib.check(lib.bssn_enforce_f64(d.data_ptr(), 11 * 12 * 13, 1e-4, 1e-4, stream_handle(d.device)))
    assert normwise(to_host(d), O.enforce(F)).max() < TOL


def test_constraints_match_oracle():
    say("constraints and interior norms, 24^3 single puncture, order 4")
    grid, st, _ = idata.config_single_puncture(seed=6, N=24)
    F = st.numpy()
    cs = ConstraintSolver(grid, order=4)
    got = cs.constraints(to_dev(F))
    HM = np.stack([to_host(got[k]) for k in ("H", "M0", "M1", "M2")])
    h3 = [grid.dx, grid.dy, grid.dz]
    assert normwise(HM, O.constraints(F, h3, 4)).max() < TOL
    H, M = cs.l2(to_dev(F))
    Hr, Mr = O.constraint_l2(F, h3, 4, grid.ng)
    assert abs(H - Hr) <= 1e-12 * Hr and abs(M - Mr) <= 1e-12 * Mr


def test_solver_interface():
    say("BSSNSolver.rhs_dict fused == twopass, device_rhs_fused")
    grid, st, _ = idata.config_single_puncture(seed=8, N=16)
    F = to_dev(st.numpy())
    a = BSSNSolver(grid, order=4, scheme="cuda_fused").rhs(F, t=0.3).data
    b = BSSNSolver(grid, order=4, scheme="verbatim").rhs(F, t=0.3).data
    assert normwise(to_host(a), to_host(b)).max() < TOL
    from bssn3d_hip.rhs import device_rhs_fused
    d = device_rhs_fused(F, PhysicsParams(), grid.dx, grid.dy, grid.dz, 0.25 * grid.dx, 0.3,
                         order=4)
    np.testing.assert_array_equal(to_host(d["alpha"]), to_host(a[0]))


def test_full_size_binary_sampled():
    """The 256^3 binary at full size: sampled points against the oracle (random points, the
    7^3 neighbourhoods of both punctures where the fields vary most, edges, faces and corners),
    normalised by each field's full-grid maximum; Minkowski staticity and bitwise determinism."""
    say("256^3 binary workload: sampled parity + properties")
    grid, st, pos = idata.config_binary(seed=0, N=256, device=DEV)
    S16 = scalars(PhysicsParams(), grid.dx, grid.dy, grid.dz, 0.25 * grid.dx, 0.0)
    out = torch.empty_like(st)
    rhs_into(st, out, S16, 4)
    out2 = torch.empty_like(st)
    rhs_into(st, out2, S16, 4)
    torch.cuda.synchronize()
    assert torch.equal(out, out2)
    del out2
    fullmax = out.abs().reshape(24, -1).max(dim=1).values.cpu().numpy()
    Sx, Sy, Sz = grid.shape
    rng = np.random.default_rng(1)
    pts = [np.stack([rng.integers(0, n, size=3000) for n in (Sx, Sy, Sz)])]
    pts.append(np.array([[x, y, z] for x in (0, 1, Sx - 2, Sx - 1) for y in (0, 2, Sy - 1)
                         for z in (0, 3, Sz - 1)]).T)
    for p in pos:
        c = [int(round((p[a] - lo) / grid.dx)) + grid.ng
             for a, lo in enumerate((grid.xmin, grid.ymin, grid.zmin))]
        r = np.arange(-3, 4)
        X, Y, Z = np.meshgrid(c[0] + r, c[1] + r, c[2] + r, indexing="ij")
        pts.append(np.stack([X.ravel(), Y.ravel(), Z.ravel()]))
    P = np.concatenate(pts, axis=1)
    idx = (P[0] * Sy + P[1]) * Sz + P[2]
    F = to_host(st)
    ref = O.rhs_points(F, S16, idx, 4)
    truth, _ = O.rhs_points_ld(F, S16, idx, 4)
    got = to_host(out).reshape(24, -1)[:, idx]
    # Near the punctures (chi ~ 1e-4) the RHS is a difference of terms ~1e8 times larger than
    # its value, so FP64 itself is good to only ~1e-11 there: the reference algorithm in FP64
    # (the oracle) misses the long-double evaluation by e_ref. The HIP kernel must be as
    # accurate as that: e_hip <= max(1e-12, 2 e_ref), normwise by each field's full-grid max.
    e_ref = normwise(ref, truth, den=fullmax)
    e_hip = normwise(got, truth, den=fullmax)
    print("e_ref", np.array2string(e_ref, precision=2), "\ne_hip",
          np.array2string(e_hip, precision=2), flush=True)
    assert np.all(e_hip <= np.maximum(TOL, 2.0 * e_ref)), (e_hip, e_ref)
    del F
    mink = idata.minkowski(grid, device=DEV)
    rhs_into(mink, out, S16, 4, 0.1)
    torch.cuda.synchronize()
    assert float(out.abs().max()) < 1e-12


def test_rk4_stage_times_with_fast_ssl():
    """RK4 step with sig_ssl = 0.4 (the SSL factor changes ~20 % within the step, so each
    stage's time matters): the 20^3 puncture against the oracle (1e-12), and a uniform lapse
    state against the closed-form solution of its slice-locking ODE (oracle-pinned bound)."""
    say("RK4 step with a fast SSL factor: stage times")
    from test_oracle import _uniform_lapse_state, ssl_exact
    p = PhysicsParams(ssl_sigma=0.4)
    grid, st, _ = idata.config_single_puncture(seed=9, N=20)
    F = st.numpy()
    dt, t0 = 0.25 * grid.dx, 0.35
    evo = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="periodic")
    got = to_host(evo.step(to_dev(F), t0, dt).data)
    S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, t0)
    ref = O.periodic_sync(O.rk4_step(F, S16, t0, dt, order=4, ko_sigma=0.1, ng=3), 3)
    # error per field relative to the step's increment (a wrong stage time moves alpha's
    # increment by percents); the bar is floored at four ulps of the field's magnitude over its
    # increment, the resolution with which a stored O(1) value (gt_xx ~ 1, increment ~4e-4)
    # can carry its increment at all
    inc = np.abs(ref - F).reshape(24, -1).max(axis=1)
    mag = np.abs(ref).reshape(24, -1).max(axis=1)
    bar = np.maximum(TOL, 4.0 * np.finfo(np.float64).eps * mag / inc)
    e = normwise(got, ref, den=inc)
    assert np.all(e < bar), (e, bar)
    assert e[0] < TOL                  # alpha: its increment is O(1e-1), far above the floor
    a0, dtu = 1.3, 0.02
    g2 = Grid.from_domain(8, order=4, ko_order=6)
    U = _uniform_lapse_state(a0, g2.shape)
    evo2 = BSSNEvolution(g2, p, order=4, ko_sigma=0.1, bc="periodic")
    out = to_host(evo2.step(to_dev(U), t0, dtu).data)
    exact = ssl_exact(a0, t0, t0 + dtu, p.ssl_h, p.ssl_sigma)
    assert abs(out[0, 7, 7, 7] - exact) / (a0 - 1.0) < 1e-9
