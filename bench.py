"""Benchmark: the fused BSSN RHS + RK4 step at 256^3 FP64 on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu] [--no-twopass]
                    [--no-int8] [--no-batch] [--no-longrun] [--no-amr] [--no-bfp]

A step is one full RK4 step (four fused stage kernels: RHS + Kreiss-Oliger + stage update,
enforcement folded into the last; then the periodic ghost refill) of the 256^3-interior,
h = 1/16 equal-mass Brill-Lindquist binary (m = 0.5 each, 6M apart, 1e-4 seeded sinusoid
perturbations on every field), FD order 4 / KO order 6, sigma_KO = 0.1. Each step feeds the next,
so no input is reused from cache. ``value`` is interior points per second of the whole step
(Mpts/s). Also reported: the bare fused RHS and the materialised two-pass RHS (138 derivative
arrays through HBM, then the algebra) on the same state, the roofline position of the fused
stage kernel (timed live with HIP events on the launch stream), and the CPU oracle timed on
this host's cores on a bounded 48^3 sample.

Secondary forms (extra keys, same metric family):
  int8_split  the RHS with its 138 derivatives from 7 int8 digits on the INT8 matrix cores
              (Ozaki scheme I) and the FP64 algebra, on the single-puncture construction at
              256^3 (h = 12/256), with its normwise error against the FP64 fused RHS, and the
              RK4 step composed from it (per stage: sync, bundle, algebra, KO; full-array
              updates);
  batched     640 uneven blocks (extents 8..32 per axis) on three nested levels (h = 1/4,
              1/8, 1/16) around the binary 0.909 / 0.0909 4M apart: the RHS of all blocks in
              one launch and the hierarchy RK4 step (one batched launch per stage plus the
              ghost fill program: periodic root, same-level copies, parent prolongation),
              launch shape autotuned, eager and replayed from a captured graph; Mpts/s over
              interior points;
  longrun     the tracking gate (fused and INT8-split schemes against the two-pass scheme on
              the reference's three N=24 configurations) with its verdicts and wall time, and
              a 128^3 long run with the reference's carried ghost values; the carried-ghost
              RK4 step at 256^3 is reported beside the default one (carried_ghost_step_ms),
              and so is the RK4 step with the radiative outer boundary in place of periodic
              ghosts (sommerfeld_step_ms);
  amr         a live three-level block hierarchy (16^3 blocks) around the binary: RK4 step,
              ghost fill program, a refining regrid pass;
  bfp         the two-pass RHS with the bundle stored at 32 and 48 mantissa bits per value
              (block floating point, one exponent per derivative array) against the fp64
              bundle: time, stored bytes, error.

Rank 0 prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "integer-schemes_amd"))

METRIC = "Mpts/s BSSN RHS+RK4 at 256^3 FP64 vs INT8-split/batched; % HBM & FP64 peak"
PEAK_HBM = 8.0e12          # B/s, MI355X HBM3E
PEAK_FP64 = 78.6e12        # flop/s, MI355X FP64 vector (= FP64 matrix) peak

# Algorithmic work per interior point (order 4 FD, KO order 6), counted from the operators:
#   stencils: 72 first derivatives x 8 flops (4 taps, multiply + add), 33 diagonal second
#   derivatives x 10 (5 taps), 33 mixed x 36 (4 x 4 nested taps)              = 2094
#   algebra: the reference CSE's operation count (2486 mul + 1259 add + 482 sub + 52 div
#   + 30 for the integer powers, sqrt, exp; vendor/bssneqs_SSL_HD_dxsq.cpp)     = 4309
#   KO: 24 fields x 3 axes x 7 taps x 2 + 2 adds per field                         = 1056
#   stage update: 24 fields x ~4 flops                                            =   96
FLOP_RHS = 2094 + 4309
FLOP_STAGE = FLOP_RHS + 1056 + 96
# HBM bytes per interior point per stage launch (24 fields x 8 B per array touched):
#   stage 1: read y, write Y and S (576); stages 2, 3: read Y, y, S, write Y', S (960);
#   stage 4: read Y, y, S, write y (768). RHS alone: read 24 + write 24 fields = 384.
BYTES_STAGES = (576, 960, 960, 768)
BYTES_RHS = 384


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seconds: float = 12.0):
    """The oracle (oracle/bssn_oracle.c, its -O3 AVX2/FMA timing build liboracle_perf.so) on this
    host's cores, bounded samples of the workload:
      value  RK4 steps of the 48^3 single-puncture workload (FD4/KO6) on all host cores, repeated
             for about ``seconds``;
      runs   the same step on one core, the RHS (+KO) on one core and on all cores at 48^3, and
             one RHS evaluation of the 256^3 binary (262^3 padded) on all cores.
    Mpts/s counts interior points for steps and padded points for RHS evaluations."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    from bssn3d_hip import initial_data as idata
    from bssn3d_hip.state import PhysicsParams, scalars

    O.use_build("perf")
    ncores = O.num_threads()
    grid, st, _ = idata.config_single_puncture(seed=0, N=48)
    F = st.numpy()
    dt = 0.25 * grid.dx
    p = PhysicsParams()
    S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, 0.0)
    pad3 = grid.shape[0] ** 3

    def time_loop(fn, budget, min_reps=1):
        fn()                                   # warm (page-in, thread pool)
        n, t0 = 0, time.perf_counter()
        while True:
            fn()
            n += 1
            el = time.perf_counter() - t0
            if el >= budget and n >= min_reps:
                return n, el

    def step():
        return O.rk4_step(F, S16, 0.0, dt, order=4, ko_sigma=0.1, ng=3)

    def rhs():
        return O.rhs(F, S16, 4, ko_sigma=0.1)

    runs = []
    O.set_threads(ncores)
    n, el = time_loop(step, seconds)
    value = 48 ** 3 * n / el / 1e6
    sample = (f"{n} RK4 steps of the 48^3 single-puncture workload (54^3 padded), FD4/KO6, "
              f"{el:.1f} s")
    n, el = time_loop(rhs, 2.0)
    runs.append({"what": "rhs+ko 48^3", "cores": ncores, "mpts_s": round(pad3 * n / el / 1e6, 4),
                 "reps": n, "s": round(el, 2)})
    O.set_threads(1)
    n, el = time_loop(step, 3.0)
    runs.append({"what": "rk4 step 48^3", "cores": 1,
                 "mpts_s": round(48 ** 3 * n / el / 1e6, 4), "reps": n, "s": round(el, 2)})
    n, el = time_loop(rhs, 2.0)
    runs.append({"what": "rhs+ko 48^3", "cores": 1, "mpts_s": round(pad3 * n / el / 1e6, 4),
                 "reps": n, "s": round(el, 2)})
    O.set_threads(ncores)
    del F, st
    g2, st2, _ = idata.config_binary(seed=0, N=256)
    F2 = st2.numpy()
    del st2
    S2 = scalars(p, g2.dx, g2.dy, g2.dz, 0.25 * g2.dx, 0.0)
    t0 = time.perf_counter()
    O.rhs(F2, S2, 4, ko_sigma=0.1)
    el = time.perf_counter() - t0
    runs.append({"what": "rhs+ko 256^3 binary (262^3 padded)", "cores": ncores,
                 "mpts_s": round(F2[0].size / el / 1e6, 4), "reps": 1, "s": round(el, 2)})
    del F2
    O.use_build("check")
    return {"value": round(value, 4), "unit": "Mpts/s (RK4 step)", "cores": ncores,
            "kind": "port", "sample": sample, "build": "oracle/liboracle_perf.so (-O3 "
            "-march=x86-64-v3, FMA contraction on)", "cpu_model": _cpu_model(), "runs": runs}


def schedule() -> dict:
    """The stage kernels' schedule knobs the library was built with: the autotuner's choice
    (integer-schemes_amd/tuning.json) or the built-in defaults."""
    path = ROOT / "integer-schemes_amd" / "tuning.json"
    if not path.exists():
        return {"defines": [], "source": "built-in defaults"}
    t = json.loads(path.read_text())
    return {"defines": t.get("defines", []), "source": "integer-schemes_amd/tuning.json",
            "tuned_on": t.get("device"), "tuned_step_ms": t.get("candidates", {}).get(
                t.get("best", ""), {}).get("step_ms")}


def measured_traffic(workload: str):
    """HBM bytes per stage launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    (each scaled by the same counter on a 1 GiB elementwise copy), or None for another
    workload."""
    path = ROOT / "profiles" / "rk4_stage_256_fd4_traffic.json"
    if not path.exists():
        return None
    t = json.loads(path.read_text())
    if t.get("workload") != workload:
        return None
    return t["read_bytes_per_launch"] + t["write_bytes_per_launch"]


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def bench_int8(N: int, reps: int) -> dict:
    """INT8-split RHS (bundle from 7 int8 digits + FP64 algebra) vs the fused FP64 RHS."""
    import numpy as np
    import torch
    from bssn3d_hip import _lib
    from bssn3d_hip import initial_data as idata
    from bssn3d_hip.rhs import rhs_int8_into, rhs_into, stream_handle
    from bssn3d_hip.state import PhysicsParams, scalars

    dev = torch.device("cuda:0")
    grid, U, _ = idata.config_single_puncture(seed=0, N=N, device=dev)
    S16 = scalars(PhysicsParams(), grid.dx, grid.dy, grid.dz, 0.25 * grid.dx, 0.0)
    Sx, Sy, Sz = grid.shape
    out8 = torch.empty_like(U)
    ref = torch.empty_like(U)
    D = torch.empty((138, Sx, Sy, Sz), dtype=torch.float64, device=dev)
    rhs_into(U, ref, S16, 4)
    rhs_int8_into(U, out8, D, S16, 4)
    torch.cuda.synchronize()
    a, b = _events()
    a.record()
    for _ in range(reps):
        rhs_int8_into(U, out8, D, S16, 4)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    h3 = _lib.dbl([grid.dx, grid.dy, grid.dz])
    lib = _lib.load()
    a.record()
    for _ in range(reps):
        _lib.check(lib.bssn_deriv_bundle_i8_f64(U.data_ptr(), D.data_ptr(), Sx, Sy, Sz, 4, h3, 7,
                                                stream_handle(dev)))
    b.record()
    torch.cuda.synchronize()
    ms_bundle = a.elapsed_time(b) / reps
    err = []
    for f in range(24):
        den = max(float(ref[f].abs().max()), 1e-12)
        err.append(float((out8[f] - ref[f]).abs().max()) / den)
    # per derivative: the INT8 bundle (still in D) against the FP64 bundle
    D64 = torch.empty_like(D)
    _lib.check(lib.bssn_deriv_bundle_f64(U.data_ptr(), D64.data_ptr(), Sx, Sy, Sz, 4, h3,
                                         stream_handle(dev)))
    _lib.check(lib.bssn_deriv_bundle_i8_f64(U.data_ptr(), D.data_ptr(), Sx, Sy, Sz, 4, h3, 7,
                                            stream_handle(dev)))
    derr = []
    for i in range(138):
        den = max(float(D64[i].abs().max()), 1e-12)
        derr.append(float((D[i] - D64[i]).abs().max()) / den)
    npts = N ** 3
    del D, D64, out8, ref
    torch.cuda.empty_cache()
    # the RK4 step in this form: per stage sync + INT8 bundle + FP64 algebra + KO, full-array
    # RK4 updates, enforce, sync (the reference's composition)
    from bssn3d_hip.evolve import BSSNEvolution
    evo = BSSNEvolution(grid, PhysicsParams(), order=4, ko_sigma=0.1, bc="periodic",
                        scheme="int8")
    dt = 0.25 * grid.dx
    evo.step_(U, 0.0, dt)
    torch.cuda.synchronize()
    nstep = max(2, reps // 3)
    a.record()
    for i in range(nstep):
        evo.step_(U, (i + 1) * dt, dt)
    b.record()
    torch.cuda.synchronize()
    ms_step = a.elapsed_time(b) / nstep
    assert bool(torch.isfinite(U).all()), "non-finite INT8-split state"
    del evo, U
    return {"workload": f"single_puncture_{N}^3_h{12.0 / N:g}", "slices": 7,
            "rhs_ms": round(ms, 4), "rhs_mpts_s": round(npts / (ms * 1e-3) / 1e6, 2),
            "step_ms": round(ms_step, 4), "step_mpts_s": round(npts / (ms_step * 1e-3) / 1e6, 2),
            "steps_timed": nstep,
            "bundle_ms": round(ms_bundle, 4),
            "bundle_hbm_frac": round(138 * 8 * Sx * Sy * Sz / (ms_bundle * 1e-3) / PEAK_HBM, 4),
            "max_normwise_err_vs_fp64_rhs": float("%.3e" % max(err)),
            "median_normwise_err_vs_fp64_rhs": float("%.3e" % float(np.median(err))),
            "max_normwise_err_vs_fp64_derivative": float("%.3e" % max(derr)),
            "median_normwise_err_vs_fp64_derivative": float("%.3e" % float(np.median(derr))),
            "worst_derivative_index": int(np.argmax(derr))}


def bench_batched(reps: int) -> dict:
    """The batched block-AMR workload: 640 uneven blocks (interior extents per axis from
    {8, 12, 16, 24, 32}) on three nested levels (h = 1/4, 1/8, 1/16) around the 1:10 binary
    (m = 0.909, 0.0909, 4M apart); per RK stage one batched launch over every block, then the
    ghost fill program (root periodic, same-level copies, prolongation from the parent level).
    The launch shape is autotuned; steps are timed eagerly and replayed from a captured graph,
    and again as linked steps (stages read copied ghosts at their sources)."""
    import torch
    from bssn3d_hip.batch import timed
    from bssn3d_hip.patches import config_hierarchy, fill_binary
    from bssn3d_hip.state import PhysicsParams

    dev = torch.device("cuda:0")
    w0 = time.perf_counter()
    H, _ = config_hierarchy(seed=0, device=dev)
    prog = H.program
    build_s = time.perf_counter() - w0
    U0 = fill_binary(H, H.empty_pool())
    p = PhysicsParams()
    dt = 0.25 * H.boxes[-1].h
    kw = dict(params=p, order=4, ko_sigma=0.1)
    tune = H.tune_launch(U0, dt, candidates=[0, 1024, 2048, 4096], reps=2, **kw)
    batch = H.batch
    prog_shape = batch.persistent_waves
    U = U0.clone()
    OUT = H.empty_pool()
    batch.rhs_into(U, OUT, p, dt, 0.0, 4, 0.0)
    torch.cuda.synchronize()
    ms_rhs = timed(lambda: batch.rhs_into(U, OUT, p, dt, 0.0, 4, 0.0), reps, dev)
    # what the batched launch replaces: the single-grid RHS (bssn_rhs_f64) block by block, one
    # launch per block, eagerly and with the 640 launches captured in one graph; every block's
    # result must equal the batched one bit for bit
    from bssn3d_hip.rhs import rhs_into
    from bssn3d_hip.state import scalars
    OUT2 = H.empty_pool()
    per = [(H.view(U, b), H.view(OUT2, b), scalars(p, *H.blocks[b].h, dt, 0.0))
           for b in range(len(H))]

    def per_block():
        for F, O, s16 in per:
            rhs_into(F, O, s16, 4, 0.0)
    per_block()
    torch.cuda.synchronize()
    assert all(torch.equal(H.view(OUT, b), H.view(OUT2, b)) for b in range(len(H))), \
        "per-block RHS differs from the batched RHS"
    ms_per_block = timed(per_block, reps, dev)
    gpb = torch.cuda.CUDAGraph()
    with torch.cuda.graph(gpb):
        per_block()
    gpb.replay()
    torch.cuda.synchronize()
    ms_per_block_graph = timed(gpb.replay, reps, dev)
    del gpb, per, OUT, OUT2
    ms_sync = timed(lambda: H.sync_(U), reps, dev)
    t = [0.0]

    def step():
        H.step_(U, t[0], dt, **kw)
        t[0] += dt
    step()
    torch.cuda.synchronize()
    U.copy_(U0)
    ms_step = timed(step, reps, dev)
    assert bool(torch.isfinite(U).all()), "non-finite hierarchy state"
    U.copy_(U0)
    tdev = torch.full((1,), 0.0, dtype=torch.float64, device=dev)
    g = H.capture_steps(U, tdev, dt, 1, **kw)
    g.replay()
    torch.cuda.synchronize()
    U.copy_(U0)
    ms_graph = timed(g.replay, reps, dev)
    assert bool(torch.isfinite(U).all()), "non-finite hierarchy state"
    # linked steps: copied ghosts read at their sources, only the prolongations (and the copies
    # their taps need) between stages; sync_end=False leaves the other ghosts to an explicit sync
    linked = H.linked_program
    H.link
    lk = dict(kw, ghosts="linked")
    # the linked step's stages read ghosts at their sources: tune its launch shape on its own
    tune_link = H.tune_launch(U0, dt, candidates=[0, 1024, 2048, 4096], reps=2,
                              sync_end=False, **lk)
    U.copy_(U0)
    H.step_(U, 0.0, dt, **lk)
    torch.cuda.synchronize()
    U.copy_(U0)
    ms_link_synced = timed(lambda: H.step_(U, 0.0, dt, **lk), reps, dev)
    U.copy_(U0)
    ms_link = timed(lambda: H.step_(U, 0.0, dt, sync_end=False, **lk), reps, dev)
    assert bool(torch.isfinite(U).all()), "non-finite hierarchy state"
    U.copy_(U0)
    tdev.fill_(0.0)
    g = H.capture_steps(U, tdev, dt, 1, sync_end=False, **lk)
    g.replay()
    torch.cuda.synchronize()
    U.copy_(U0)
    ms_link_graph = timed(g.replay, reps, dev)
    pts = batch.interior_points
    return {"workload": "640_blocks_3_nested_levels_h1/4_1/8_1/16_sizes_8-32_masses_0.909_0.0909"
                        "_sep4",
            "blocks": len(H), "blocks_per_level": [b.nblocks for b in H.boxes],
            "interior_points": pts, "padded_points": batch.padded_points,
            "ghost_program": {"copies": prog.n_copy, "prolongations": prog.n_prolong,
                              "fill_ms": round(ms_sync, 4)},
            "host_build_s": round(build_s, 2),
            "launch_persistent_waves": prog_shape,
            "autotune_step_ms": {str(k): round(v, 4) for k, v in tune.items()},
            "rhs_ms": round(ms_rhs, 4), "rhs_mpts_s": round(pts / (ms_rhs * 1e-3) / 1e6, 2),
            "per_block_rhs_ms": round(ms_per_block, 4),
            "per_block_rhs_graph_ms": round(ms_per_block_graph, 4),
            "batched_vs_per_block_rhs": round(ms_per_block / ms_rhs, 3),
            "batched_vs_per_block_rhs_graph": round(ms_per_block_graph / ms_rhs, 3),
            "steps_timed": reps,
            "step_ms": round(ms_step, 4), "step_mpts_s": round(pts / (ms_step * 1e-3) / 1e6, 2),
            "graph_step_ms": round(ms_graph, 4),
            "graph_step_mpts_s": round(pts / (ms_graph * 1e-3) / 1e6, 2),
            "linked": {"program_copies": linked.n_copy, "prolongations": linked.n_prolong,
                       "launch_persistent_waves": batch.persistent_waves,
                       "autotune_step_ms": {str(k): round(v, 4) for k, v in tune_link.items()},
                       "step_ms": round(ms_link, 4),
                       "step_mpts_s": round(pts / (ms_link * 1e-3) / 1e6, 2),
                       "graph_step_ms": round(ms_link_graph, 4),
                       "graph_step_mpts_s": round(pts / (ms_link_graph * 1e-3) / 1e6, 2),
                       "synced_step_ms": round(ms_link_synced, 4),
                       "synced_step_mpts_s": round(pts / (ms_link_synced * 1e-3) / 1e6, 2)}}


def bench_longrun() -> dict:
    """The tracking gate on the reference's suite (robust / gauge / hambump, N=24, 2 crossing
    times, 8 samples each; fused against the two-pass scheme at FD order 6 and the INT8-split
    scheme at FD order 4), and a long run at 128^3 (robust noise 1e-8, 1 crossing time = 512
    RK4 steps, 4 samples) with the reference's carried ghosts."""
    import torch
    from bssn3d_hip import tracking_gate as tg
    from bssn3d_hip.series import SeriesSpec, record

    out = {}
    for name, scheme, order in (("gate_fused_vs_twopass_fd6", "fused", 6),
                                ("gate_int8_vs_twopass_fd4", "int8", 4)):
        w0 = time.perf_counter()
        rep = tg.run_gate(scheme, order=order, device="cuda:0")
        torch.cuda.synchronize()
        out[name] = {"pass": bool(rep.passed), "wall_s": round(time.perf_counter() - w0, 3),
                     "worst_dev": float("%.3e" % rep.worst)}
    N = 128
    record(SeriesSpec("robust", N, 0.05, 1, 0.1, 1e-8), device="cuda:0")
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    s = record(SeriesSpec("robust", N, 1.0, 4, 0.1, 1e-8), device="cuda:0")
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    steps = int(s.step[-1])
    out["longrun_robust_128^3"] = {
        "steps": steps, "samples": len(s), "wall_s": round(wall, 3),
        "steps_per_s": round(steps / wall, 2), "mpts_s": round(steps * N ** 3 / wall / 1e6, 2),
        "final_H": float("%.4e" % s.H[-1]), "finite": s.all_finite}
    return out


def bench_amr(reps: int) -> dict:
    """A live three-level hierarchy: 4x4x4 root blocks of 16^3 nodes on [-8, 8)^3 (h = 1/4,
    1/8, 1/16), Brill-Lindquist binary (0.5 + 0.5, 6 apart; no perturbations), refined by two
    regrid passes on the chi indicator; shared-dt RK4 steps with the ghost program after every
    stage (root periodic, parent prolongation, face copies), the program alone, and one regrid
    pass (indicator, flags, transfers, new plan and program)."""
    import torch
    from bssn3d_hip import initial_data as idata
    from bssn3d_hip.amr import AMRHierarchy, RegridPolicy
    from bssn3d_hip.batch import timed
    from bssn3d_hip.state import CHI, PhysicsParams

    dev = torch.device("cuda:0")
    H = AMRHierarchy(nb_root=(4, 4, 4), BS=16, ng=3, levels=3, L=16.0, lo=(-8.0,) * 3,
                     device=dev)
    pos = [(-3.0 + 0.01, 0.013, 0.007), (3.0 + 0.01, 0.013, 0.007)]
    fn = lambda g: idata.punctures(g, [0.5, 0.5], pos, amp=0.0, device=dev)
    U = H.fill(H.empty_pool(), fn)
    pol = RegridPolicy(refine_above=0.02, coarsen_below=1e-6, patience=1, buffer=0)
    for _ in range(2):
        U = H.regrid(U, pol, fields=[CHI])
    counts = [H.tree.count(L) for L in range(3)]
    p = PhysicsParams()
    dt = 0.25 * H.spacing(2)
    t = [0.0]

    def step():
        H.step_(U, t[0], dt, p, 4, 0.1)
        t[0] += dt
    kw = dict(params=p, order=4, ko_sigma=0.1)
    tune = H.tune_launch(U, dt, candidates=[0, 1024, 2048], reps=3, **kw)
    prog_shape = H.batch.persistent_waves
    step()
    torch.cuda.synchronize()
    ms_step = timed(step, reps, dev)
    assert bool(torch.isfinite(U).all()), "non-finite hierarchy state"
    ms_sync = timed(lambda: H.sync_(U), reps, dev)
    # linked steps: face-copied ghosts read at their sources, only the prolongations (and the
    # copies their taps read) between stages; launch shape tuned on its own
    H.linked_program, H.link
    lk = dict(kw, ghosts="linked", sync_end=False)
    tune_link = H.tune_launch(U, dt, candidates=[0, 1024, 2048], reps=3, **lk)
    H.step_(U, 0.0, dt, **lk)
    torch.cuda.synchronize()
    ms_link = timed(lambda: H.step_(U, 0.0, dt, **lk), reps, dev)
    assert bool(torch.isfinite(U).all()), "non-finite hierarchy state"
    H.sync_(U)
    pts = sum(counts) * 16 ** 3
    prog = H.program
    # a regrid pass that changes nothing costs the indicator and the decisions; time one that
    # refines (a fresh hierarchy), including the new plan and program
    torch.cuda.synchronize()
    H2 = AMRHierarchy(nb_root=(4, 4, 4), BS=16, ng=3, levels=3, L=16.0, lo=(-8.0,) * 3,
                      device=dev)
    V = H2.fill(H2.empty_pool(), fn)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    V = H2.regrid(V, pol, fields=[CHI])
    H2.batch, H2.program
    torch.cuda.synchronize()
    ms_regrid = (time.perf_counter() - w0) * 1e3
    del V, H2
    return {"workload": "bl_binary_amr_3lev_root4x4x4_bs16_box16",
            "blocks_per_level": counts, "interior_points": pts,
            "ghost_program": {"prolongations": prog.n_prolong, "copies": prog.n_copy},
            "launch_persistent_waves": prog_shape,
            "autotune_step_ms": {str(k): round(v, 4) for k, v in tune.items()},
            "steps_timed": reps, "step_ms": round(ms_step, 4),
            "step_mpts_s": round(pts / (ms_step * 1e-3) / 1e6, 2),
            "ghost_fill_ms": round(ms_sync, 4),
            "linked": {"copies": H.linked_program.n_copy,
                       "launch_persistent_waves": H.batch.persistent_waves,
                       "autotune_step_ms": {str(k): round(v, 4) for k, v in tune_link.items()},
                       "step_ms": round(ms_link, 4),
                       "step_mpts_s": round(pts / (ms_link * 1e-3) / 1e6, 2)},
            "regrid_refine_pass_ms": round(ms_regrid, 1)}


def bench_bfp(N: int, reps: int) -> dict:
    """The two-pass RHS with its bundle stored in block floating point (BFP32, BFP48) against
    the fp64 bundle, on the 256^3 binary: per evaluation a maxima pass, the quantizing bundle
    pass, the dequantizing algebra (and the 138 maxima to the host for the exponents); and with
    exponents fixed from an earlier evaluation (no maxima pass)."""
    import numpy as np
    import torch
    from bssn3d_hip import initial_data as idata
    from bssn3d_hip.rhs import bfp_buffers, rhs_bfp_into, rhs_twopass_into
    from bssn3d_hip.state import PhysicsParams, scalars

    dev = torch.device("cuda:0")
    grid, U, _ = idata.config_binary(seed=0, N=N, device=dev)
    S16 = scalars(PhysicsParams(), grid.dx, grid.dy, grid.dz, 0.25 * grid.dx, 0.0)
    Sx, Sy, Sz = grid.shape
    P = Sx * Sy * Sz
    ref = torch.empty_like(U)
    D = torch.empty((138, Sx, Sy, Sz), dtype=torch.float64, device=dev)
    rhs_twopass_into(U, ref, D, S16, 4)
    torch.cuda.synchronize()
    a, b = _events()
    a.record()
    for _ in range(reps):
        rhs_twopass_into(U, ref, D, S16, 4)
    b.record()
    torch.cuda.synchronize()
    out = {"workload": f"bssn_rhs_{N}^3_bl_binary_fd4", "fp64_bundle_bytes": 138 * 8 * P,
           "twopass_fp64_ms": round(a.elapsed_time(b) / reps, 4)}
    del D
    torch.cuda.empty_cache()
    o = torch.empty_like(U)
    for k in (32, 48):
        Q, amax = bfp_buffers(U, k)
        rhs_bfp_into(U, o, Q, amax, S16, 4, mant_bits=k)
        torch.cuda.synchronize()
        a.record()
        for _ in range(reps):
            rhs_bfp_into(U, o, Q, amax, S16, 4, mant_bits=k)
        b.record()
        torch.cuda.synchronize()
        I = (slice(None),) + (slice(3, -3),) * 3
        d = (o[I] - ref[I]).double()
        err = float(torch.sqrt(torch.mean(d * d)) / torch.sqrt(torch.mean(ref[I] * ref[I])))
        ms = a.elapsed_time(b) / reps
        e_fix = rhs_bfp_into(U, o, Q, amax, S16, 4, mant_bits=k)
        torch.cuda.synchronize()
        a.record()
        for _ in range(reps):
            rhs_bfp_into(U, o, Q, amax, S16, 4, mant_bits=k, exps=e_fix)
        b.record()
        torch.cuda.synchronize()
        out[f"bfp{k}"] = {"ms": round(ms, 4),
                          "fixed_exponents_ms": round(a.elapsed_time(b) / reps, 4),
                          "stored_bundle_bytes": int(Q.numel()),
                          "rel_l2_err_vs_fp64": float("%.3e" % err)}
        del Q, amax
        torch.cuda.empty_cache()
    return out


def bench_order(order: int, sizes, reps: int) -> dict:
    """FD order 6 or 8 with KO order 8 (4 ghosts): order 6 is the reference's production pairing
    and the configuration of its published fused-kernel figure (gen_fused.py:63-66 at 128^3);
    order 8 its highest. The same equal-mass binary construction (h = 16/N), RK4 step and bare
    RHS."""
    import torch
    from bssn3d_hip import initial_data as idata
    from bssn3d_hip.evolve import BSSNEvolution
    from bssn3d_hip.rhs import rhs_into
    from bssn3d_hip.state import PhysicsParams, scalars

    dev = torch.device("cuda:0")
    p = PhysicsParams()
    out = {}
    for N in sizes:
        grid, U, _ = idata.config_binary(seed=0, N=N, device=dev, ng=4)
        dt = 0.25 * grid.dx
        evo = BSSNEvolution(grid, p, order=order, ko_sigma=0.1, bc="periodic", integrator="rk4",
                            ko_order=8)
        t = 0.0
        for _ in range(2):
            evo.step_(U, t, dt)
            t += dt
        a, b = _events()
        torch.cuda.synchronize()
        a.record()
        for _ in range(reps):
            evo.step_(U, t, dt)
            t += dt
        b.record()
        torch.cuda.synchronize()
        ms_step = a.elapsed_time(b) / reps
        finite = bool(torch.isfinite(U).all())
        S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, t)
        o = torch.empty_like(U)
        rhs_into(U, o, S16, order)
        a.record()
        for _ in range(reps):
            rhs_into(U, o, S16, order)
        b.record()
        torch.cuda.synchronize()
        ms_rhs = a.elapsed_time(b) / reps
        n3 = N ** 3
        out[f"{N}^3"] = {"step_ms": round(ms_step, 4),
                         "step_mpts_s": round(n3 / (ms_step * 1e-3) / 1e6, 2),
                         "rhs_ms": round(ms_rhs, 4),
                         "rhs_mpts_s": round(grid.shape[0] ** 3 / (ms_rhs * 1e-3) / 1e6, 2),
                         "finite": finite}
        del U, o, evo
        torch.cuda.empty_cache()
    out["workload"] = f"bl_binary_fd{order}_ko8_ng4_h16/N"
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-twopass", action="store_true")
    ap.add_argument("--no-int8", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-longrun", action="store_true")
    ap.add_argument("--no-amr", action="store_true")
    ap.add_argument("--no-bfp", action="store_true")
    ap.add_argument("--no-fd6", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-gib", type=float, default=0.0,
                    help="also run one plain device copy of this many GiB (profiler calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the library has no multi-GPU path (sharding "
              "would be a z-slab decomposition with a ghost-plane exchange per RK stage, "
              "not built).", file=sys.stderr)
        return 2

    import torch
    from bssn3d_hip import initial_data as idata
    from bssn3d_hip import _lib
    from bssn3d_hip.constraints import ConstraintSolver
    from bssn3d_hip.evolve import BSSNEvolution
    from bssn3d_hip.rhs import rhs_into, rhs_twopass_into, stream_handle
    from bssn3d_hip.state import PhysicsParams, scalars

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    N = args.n
    grid, U, _ = idata.config_binary(seed=0, N=N, device=dev)
    p = PhysicsParams()
    dt = 0.25 * grid.dx
    evo = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="periodic", integrator="rk4")
    npts = N ** 3
    Sx, Sy, Sz = grid.shape
    lib = _lib.load()

    def ev():
        return torch.cuda.Event(enable_timing=True)

    # Hamiltonian / momentum constraint norms of the workload before and after the timed
    # steps (interior RMS, FD order 4 like the evolution; ConstraintSolver.l2 is the
    # reference's constraints.py:53-62)
    cons = ConstraintSolver(grid, order=4)
    h0, m0 = cons.l2(U)

    # ---------------------------------------------------------------- RK4 step (the metric)
    t = 0.0
    for _ in range(args.warmup):
        evo.step_(U, t, dt)
        t += dt
    torch.cuda.synchronize()
    e0, e1 = ev(), ev()
    w0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        evo.step_(U, t, dt)
        t += dt
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ms_step = e0.elapsed_time(e1) / args.steps
    assert bool(torch.isfinite(U).all()), "non-finite state after the timed steps"
    h1, m1 = cons.l2(U)
    nsteps_done = args.warmup + args.steps

    # periodic refill alone, to separate the stage kernels inside the step
    reps = max(args.steps, 5)
    s0, s1 = ev(), ev()
    s0.record()
    for _ in range(reps):
        _lib.check(lib.bssn_periodic_sync_f64(U.data_ptr(), Sx, Sy, Sz, grid.ng,
                                              stream_handle(dev)))
    s1.record()
    torch.cuda.synchronize()
    ms_sync = s0.elapsed_time(s1) / reps
    ms_stage = (ms_step - ms_sync) / 4.0

    # the same step keeping the reference's carried ghost values (per stage: sync of the stage
    # input, every padded point advanced), on a copy of the state
    evo_c = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="periodic", integrator="rk4",
                          ghosts="carried")
    Uc = U.clone()
    evo_c.step_(Uc, t, dt)
    c0, c1 = ev(), ev()
    c0.record()
    for i in range(reps):
        evo_c.step_(Uc, t + i * dt, dt)
    c1.record()
    torch.cuda.synchronize()
    ms_carried = c0.elapsed_time(c1) / reps
    del Uc, evo_c

    # the same step with the radiative outer boundary instead of periodic ghosts (per stage:
    # the fused kernel over the interior with clamped stencils, then the boundary kernel over
    # the ghost zones), on a copy of the state
    evo_s = BSSNEvolution(grid, p, order=4, ko_sigma=0.1, bc="sommerfeld", integrator="rk4")
    Us = U.clone()
    evo_s.step_(Us, t, dt)
    b0, b1 = ev(), ev()
    b0.record()
    for i in range(reps):
        evo_s.step_(Us, t + i * dt, dt)
    b1.record()
    torch.cuda.synchronize()
    ms_sommerfeld = b0.elapsed_time(b1) / reps
    sommerfeld_finite = bool(torch.isfinite(Us).all())
    del Us, evo_s

    # ---------------------------------------------------------------- bare fused RHS
    S16 = scalars(p, grid.dx, grid.dy, grid.dz, dt, t)
    out = torch.empty_like(U)
    for _ in range(2):
        rhs_into(U, out, S16, 4)
    r0, r1 = ev(), ev()
    r0.record()
    for _ in range(reps):
        rhs_into(U, out, S16, 4)
    r1.record()
    torch.cuda.synchronize()
    ms_rhs = r0.elapsed_time(r1) / reps

    # ---------------------------------------------------------------- two-pass RHS
    ms_twopass = None
    if not args.no_twopass:
        D = torch.empty((138, Sx, Sy, Sz), dtype=torch.float64, device=dev)
        rhs_twopass_into(U, out, D, S16, 4)
        q0, q1 = ev(), ev()
        q0.record()
        for _ in range(reps):
            rhs_twopass_into(U, out, D, S16, 4)
        q1.record()
        torch.cuda.synchronize()
        ms_twopass = q0.elapsed_time(q1) / reps
        del D

    # ---------------------------------------------------------------- roofline of the stage kernel
    flops_launch = FLOP_STAGE * npts
    bytes_launch = sum(BYTES_STAGES) / 4.0 * npts
    t_launch = ms_stage * 1e-3
    hbm_frac = bytes_launch / t_launch / PEAK_HBM
    fp64_frac = flops_launch / t_launch / PEAK_FP64
    if bytes_launch / PEAK_HBM >= flops_launch / PEAK_FP64:
        roof = {"bound": "hbm", "achieved": round(bytes_launch / t_launch / 1e9, 2),
                "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": round(hbm_frac, 4)}
    else:
        roof = {"bound": "mfma", "achieved": round(flops_launch / t_launch / 1e12, 3),
                "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": round(fp64_frac, 4)}
    roof["traffic"] = measured_traffic(f"bssn_rk4_step_{N}^3_fp64_bl_binary")
    roof["kernel"] = "k_fused<4,KO,WRAP> (RK4 stage)"
    roof["fp64_frac"] = round(fp64_frac, 4)
    roof["hbm_frac"] = round(hbm_frac, 4)
    roof["ms_per_launch"] = round(ms_stage, 4)

    if args.copy_gib > 0:
        n = int(args.copy_gib * (1 << 30)) // 8
        a = torch.ones(n, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        torch.neg(a, out=b)          # an elementwise kernel: reads and writes copy_gib GiB
        torch.cuda.synchronize()
        del a, b

    value = npts / (ms_step * 1e-3) / 1e6
    res = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"bssn_rk4_step_{N}^3_fp64_bl_binary", "interior": N,
                   "padded": Sx, "h": grid.dx, "fd_order": 4, "ko_order": 6, "ko_sigma": 0.1,
                   "integrator": "rk4", "masses": [0.5, 0.5], "separation": 6.0,
                   "bc": "periodic"},
        "wall_s_timed": round(wall, 4),
        "rhs_ms": round(ms_rhs, 4),
        "rhs_mpts_s": round(npts / (ms_rhs * 1e-3) / 1e6, 2),
        "sync_ms": round(ms_sync, 4),
        "carried_ghost_step_ms": round(ms_carried, 4),
        "sommerfeld_step_ms": round(ms_sommerfeld, 4),
        "sommerfeld_step_mpts_s": round(npts / (ms_sommerfeld * 1e-3) / 1e6, 2),
        "sommerfeld_finite": sommerfeld_finite,
        "roofline": roof,
        "schedule": schedule(),
        "constraints": {"order": 4, "H_l2_initial": float("%.6e" % h0),
                        "M_l2_initial": float("%.6e" % m0),
                        "H_l2_final": float("%.6e" % h1), "M_l2_final": float("%.6e" % m1),
                        "steps": nsteps_done, "t_final": round(t, 6)},
    }
    if ms_twopass is not None:
        res["twopass_rhs_ms"] = round(ms_twopass, 4)
        res["twopass_rhs_mpts_s"] = round(npts / (ms_twopass * 1e-3) / 1e6, 2)
        res["fused_vs_twopass_rhs"] = round(ms_twopass / ms_rhs, 3)
        # the two-pass RHS at the HBM roofline: the bundle pass reads 24 fields and writes 138
        # arrays, the algebra pass reads both and writes 24 fields, 348 doubles per padded point
        floor_ms = 348 * 8 * (N + 6) ** 3 / PEAK_HBM * 1e3
        res["twopass_hbm_floor_ms"] = round(floor_ms, 4)
        res["fused_vs_twopass_hbm_floor"] = round(floor_ms / ms_rhs, 3)
    del out, U
    torch.cuda.empty_cache()
    if not args.no_fd6:
        res["fd6_ko8"] = bench_order(6, (128, N), reps)
        res["fd8_ko8"] = bench_order(8, (128,), reps)
        torch.cuda.empty_cache()
    if not args.no_int8:
        res["int8_split"] = bench_int8(N, reps)
        torch.cuda.empty_cache()
    if not args.no_batch:
        res["batched"] = bench_batched(reps)
        torch.cuda.empty_cache()
    if not args.no_bfp:
        res["bfp"] = bench_bfp(N, reps)
        torch.cuda.empty_cache()
    if not args.no_amr:
        res["amr"] = bench_amr(reps)
        torch.cuda.empty_cache()
    if not args.no_longrun:
        res["longrun"] = bench_longrun()
        torch.cuda.empty_cache()
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
