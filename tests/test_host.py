"""CPU: the C ABI library loads and exports what include/bssn_hip.h declares, its argument
checks fire before any device work, and the Python host mirror keeps the reference's
conventions (scalar pack, output order, grid, initial data, error behaviour)."""
from __future__ import annotations

import ctypes
import json
import math
import re

import numpy as np
import pytest
import torch

from conftest import ROOT

from bssn3d_hip import _lib, initial_data as idata
from bssn3d_hip.grid import Grid, ghost_width
from bssn3d_hip.rhs import BSSNSolver, as_tensor
from bssn3d_hip.state import PhysicsParams, VAR_NAMES, pack_S, scalars

HEADER = ROOT / "include" / "bssn_hip.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|long|long long|size_t|void|const char\*)\s+(bssn_\w+)\(",
                                 text, re.M)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.SIGNATURES)


def test_library_has_no_undefined_internal_symbols():
    """Every launcher the C ABI calls is linked in (a compile unit missing from the build would
    leave an undefined bssn:: symbol that only fails at load time)."""
    import shutil
    import subprocess
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--undefined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    assert not [l for l in out.splitlines() if "4bssn" in l]


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert _lib.load().bssn_abi_version() == _lib.ABI_VERSION


def test_argument_errors_before_device_work():
    lib = _lib.load()
    S = _lib.dbl(pack_S(2.0, 0.06, 0.01, 0.1, 0.6, 20.0, 0.0, (1, 1, 1, 1), (1, 0), 0.1, 0.1))
    # null pointers, bad order, bad shape, bad sigma: BSSN_E_ARG with a message
    assert lib.bssn_rhs_f64(None, S, None, 8, 8, 8, 4, 0.0, None) == 1
    assert b"null" in lib.bssn_last_error()
    fake = ctypes.c_void_p(16)
    assert lib.bssn_rhs_f64(fake, S, fake, 8, 8, 8, 5, 0.0, None) == 1
    assert b"fd_order" in lib.bssn_last_error()
    assert lib.bssn_rhs_f64(fake, S, fake, 0, 8, 8, 4, 0.0, None) == 1
    assert lib.bssn_rhs_f64(fake, S, fake, 8, 8, 8, 4, -1.0, None) == 1
    bad = _lib.dbl([2.0, 0.06, 0.0] + [0.1] * 13)   # dt = 0: CAHD divides by it
    assert lib.bssn_rhs_f64(fake, bad, fake, 8, 8, 8, 4, 0.0, None) == 1
    assert b"dt" in lib.bssn_last_error()
    assert lib.bssn_rk4_step_f64(fake, fake, 8, 8, 8, S, 4, 0.1, 3, 1, 0.0, 0.0, 1e-4, 1e-4, None) == 1
    assert lib.bssn_periodic_sync_f64(fake, 5, 5, 5, 3, None) == 1   # interior < ng
    assert lib.bssn_msrk_step_f64(fake, fake, fake, fake, 12, 12, 12, S, 4, 0.1, 3, 1, 0.0, 0.1,
                                  1e-4, 1e-4, None) == 1
    assert b"distinct" in lib.bssn_last_error()
    # device scalar pack: null pack, bad order, bad shape are caught on the host
    assert lib.bssn_rhs_sdev_f64(fake, None, fake, 8, 8, 8, 4, 0.0, None) == 1
    assert b"null" in lib.bssn_last_error()
    assert lib.bssn_rhs_sdev_f64(fake, fake, fake, 8, 8, 8, 10, 0.0, None) == 1
    assert b"fd_order" in lib.bssn_last_error()
    # FD order 8 needs 4 ghosts for its KO 8 stencil on the Sommerfeld step
    xm = _lib.dbl([0.0, 0.0, 0.0])
    assert lib.bssn_rk4_step_sommerfeld_f64(fake, fake, 16, 16, 16, S, 8, 0.1, 3, xm, 0.0,
                                            0.1, 1e-4, 1e-4, None) == 1
    assert b"reach" in lib.bssn_last_error()
    # the INT8-split derivatives are FD order 4 only
    assert lib.bssn_deriv_bundle_i8_f64(fake, fake, 16, 16, 16, 8, _lib.dbl([0.1] * 3), 7,
                                        None) == 1
    assert b"fd_order" in lib.bssn_last_error()
    assert lib.bssn_rk4_step_sdev_f64(fake, fake, 8, 8, 8, None, 4, 0.1, 3, 1, 1e-4, 1e-4,
                                      None) == 1
    assert lib.bssn_rk4_step_sdev_f64(fake, fake, 5, 5, 5, fake, 4, 0.1, 3, 1, 1e-4, 1e-4,
                                      None) == 1   # interior < ng
    assert lib.bssn_step_workspace_bytes(0, 10, 10, 10) == 3 * 24 * 1000 * 8
    assert lib.bssn_step_workspace_bytes(1, 10, 10, 10) == 2 * 24 * 1000 * 8


def test_library_errors_raise():
    with pytest.raises(_lib.BSSNError, match="fd_order"):
        _lib.check(_lib.load().bssn_rhs_f64(ctypes.c_void_p(16), _lib.dbl([1.0] * 16),
                                            ctypes.c_void_p(16), 8, 8, 8, 3, 0.0, None))


def test_scalar_pack_and_output_map():
    """16 scalars in the reference's order (fused_rhs_cuda.py:52-55); 24 outputs in state order."""
    s = scalars(PhysicsParams(), 0.06, 0.07, 0.08, 0.01, 0.5)
    assert len(s) == 16
    assert s[:7] == [2.0, 0.06, 0.01, 0.06, 0.6, 20.0, 0.5]
    assert s[7:13] == [1.0, 1.0, 1.0, 1.0, 1.0, 0.0]
    assert s[13:] == [0.06, 0.07, 0.08]
    assert VAR_NAMES[0] == "alpha" and len(VAR_NAMES) == 24


def test_missing_library_error(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "nope.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(FileNotFoundError, match="build.py"):
        _lib.load()


def test_host_tensors_rejected():
    g = Grid.from_domain(8, order=4, ko_order=6)
    s = idata.minkowski(g)
    with pytest.raises(ValueError, match="HIP device"):
        as_tensor(s)
    with pytest.raises(TypeError):
        as_tensor(s.float())
    with pytest.raises(ValueError, match="shape"):
        as_tensor(torch.zeros(23, 4, 4, 4, dtype=torch.float64))


def test_solver_validation():
    g = Grid.from_domain(8, order=4, ko_order=6)
    with pytest.raises(ValueError, match="ng"):
        BSSNSolver(g, order=6)
    with pytest.raises(ValueError, match="scheme"):
        BSSNSolver(g, order=4, scheme="nope")
    with pytest.raises(ValueError, match="precision"):
        BSSNSolver(g, order=4, precision="fp32_contraction")
    g8 = Grid.from_domain(8, order=8)
    assert g8.ng == 4
    s8 = BSSNSolver(g8, order=8)
    assert s8.ko_order == 8                  # FD 8 pairs with KO 8 (the widest KO stencil)
    with pytest.raises(ValueError, match="KO order"):
        BSSNSolver(g8, order=8, ko_order=6)
    with pytest.raises(ValueError, match="order"):
        BSSNSolver(g8, order=10)
    with pytest.raises(ValueError, match="int8"):
        BSSNSolver(g8, order=8, scheme="int8")
    s = BSSNSolver(g, order=4, scheme="verbatim")
    assert s.scheme == "twopass" and s.dt == 0.25 * g.dx


def test_grid_matches_reference_convention():
    g = Grid.from_domain(16, order=6)
    assert g.ng == 4 and g.shape == (24, 24, 24)
    assert g.x[g.ng] == -0.5 and abs(g.x[1] - g.x[0] - 1 / 16) < 1e-15
    assert ghost_width(4, 6) == 3


def test_puncture_closed_form():
    g, st, pos = idata.config_single_puncture(seed=1, N=12)
    d = st.numpy()
    x, y, z = g.mesh()
    r = np.sqrt((x - pos[0][0]) ** 2 + (y - pos[0][1]) ** 2 + (z - pos[0][2]) ** 2)
    psi = 1 + 1.0 / (2 * np.maximum(r, g.dx / 2))
    # perturbations are at most 1e-4
    assert np.abs(d[1] - psi ** -4).max() <= 1e-4 + 1e-15
    assert np.abs(d[0] - psi ** -2).max() <= 1e-4 + 1e-15
    assert np.abs(d[3] - 1).max() <= 1e-4 and np.abs(d[4]).max() <= 1e-4
    assert np.abs(d[2]).max() > 0   # every field is perturbed
    # punctures sit at least h/4 off the grid planes
    for c, ax in zip(pos[0], (g.x, g.y, g.z)):
        assert np.min(np.abs(ax - c)) >= 0.25 * g.dx - 1e-12
    # the perturbation of one field is periodic on the domain (integer wavenumbers)
    pert = idata.Perturbation.draw(np.random.default_rng(3))
    assert np.all(np.abs(pert.k).max(axis=2) >= 1) and np.all(np.abs(pert.k) <= 4)


def test_binary_config_shape():
    g, st, pos = idata.config_binary(seed=0, N=16, L=16.0)
    assert st.shape == (24, 22, 22, 22) and g.dx == 1.0
    assert abs((pos[1][0] - pos[0][0]) - 6.0) < 1e-12


def _det_gt(d):
    g = d[3:9]
    return (g[0] * (g[3] * g[5] - g[4] * g[4]) - g[1] * (g[1] * g[5] - g[4] * g[2]) +
            g[2] * (g[1] * g[4] - g[3] * g[2]))


def _conformal_trace_At(d):
    g, A = d[3:9], d[12:18]
    iu = [g[3] * g[5] - g[4] * g[4], g[2] * g[4] - g[1] * g[5], g[1] * g[4] - g[2] * g[3],
          g[0] * g[5] - g[2] * g[2], g[1] * g[2] - g[0] * g[4], g[0] * g[3] - g[1] * g[1]]
    return (iu[0] * A[0] + iu[3] * A[3] + iu[5] * A[5] +
            2.0 * (iu[1] * A[1] + iu[2] * A[2] + iu[4] * A[4]))


def test_gauge_wave_invariants():
    """3D/tests/unit/test_bssn_state.py:93-102, and the wave on each axis (:105-...): unit-
    determinant conformal metric, trace-free At, alpha^2 = chi^-3, Gt only along the axis."""
    g = Grid.from_domain(12, order=4)
    for axis, amp in ((0, 0.01), (1, 0.02), (2, 0.02)):
        d = idata.gauge_wave(g, amplitude=amp, axis=axis).numpy()
        assert np.max(np.abs(_det_gt(d) - 1.0)) < 1e-12
        assert np.max(np.abs(_conformal_trace_At(d))) < 1e-12
        assert np.all(d[1] > 0.0) and np.all(d[0] > 0.0)
        assert np.max(np.abs(d[0] ** 2 - d[1] ** -3.0)) < 1e-12
        H = d[0] ** 2                                    # alpha = sqrt(H)
        for a, slot in enumerate((3, 6, 8)):             # H^(2/3) on the wave axis, else chi
            want = H ** (2.0 / 3.0) if a == axis else d[1]
            assert np.max(np.abs(d[slot] - want)) < 1e-12
        for a in range(3):
            gt = np.abs(d[18 + a]).max()
            assert (gt > 0.0) if a == axis else (gt == 0.0)


def _load_module(name, path):
    import importlib.util
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_autotune_space_and_installed_choice():
    """The autotuner's knobs are the ones the kernel headers read (#ifndef NAME), candidate keys are
    unique, and the installed tuning.json names a candidate of the full space that passed its
    check; build.tuned_defines() hands exactly those -D knobs to the installed build."""
    pkg = ROOT / "integer-schemes_amd"
    at = _load_module("bssn_autotune", pkg / "autotune.py")
    bd = _load_module("bssn_build", pkg / "build.py")
    stage_h = "".join((pkg / "csrc" / h).read_text()
                      for h in ("bssn_stage.h", "bssn_brick.h", "bssn_params.h"))
    for space in at.SPACES.values():
        keys = [at.key_of(d) for d in space]
        assert len(set(keys)) == len(keys) and keys[0] == "default"
        for d in space:
            shape = [d.get(f"BSSN_BRICK_{a}", 4) for a in "XYZ"]
            assert shape[0] * shape[1] * shape[2] == 64
            for name, v in d.items():
                assert f"#ifndef {name}" in stage_h, name
                # the brick order's 0 means rows of columns; every other knob counts from 1
                assert isinstance(v, int) and v >= (0 if name == "BSSN_BRICK_TILE_XY" else 1)
    t = json.loads((pkg / "tuning.json").read_text())
    full = {at.key_of(d): d for d in at.SPACES["full"]}
    assert t["best"] in full and t["candidates"][t["best"]]["passed"]
    assert t["defines"] == [f"{a}={b}" for a, b in full[t["best"]].items()]
    # the installed record is the selection rule applied to its own measurements
    assert at.choose(t["candidates"], margin=t["margin"])["best"] == t["best"]
    assert bd.tuned_defines() == t["defines"]
    # the rule: beat the defaults' step by more than the margin, without a slower bare RHS
    def cand(step, rhs, passed=True, d=None):
        return {"step_ms": step, "rhs_ms": rhs, "passed": passed, "defines": d or {}}
    base = cand(40.0, 6.0)
    r = {"default": base, "a": cand(39.9, 6.0, d={"BSSN_WB_LEAD": 3}),
         "b": cand(38.0, 6.5, d={"BSSN_RHS_DEPTH": 1}), "c": cand(37.0, 6.0, passed=False)}
    assert at.choose(r)["best"] == "default"        # a: noise, b: slower RHS, c: failed
    r["d"] = cand(39.0, 6.01, d={"BSSN_STAGE_DEPTH": 2})
    rec = at.choose(r)
    assert rec["best"] == "d" and rec["defines"] == ["BSSN_STAGE_DEPTH=2"]
    # merging a second run: its candidates join, its defaults are the baseline
    new = at.choose({"default": cand(39.5, 6.0), "e": cand(38.5, 5.9, d={"BSSN_BRICK_X": 2})},
                    workload="w")
    m = at.merge(dict(rec, workload="w"), new)
    assert set(m["candidates"]) == {"default", "a", "b", "c", "d", "e"}
    assert m["candidates"]["default"]["step_ms"] == 39.5 and m["best"] == "e"
    with pytest.raises(ValueError):
        at.merge(dict(rec, workload="other"), new)
