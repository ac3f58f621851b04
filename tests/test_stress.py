"""GPU: the extreme-mass-ratio stress run's code path at a small size (48^3 with the same
masses, separation and spacing rule, 20 steps): the INT8-split run and the INT8-split run under
the batched-block schedule (6^3 blocks of 8^3 with halo copies) track the FP64 run's constraint
norms within the apples gate's 1e-6 and stay finite; the batched-block FP64 run gives the FP64
run's samples bit for bit."""
from __future__ import annotations

import pytest

from conftest import gpu_available

from bssn3d_hip.stress import stress

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a HIP device")]


def test_stress_small():
    print("[case] stress run 48^3, 20 steps", flush=True)
    res = stress(N=48, steps=20, sample_every=5, device="cuda:0")
    f, i8, bt, ib = res["fused"], res["int8"], res["batched"], res["int8_batched"]
    assert f["steps"] == i8["steps"] == bt["steps"] == ib["steps"] == 20
    assert len(f["samples"]) == 5
    assert all(s["finite"] for s in f["samples"])
    print({k: f"{v:.2e}" for k, v in i8["vs_fused"]["dev"].items()})
    assert i8["vs_fused"]["finite"] and i8["vs_fused"]["same_length"]
    assert max(i8["vs_fused"]["dev"].values()) <= 1e-6
    print({k: f"{v:.2e}" for k, v in ib["vs_fused"]["dev"].items()})
    assert ib["vs_fused"]["finite"] and ib["vs_fused"]["same_length"]
    assert max(ib["vs_fused"]["dev"].values()) <= 1e-6
    for a, b in zip(f["samples"], bt["samples"]):
        assert (a["H"], a["M"], a["max_alpha"]) == (b["H"], b["M"], b["max_alpha"])
