"""Build libbssn_hip.so (gfx950) in-tree with hipcc.

    python integer-schemes_amd/build.py          # compile every source, link the library

Each HIP source is compiled separately (in parallel) for --offload-arch=gfx950 and linked into
integer-schemes_amd/bssn3d_hip/libbssn_hip.so, which the Python host layer loads.
``build(out=..., defines=[...])`` builds a variant with extra -D knobs into another path; the
installed library is built with the schedule knobs of tuning.json when the autotuner
(autotune.py) has written one.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "bssn3d_hip" / "libbssn_hip.so"
ARCH = os.environ.get("BSSN_OFFLOAD_ARCH", "gfx950")

# largest compile units first (the fused stage kernels are instantiated per kind)
SOURCES = ["bssn_fused4.hip", "bssn_fused6.hip", "bssn_fused8.hip", "bssn_batch4.hip",
           "bssn_batch6.hip", "bssn_batch8.hip", "bssn_fused4s.hip", "bssn_fused6s.hip",
           "bssn_fused8s.hip",
           "bssn_aux.hip", "bssn_batch.hip", "bssn_i8.hip", "bssn_bc.hip", "bssn_amr.hip",
           "bssn_abi.cpp"]

# Packed-FP32 instructions are disabled for every HIP compile (they are not used by this
# FP64 library; the flag keeps hipcc from forming them from any float code).
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def _compile(src: str, objdir: Path, defines) -> Path:
    obj = objdir / (Path(src).stem + ".o")
    cmd = [hipcc(), *HIP_FLAGS, *[f"-D{d}" for d in defines], "-c", str(CSRC / src), "-o",
           str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


TUNING = PKG / "tuning.json"


def tuned_defines() -> list:
    """The schedule knobs the autotuner chose on the device (tuning.json), [] without one."""
    if not TUNING.exists():
        return []
    import json
    return list(json.loads(TUNING.read_text()).get("defines", []))


def build(verbose: bool = False, out: Path | None = None, defines=None) -> Path:
    """Compile every source and link the library. defines: -D knobs; None = the tuned ones
    (tuning.json) for the installed library, none for a variant build."""
    if defines is None:
        defines = tuned_defines() if out is None else []
    out = Path(out) if out is not None else OUT
    objdir = PKG / "_build" / (out.parent.name if out != OUT else "lib")
    objdir.mkdir(parents=True, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, objdir, defines), SOURCES))
    out.parent.mkdir(parents=True, exist_ok=True)
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}" + (f" with {' '.join(defines)}" if defines else ""))
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
