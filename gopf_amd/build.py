"""Build libgopf_amd.so (gfx950) in-tree with hipcc.

Run ``python -m gopf_amd.build`` or call ``build()``; every call recompiles every source.
``python -m gopf_amd.build --incremental`` recompiles only the sources whose object is
older than the source or one of the headers it includes (from hipcc's dependency files).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgopf_amd.so")
# heaviest units first: the pool starts them before the light ones
SOURCES = ["spectral_ext_rows.hip", "spectral_ext_rows64.hip", "spectral_ext_fwd.hip", "spectral_ext_inv.hip",
           "spectral_ext_fwd64.hip", "spectral_ext_inv64.hip", "spectral_pencils_fwd.hip", "spectral_pencils_inv.hip",
           "spectral_sync.hip", "spectral_r2c.hip", "spectral_c2r.hip", "spectral_update.hip", "spectral_rows.hip",
           "spectral_pencils_fwd64.hip", "spectral_pencils_inv64.hip", "spectral_sync64.hip", "spectral_r2c64.hip",
           "spectral_c2r64.hip", "spectral_update64.hip", "spectral_rows64.hip", "spectral_group.hip",
           "spectral_group64.hip", "spectral_stage.hip", "spectral_persist.hip", "spectral_persist64.hip", "spectral_pencils.hip", "spectral_pencils64.hip", "spectral.hip", "engine.hip",
           "pointwise.hip", "batch.hip", "krylov.hip", "io.hip", "reduce.hip", "model.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
         # packed FP32 instructions stay off in every device compile
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def _obj(src: str) -> str:
    return os.path.join(BUILD, os.path.splitext(src)[0] + ".o")


def _stale(src: str) -> bool:
    """True when the object is missing or older than a file it was built from."""
    obj = _obj(src)
    dep = obj + ".d"
    if not (os.path.exists(obj) and os.path.exists(dep)):
        return True
    t = os.path.getmtime(obj)
    with open(dep) as f:
        text = f.read().replace("\\\n", " ")
    files = text.split(":", 1)[1].split() if ":" in text else [os.path.join(CSRC, src)]
    return any(not os.path.exists(p) or os.path.getmtime(p) > t for p in files)


def _compile(src: str) -> str:
    obj = _obj(src)
    dep = ["-MMD", "-MF", obj + ".d"]
    cmd = [HIPCC] + FLAGS + dep + ["-I", os.path.join(HERE, "..", "include"), "-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-std=c++17", "-O2", "-fPIC", "-Wall"] + dep + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True, incremental: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    todo = [s for s in SOURCES if not incremental or _stale(s)]
    workers = max(1, min(len(todo), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(_compile, todo))
    objs = [_obj(s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} ({len(todo)} of {len(SOURCES)} sources compiled)")
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv[1:])
    sys.exit(0)
