"""Builds the gfx950 library ``vira_amd/lib/libvira_gpu.so`` with hipcc.

Every HIP source is compiled for gfx950 only, without FP contraction (the
renderer's results are compared bit for bit with a CPU restatement) and with
packed-FP32 instruction selection disabled.  Objects are compiled in parallel
and linked into one shared library exporting the C-ABI of include/vira_gpu.h.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "lib"
LIB = OUT_DIR / "libvira_gpu.so"
OBJ_DIR = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
    f"-I{ROOT / 'include'}",
]
SOURCES = ["capi.hip", "dem.hip", "bvh.hip", "render.hip", "adjoint.hip", "batch.hip", "sensor.hip", "fftconv.hip", "denoise.hip",
           "quipu.hip", "scan.hip", "catalogue.hip"]
# C++ programs over include/vira_gpu.hpp, linked against the library (host code only)
EXAMPLES = ["render_dem.cpp", "quipu_scene.cpp"]


def _compile(src: Path, obj: Path) -> None:
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    # keep compiler warnings visible to whoever builds
    noise = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build(verbose: bool = True) -> Path:
    OUT_DIR.mkdir(exist_ok=True)
    OBJ_DIR.mkdir(exist_ok=True)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        futs = []
        for s in SOURCES:
            obj = OBJ_DIR / (Path(s).stem + ".o")
            objs.append(obj)
            futs.append(ex.submit(_compile, CSRC / s, obj))
        for f in futs:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} ({LIB.stat().st_size // 1024} KiB)")
    for ex_src in EXAMPLES:
        exe = OUT_DIR / Path(ex_src).stem
        cmd = [HIPCC, "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(PKG / "examples" / ex_src), "-o",
               str(exe), f"-L{OUT_DIR}", "-lvira_gpu", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {ex_src} failed:\n{r.stdout}\n{r.stderr}")
    return LIB


if __name__ == "__main__":
    build()
