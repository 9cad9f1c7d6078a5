"""Builds nec2c_amd/lib/libvie_hip.so for gfx950 with hipcc (no torch, no JIT).

    python -m nec2c_amd.build        # or nec2c_amd.build.build()

Objects go to nec2c_amd/_build/, the shared library to nec2c_amd/lib/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
SOURCES = ["vie_passes_f32.hip", "vie_passes_f64.hip", "spectrum.hip", "krylov.hip", "farfield.hip",
           "conductor_rows.hip", "wire_volume.hip", "mesh_check.hip", "voxels.cpp", "api.hip"]
ARCH = os.environ.get("VIE_HIP_ARCH", "gfx950")
FLAGS = [f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
         # packed-FP32 instructions are kept out of every kernel
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-I" + os.path.join(HERE, "..", "include")]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    out = os.path.join(LIB, "libvie_hip.so")
    tmp = out + ".tmp"
    # -Bsymbolic: the library's calls to its own exported names (vie_fft_create,
    # dielectric_fft_prepare, ...) bind inside it, so a host program that also
    # defines those names (nec2c with its CPU backends) never interposes them.
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
