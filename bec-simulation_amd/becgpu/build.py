"""Build libbecgpu.so (gfx950) in-tree with hipcc.

Usage: python -m becgpu.build   (from bec-simulation_amd/)  or  build() from Python.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)                      # bec-simulation_amd/
CSRC = os.path.join(ROOT, "csrc")
REPO = os.path.dirname(ROOT)
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG_DIR, "libbecgpu.so")

# (source, object, defines): bec_kernels.hip is compiled once for its shared code and once
# per component count D = 2F+1, so that the parts compile in parallel
KERNEL_DS = (1, 3, 5, 7, 9, 11, 13)
# bec_group.hip likewise: shared part and the grouped x pass of 1, 3 and 5 components
GROUP_DS = (1, 3, 5)
UNITS = ([("bec_kernels.hip", f"bec_kernels_p{d}.o", [f"-DBEC_PART={d}"]) for d in (0,) + KERNEL_DS]
         + [("bec_group.hip", f"bec_group_p{d}.o", [f"-DBEC_PART={d}"]) for d in (0,) + GROUP_DS]
         + [(src, os.path.splitext(src)[0] + ".o", [])
            for src in ("bec_strided.hip", "bec_xsite.hip", "bec_aux.hip", "bec_generic.hip",
                        "bec_capi.hip")])
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed-FP32 instructions (v_pk_*_f32) anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC, "-I", os.path.join(REPO, "include"),
    "-Wno-unused-result",
]


def _compile(unit) -> str:
    src, objname, defs = unit
    obj = os.path.join(BUILD, objname)
    cmd = [HIPCC, *FLAGS, *defs, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src} {' '.join(defs)}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(UNITS), os.cpu_count() or 8)) as ex:
        objs = list(ex.map(_compile, UNITS))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
