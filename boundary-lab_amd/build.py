"""Build libbemhip.so (all HIP sources, gfx950) in-tree: ``python boundary-lab_amd/build.py``."""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "beat_hip" / "libbemhip.so"
SOURCES = ["bem_assembly.hip", "bem_lu.hip", "bem_field.hip", "bem_capi.hip", "bem_tables.hip", "bem_coupled.hip"]
ARCH = os.environ.get("BEM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def build(verbose: bool = False) -> Path:
    objs = []
    flags = [
        f"--offload-arch={ARCH}",
        "-O3",
        "-fPIC",
        "-std=c++17",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        # register-resident LU sub-panels (up to 2 x 16 complex per thread) stay in VGPRs instead of scratch
        "-mllvm", "-amdgpu-promote-alloca-to-vector-limit=1024",
        "-I", str(HERE.parent / "include"),
    ]
    build_dir = HERE / "_build"
    build_dir.mkdir(exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = build_dir / (src + ".o")
        objs.append(obj)
        cmd = [hipcc(), *flags, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        for src, out in failed:
            sys.stderr.write(f"--- {src} ---\n{out}\n")
        raise RuntimeError("HIP build failed: " + ", ".join(s for s, _ in failed))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
