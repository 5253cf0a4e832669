"""Build libcube2_short.so (HIP for gfx950) in-tree with hipcc."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "libcube2_short.so"
ARCH = os.environ.get("CUBE2_OFFLOAD_ARCH", "gfx950")

# No packed FP32 instructions (v_pk_*_f32); IEEE-rounded fp32 divide/sqrt so
# the exact PP mode can reproduce the reference's arithmetic.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-fast-math", "-munsafe-fp-atomics",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


def build(verbose: bool = False, jobs: int = 8) -> Path:
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    procs = []
    objs = []
    for src in sources():
        obj = objdir / (src.name + ".o")
        objs.append(obj)
        cmd = [hipcc, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
        if src.suffix == ".cpp":
            cmd = [hipcc, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
        procs.append((src, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0), verbose)
    for p in procs:
        _wait(p, verbose)
    tmp = OUT.with_suffix(".so.tmp")
    link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


def _wait(item, verbose):
    src, cmd, p = item
    out, _ = p.communicate()
    text = out.decode(errors="replace")
    if p.returncode:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{text}")
    if verbose and text.strip():
        print(text, file=sys.stderr)


if __name__ == "__main__":
    print(build(verbose=True))
