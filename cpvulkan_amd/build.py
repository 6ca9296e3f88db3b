"""Build of the native libraries (no cmake/ninja: hipcc and gcc directly).

    python -m cpvulkan_amd.build            # everything
    python -m cpvulkan_amd.build libcpv     # one target

Products (git-ignored, built in-tree so they travel with the source tree):
    cpvulkan_amd/lib/libcpv.so              HIP draw path + C-ABI (include/cpv.h)
    cpvulkan_amd/lib/libcpvulkan_amd_icd.so Vulkan ICD front end
    oracle/build/libcpvref.so               CPU oracle (test infrastructure)
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "cpvulkan_amd"
CSRC = PKG / "csrc"
LIB = PKG / "lib"
ORACLE = ROOT / "oracle"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Strict IEEE fp32: no contraction, correctly rounded divide/sqrt, denormals
# kept, and packed-fp32 instructions disabled.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", f"-I{ROOT / 'include'}", f"-I{CSRC}"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ... (exit {res.returncode})")


def _obj(src: Path, out_dir: Path, hip: bool) -> Path:
    out_dir.mkdir(parents=True, exist_ok=True)
    obj = out_dir / (src.name + ".o")
    if hip:
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    else:
        _run([HIPCC, *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-x", "c++", "-c", str(src), "-o", str(obj)])
    return obj


def build_libcpv() -> Path:
    objdir = PKG / "_build" / "libcpv"
    srcs = [(CSRC / "cpv_runtime.hip", True), (CSRC / "cpv_launch.hip", True), (CSRC / "cpv_mlp.hip", True),
            (CSRC / "spirv_compile.cpp", False)]
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        objs = list(ex.map(lambda s: _obj(s[0], objdir, s[1]), srcs))
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "libcpv.so"
    tmp = LIB / "libcpv.so.tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lhiprtc"])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    out_dir = ORACLE / "build"
    out_dir.mkdir(parents=True, exist_ok=True)
    out = out_dir / "libcpvref.so"
    _run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
          "-Wall", "-o", str(out), str(ORACLE / "cpvref.c"), "-lm"])
    return out


TARGETS = {"libcpv": build_libcpv, "oracle": build_oracle}


def build_all(targets: list[str] | None = None) -> None:
    names = targets or list(TARGETS)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(names)) as ex:
        for fut in [ex.submit(TARGETS[n]) for n in names]:
            fut.result()


if __name__ == "__main__":
    build_all(sys.argv[1:] or None)
