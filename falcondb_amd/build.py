"""Build recipe for the package's native libraries (run on a host with hipcc;
no GPU needed).

    python -m falcondb_amd.build          # builds everything below
    libfalcon_gpu.so     HIP kernels + C ABI (include/falcon_gpu.h), gfx950 only
    libfgpu_datagen.so   host C++ synthetic table generators

Every HIP file is compiled with packed-FP32 instructions disabled: the compiler
would otherwise form v_pk_* FP32 ops from ordinary float code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
ROOT = PKG.parent
BUILD = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FGPU_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-pass-failed", f"-I{ROOT / 'include'}",
]
HIP_SOURCES = ["fgpu_runtime.hip", "fgpu_filter.hip", "fgpu_groupby.hip", "fgpu_gbpart.hip",
               "fgpu_join.hip", "fgpu_batch.hip", "fgpu_topk.hip", "fgpu_sort.hip", "fgpu_server.hip"]

GPU_LIB = PKG / "libfalcon_gpu.so"
GEN_LIB = PKG / "libfgpu_datagen.so"


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        # hipcc's host pass warns that it does not know the packed-fp32 feature
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}\n{msg}")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    return obj


def build_gpu_lib() -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in HIP_SOURCES if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = GPU_LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_datagen_lib() -> Path:
    tmp = GEN_LIB.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-pthread", "-shared", "-o", str(tmp),
          str(CSRC / "fgpu_datagen.cpp")])
    os.replace(tmp, GEN_LIB)
    return GEN_LIB


STREAM_READ = PKG / "fgpu_stream_read"


def build_stream_read() -> Path:
    """The standalone read-ceiling probe (bench.py --ceiling)."""
    tmp = STREAM_READ.with_suffix(".tmp")
    _run([HIPCC, *HIP_FLAGS, "-o", str(tmp), str(CSRC / "fgpu_stream_read.hip")])
    os.replace(tmp, STREAM_READ)
    return STREAM_READ


def build_all() -> list[Path]:
    return [build_gpu_lib(), build_datagen_lib(), build_stream_read()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
    sys.exit(0)
