"""Builds the native library ``liar_amd/_lib/libliar_amd.so`` for gfx950.

HIP sources (``csrc/*.hip``) are compiled by hipcc for ``--offload-arch=gfx950``;
host-only C++ sources (``csrc/*.cpp``) by g++.  Both use ``-ffp-contract=off``
so that the FP32 primitive tests evaluate the same expressions on the host and
on the GPU.  Run ``python -m liar_amd.build``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB_DIR = PKG / "_lib"
LIB = LIB_DIR / "libliar_amd.so"
INCLUDE = PKG.parent / "include"

ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA results in VGPRs: the guide kernels read every accumulator with
    # VALU code, and AGPR results cost one v_accvgpr_read each
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
CXX_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-march=x86-64-v3",
    "-pthread",
    "-Wall",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
    f"-I{ROCM / 'include'}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "unknown target feature" not in l
                    and "-packed-fp32-ops" not in l and l.strip())
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}")
    if out.strip():
        print(out, file=sys.stderr)


def _compile(src: Path) -> Path:
    obj = OBJ / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *CXX_FLAGS, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(verbose: bool = False) -> Path:
    """Compiles every source in csrc/ and links the shared library."""
    OBJ.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    if not srcs:
        raise RuntimeError("no sources under liar_amd/csrc")
    workers = min(len(srcs), max(1, (os.cpu_count() or 2) // 2), 8)
    with cf.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-lpthread", "-lz"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle(verbose: bool = False) -> Path:
    """Compiles the CPU oracle (test infrastructure) with its Makefile."""
    oracle_dir = PKG.parent / "oracle"
    make = shutil.which("make")
    if make is None:
        raise RuntimeError("make not found")
    _run([make, "-s", "-C", str(oracle_dir), "-B"])
    lib = oracle_dir / "liboracle.so"
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    build(verbose=True)
    build_oracle(verbose=True)
