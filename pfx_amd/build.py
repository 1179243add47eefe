"""Build libpfxhip.so for gfx950 in-tree (hipcc, no JIT cache).

`python -m pfx_amd.build` compiles every source under pfx_amd/csrc into
pfx_amd/_build/*.o, links pfx_amd/libpfxhip.so and builds the C++
command-line caller pfx_amd/pfxh-render against it. All device and host code
is built with -ffp-contract=off (no fused multiply-add), IEEE division and
square root (hipcc's default for HIP), and with packed FP32 instructions
disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "_build"
LIB = HERE / "libpfxhip.so"
SOURCES = ["kernels.hip", "layout.cpp", "trace.cpp", "checkpoint.cpp", "denoise.hip", "denoise_host.cpp", "batch.cpp", "atrous.hip",
           "learn.hip", "learn_host.cpp", "stage.cpp",
           "sobol_table.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src: str) -> tuple[str, int, str]:
    obj = OUT / (Path(src).stem + ".o")
    cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    log = "\n".join(
        line for line in (proc.stdout + proc.stderr).splitlines() if "packed-fp32-ops" not in line
    )
    return src, proc.returncode, log


def build(verbose: bool = False) -> Path:
    OUT.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as pool:
        results = list(pool.map(_compile, SOURCES))
    failed = [(s, log) for s, rc, log in results if rc != 0]
    for src, rc, log in results:
        if verbose and log.strip():
            print(f"[{src}]\n{log}", file=sys.stderr)
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n".join(f"{s}:\n{log}" for s, log in failed))
    objs = [str(OUT / (Path(s).stem + ".o")) for s in SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-o", str(tmp), *objs]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("link failed:\n" + proc.stdout + proc.stderr)
    os.replace(tmp, LIB)
    # the C++ command-line caller, linked against the library beside it
    cli = HERE / "pfxh-render"
    cmd = ["g++", "-O2", "-std=c++17", str(HERE / "cli" / "pfxh_render.cpp"), "-o", str(cli), "-L", str(HERE),
           "-lpfxhip", "-Wl,-rpath,$ORIGIN"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("pfxh-render failed to build:\n" + proc.stdout + proc.stderr)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
