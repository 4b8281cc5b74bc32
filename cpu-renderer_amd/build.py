"""Build the gfx950 library (lib/libcrgpu.so) in-tree with hipcc.

The kernels translation unit is compiled for gfx950 with the host pieces
(C ABI, scene preparation) compiled host-only; all three compile in parallel.
Device code is built with -ffp-contract=off (no implicit multiply-add: the
transport code must round exactly like the reference) and with packed FP32
disabled.
"""
from __future__ import annotations

import os
import pathlib
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
BUILD = HERE / "_build"
LIB = HERE / "lib" / "libcrgpu.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]

UNITS = [("kernels.hip", []), ("capi.cpp", ["--offload-host-only"]), ("engine.cpp", ["--offload-host-only"]),
         ("host_scene.cpp", ["--offload-host-only"]), ("film_host.cpp", ["--offload-host-only"])]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    jobs = []
    for src, extra in UNITS:
        obj = BUILD / (src + ".o")
        jobs.append((obj, [HIPCC, *COMMON, *extra, "-c", str(CSRC / src), "-o", str(obj)]))
    with ThreadPoolExecutor(len(jobs)) as ex:
        for out in ex.map(lambda j: _run(j[1]), jobs):
            if verbose and out:
                print(out)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", str(tmp), *[str(j[0]) for j in jobs], "-lpthread", "-lz"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
