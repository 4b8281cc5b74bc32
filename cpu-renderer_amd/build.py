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


def build(verbose: bool = False, defines=(), lib: pathlib.Path = LIB, build_dir: pathlib.Path = BUILD) -> pathlib.Path:
    """Compile and link the library.  `defines` (NAME=VALUE strings) set the
    kernels' launch and traversal knobs (the CR_* macros of the sources) for a
    variant build, written to `lib` with objects in `build_dir`."""
    build_dir.mkdir(exist_ok=True)
    lib.parent.mkdir(exist_ok=True)
    dflags = [f"-D{d}" for d in defines]
    jobs = []
    for src, extra in UNITS:
        obj = build_dir / (src + ".o")
        jobs.append((obj, [HIPCC, *COMMON, *dflags, *extra, "-c", str(CSRC / src), "-o", str(obj)]))
    with ThreadPoolExecutor(len(jobs)) as ex:
        for out in ex.map(lambda j: _run(j[1]), jobs):
            if verbose and out:
                print(out)
    tmp = lib.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-o", str(tmp), *[str(j[0]) for j in jobs], "-lpthread", "-lz"])
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("-v", action="store_true")
    ap.add_argument("--define", action="append", default=[], help="NAME=VALUE for a variant build")
    ap.add_argument("--out", help="library path of a variant build")
    a = ap.parse_args()
    if a.out:
        out = pathlib.Path(a.out).resolve()
        print(build(a.v, a.define, out, HERE / ("_build_" + out.stem)))
    else:
        print(build(a.v, a.define))
