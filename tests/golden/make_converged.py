"""Converged reference film for the relMSE test (written by the reference CPU renderer).

The reference renderer, compiled from its own sources into oracle/_ref (the CPU
restatement in oracle/ when that build is absent), renders the 10,252-triangle
box-and-spheres workload at 128x128 with 4096 samples per pixel, seed 7; each
pixel is accumulated in binary64 in sample order (reference
src/render.cpp:390-415).  The GPU test renders the same workload with another
seed and checks relMSE = mean((I - R)^2 / (R^2 + 1e-2)) against it.

    python tests/golden/make_converged.py      # writes tests/golden/converged_c1.npz
"""
from __future__ import annotations

import os
import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT / "tests"))
sys.path.insert(0, str(ROOT / "oracle"))

WIDTH = HEIGHT = 128
SPP = 4096
SEED = 7


def main():
    import conftest  # noqa: F401  (loads the package as cr_amd)
    from conftest import scenes
    import probe
    kind = "ref" if probe.available("ref") else "orc"
    p = probe.Probe(kind)
    w = scenes.config1(width=WIDTH, height=HEIGHT, samples=SPP)
    ps, pe = p.scene(w.scene), p.env(w.env)
    t = time.time()
    film, st = ps.render(pe, w.camera, w.options(samples=SPP, seed=SEED, threads=len(os.sched_getaffinity(0))))
    print(f"{kind}: {WIDTH}x{HEIGHT}x{SPP} in {time.time() - t:.1f}s, mean {film.mean():.6f}")
    np.savez_compressed(HERE / "converged_c1.npz", film=film.astype(np.float32),
                        meta=np.array([WIDTH, HEIGHT, SPP, SEED], np.int64),
                        source=np.array(["reference" if kind == "ref" else "oracle"]))


if __name__ == "__main__":
    main()
