"""Converged reference film for the relMSE test (written by the reference CPU renderer).

The reference renderer, compiled from its own sources into oracle/_ref (the CPU
restatement in oracle/ when that build is absent), renders the 10,252-triangle
box-and-spheres workload at 128x128 with 4096 samples per pixel, seed 7; each
pixel is accumulated in binary64 in sample order (reference
src/render.cpp:390-415).  The GPU test renders the same workload with another
seed and checks relMSE = mean((I - R)^2 / (R^2 + 1e-2)) against it.

    python tests/golden/make_converged.py      # writes tests/golden/converged_c1.npz
    python tests/golden/make_converged.py c2   # writes tests/golden/converged_c2.npz
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


# Config 2 (3,117,058 triangles, 1920x1080, synthetic HDR sky with a 1e4 sun):
# two 64x64 windows of the full frame, rendered with their frame pixel indices
# (the full render's sample streams) by the oracle's window render, which the
# CPU tests pin bit for bit against crops of the reference's own full-frame
# films.  Window A lies inside the largest dielectric sphere near the image
# centre (refraction, internal reflection, Beer-Lambert), window B on sun-lit
# terrain.  The reference's render() has no window, so a 32768-spp crop of a
# 1080p frame could not be rendered with it in bounded time.
C2_WINDOWS = {"sphere": (1291, 916, 64, 64), "terrain": (800, 820, 64, 64)}
C2_SPP = 32768
C2_SEED = 7


def main_c2():
    import conftest  # noqa: F401
    from conftest import scenes
    import probe
    p = probe.Probe("orc")
    w = scenes.config2()
    ps, pe = p.scene(w.scene), p.env(w.env)
    out = {"meta": np.array([C2_SPP, C2_SEED], np.int64), "source": np.array(["oracle window render"])}
    for name, win in C2_WINDOWS.items():
        t = time.time()
        film, st = ps.render_window(pe, w.camera, w.options(samples=C2_SPP, seed=C2_SEED,
                                                            threads=len(os.sched_getaffinity(0))), win)
        print(f"config2 {name} {win}: {C2_SPP} spp in {time.time() - t:.1f}s, mean {film.mean():.6f}", flush=True)
        out[f"{name}/film"] = film.astype(np.float32)
        out[f"{name}/window"] = np.array(win, np.int64)
    np.savez_compressed(HERE / "converged_c2.npz", **out)


if __name__ == "__main__":
    if "c2" in sys.argv[1:]:
        main_c2()
    else:
        main()
