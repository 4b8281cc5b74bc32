#!/usr/bin/env python3
"""Benchmark of the gfx950 wavefront path tracer (one JSON line on stdout).

Workload (default): the plain render path at scale — fBm heightfield terrain of
1024x1024 vertices (2,093,058 triangles) plus 200 GGX/dielectric level-4
icospheres (1,024,000 triangles), synthetic 1024x512 HDR sky with a 1e4 sun
lobe, 1920x1080, 64 samples per pixel, 16 bounces, seed 0.  One step is one
complete render() of that frame with the scene, BVH and environment already
resident in HBM.

Reported: Msamples/s (value) and Mrays/s of the whole render; the extend
kernel's achieved algorithmic bandwidth against the 8 TB/s HBM roofline from
a profiled render (HIP events on the library's stream); and the reference CPU
renderer (compiled from the reference's own sources, or the CPU oracle when
that build is absent) timed on this host's cores on a bounded sample of the
same workload (full frame, fewer samples per pixel).
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
PEAK_HBM_GBS = 8000.0
METRIC = "Mrays/s & Msamples/s at relMSE≤1e-3; batched/served vs plain; p50/p99 latency"


def _package():
    if "cr_amd" not in sys.modules:
        pkg = ROOT / "cpu-renderer_amd"
        spec = importlib.util.spec_from_file_location("cr_amd", pkg / "__init__.py",
                                                      submodule_search_locations=[str(pkg)])
        mod = importlib.util.module_from_spec(spec)
        sys.modules["cr_amd"] = mod
        spec.loader.exec_module(mod)
    return sys.modules["cr_amd"]


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload(name: str):
    from cr_amd import scenes
    if name == "c2":
        return scenes.config2()
    if name == "c1":
        return scenes.config1()
    if name == "c3":
        return scenes.config3()
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(w, budget_s: float, threads: int) -> dict:
    """Time the reference CPU renderer on full-frame renders of fewer samples."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import probe
    kind = "reference" if probe.available("ref") else "port"
    p = probe.Probe("ref" if kind == "reference" else "orc")
    t0 = time.time()
    ps, pe = p.scene(w.scene), p.env(w.env)
    log(f"cpu baseline ({kind}): scene prepared in {time.time() - t0:.1f}s on {threads} threads")
    spp = 1
    best = None
    while True:
        opt = w.options(samples=spp, threads=threads)
        film, st = ps.render(pe, w.camera, opt)
        samples = w.width * w.height * spp
        best = (samples, st.seconds, spp)
        log(f"cpu baseline: {w.width}x{w.height}x{spp} in {st.seconds:.2f}s")
        if st.seconds * 4 > budget_s or spp * 4 > w.samples:
            break
        spp *= 4
    samples, secs, spp = best
    return {"value": round(samples / secs / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": kind,
            "sample": f"{w.width}x{w.height} full frame at {spp} spp (pixel/sample streams identical to the "
                      f"GPU run's first {spp} samples), render() time excluding BVH build"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline rendering")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    _package()
    from cr_amd import crgpu

    t0 = time.time()
    w = workload(args.workload)
    log(f"workload {w.name}: {w.scene.triangle_count} tris generated in {time.time() - t0:.1f}s")
    ctx = crgpu.Context(0)
    sid, info = ctx.upload_scene(w.scene)
    eid, env_bytes = ctx.upload_env(w.env)
    log(f"scene resident: {info.triangles} tris, {info.wide_nodes} wide nodes, "
        f"{(info.device_bytes + env_bytes) / 2**20:.1f} MiB, host build {info.build_seconds:.2f}s, "
        f"upload {info.upload_seconds:.2f}s")
    opt = w.options()
    for i in range(args.warmup):
        r = ctx.render(sid, eid, w.camera, opt)
        log(f"warmup {i}: {r.seconds:.3f}s")
    times, rays, res = [], 0, None
    t_start = time.perf_counter()
    for i in range(args.steps):
        t = time.perf_counter()
        res = ctx.render(sid, eid, w.camera, opt)
        times.append(time.perf_counter() - t)
        rays += res.extend_rays + res.shadow_rays
        log(f"step {i}: {times[-1]:.3f}s")
    elapsed = time.perf_counter() - t_start
    sec_per_step = elapsed / args.steps
    samples = w.camera_samples
    value = samples / sec_per_step / 1e6

    # Profiled render: per-kernel device time from HIP events on the library stream.
    ctx.set_profiling(True)
    pr = ctx.render(sid, eid, w.camera, opt)
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    ext_launches, ext_ms = kt["k_extend"]
    n_node = pr.extend_nodes / max(1, pr.extend_rays)
    n_tri = pr.extend_tris / max(1, pr.extend_rays)
    bytes_per_ray = 48 + 128 * n_node + 48 * n_tri
    achieved = bytes_per_ray * pr.extend_rays / (ext_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None,
                "kernel": "k_extend", "avg_launch_ms": round(ext_ms / max(1, ext_launches), 4),
                "launches": ext_launches, "bytes_per_ray": round(bytes_per_ray, 1),
                "nodes_per_ray": round(n_node, 2), "tris_per_ray": round(n_tri, 2)}
    kernels = {k: {"launches": v[0], "ms": round(v[1], 2)} for k, v in kt.items()}
    log(f"kernels: {kernels}")

    cpu = None
    if not args.no_cpu:
        threads = len(os.sched_getaffinity(0))
        cpu = cpu_baseline(w, args.cpu_budget, threads)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(sec_per_step * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "triangles": int(info.triangles), "width": w.width, "height": w.height,
                   "spp": w.samples, "bounces": w.bounces, "seed": w.seed, "env": "synthetic 1024x512 HDR sky"},
        "mrays_per_s": round(rays / elapsed / 1e6, 2),
        "rays_per_sample": round((res.extend_rays + res.shadow_rays) / samples, 3),
        "exact_fallbacks_per_mray": round(pr.exact_fallbacks / max(1, pr.extend_rays + pr.shadow_rays) * 1e6, 2),
        "exact_rays_per_mray": round(pr.exact_rays / max(1, pr.extend_rays + pr.shadow_rays) * 1e6, 2),
        "roofline": roofline,
        "kernels_ms": kernels,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
