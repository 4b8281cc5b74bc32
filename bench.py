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

Two systems forms are measured against the plain one-render-at-a-time path on
the same request mix (films compared bit for bit):
  batched: 64 seeded requests (terrain / spheres / soup of 1k..2M triangles,
           128^2..1024^2 px, 4..256 spp, 8..16 bounces) rendered in one shared
           wavefront after autotuning the launch configuration;
  served:  200 requests arriving as a Poisson stream at 20 req/s over those 64
           resident scenes (sizes drawn from the same ranges), continuously
           batched; p50/p99 time-to-image and aggregate Msamples/s.  The plain
           path's latencies are those of a FIFO single-render server with the
           measured plain render times under the same arrivals.
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import pathlib
import sys
import threading
import time

ROOT = pathlib.Path(__file__).resolve().parent
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
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


def _compare(gpu_film, cpu_film) -> dict:
    """Per-pixel agreement of the device's film with the reference's on the same samples."""
    import numpy as np
    a, r = gpu_film.astype(np.float64), cpu_film.astype(np.float64)
    d = np.abs(a - r)
    ok = np.all((d <= 1e-4 * np.abs(r)) | (d <= 1e-7), axis=-1)
    exact = np.all(gpu_film.view(np.uint32) == cpu_film.view(np.uint32), axis=-1)
    return {"pixels": int(ok.size), "within_1e-4": round(float(ok.mean()), 6),
            "bit_exact": round(float(exact.mean()), 6), "diverged_pixels": int((~ok).sum())}


def cpu_baseline(ctx, sid, eid, w, budget_s: float, threads: int, max_spp: int = 64) -> dict:
    """Time the reference CPU renderer on full-frame renders of fewer samples (its
    pixel/sample streams are the first samples of the GPU run's), compare its film with
    the device's render of the same samples, and time it once more at the reference's
    default thread count (threads = 0: half the logical CPUs, reference src/render.cpp:363)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import probe
    kind = "reference" if probe.available("ref") else "port"
    p = probe.Probe("ref" if kind == "reference" else "orc")
    t0 = time.perf_counter()
    ps = p.scene(w.scene)
    scene_s = time.perf_counter() - t0
    pe = p.env(w.env)
    log(f"cpu baseline ({kind}): scene finalize + BVH build in {scene_s:.2f}s")
    spp = 1
    best = None
    while True:
        opt = w.options(samples=spp, threads=threads)
        film, st = ps.render(pe, w.camera, opt)
        best = (w.width * w.height * spp, st.seconds, spp, film)
        log(f"cpu baseline: {w.width}x{w.height}x{spp} in {st.seconds:.2f}s")
        if st.seconds * 4 > budget_s or spp * 4 > min(w.samples, max_spp):
            break
        spp *= 4
    samples, secs, spp, film = best
    g = ctx.render(sid, eid, w.camera, w.options(samples=spp))
    rays = g.extend_rays + g.shadow_rays
    parity = _compare(g.pixels, film)
    log(f"cpu baseline: device film at {spp} spp vs reference: {parity}")
    d_spp = max(1, spp // 4)
    _, dst = ps.render(pe, w.camera, w.options(samples=d_spp, threads=0))
    default = {"threads": "0 (reference default: hardware_concurrency / 2)", "workers": int(dst.workers),
               "msamples_per_s": round(w.width * w.height * d_spp / dst.seconds / 1e6, 4),
               "sample": f"full frame at {d_spp} spp"}
    log(f"cpu baseline at the reference default thread count: {default}")
    return {"value": round(samples / secs / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": kind,
            "cores_meaning": "logical CPUs in this process's affinity mask, one render thread each",
            "sample": f"{w.width}x{w.height} full frame at {spp} spp (pixel/sample streams identical to the "
                      f"GPU run's first {spp} samples), render() time excluding BVH build",
            "mrays_per_s": round(rays / secs / 1e6, 3),
            "mrays_note": "extend + shadow queries of these samples (replays included), counted by the device; "
                          "the reference issues the same queries",
            "reference_default_threads": default,
            "film_vs_device": parity,
            "scene_seconds": round(scene_s, 3)}


def film_encode_form(ctx, film, with_reference: bool) -> dict:
    """io::writePng + io::writeHdr of the rendered film (reference src/io/image.cpp:201-236):
    the device encoders (PCIe transfers included), the same with the files written
    (zlib deflate for the PNG on the host), and the reference's own writers on the
    host (stb, single thread) with the outputs compared."""
    import tempfile
    t = time.perf_counter()
    png8 = ctx.film_png8(film, 0.0)
    hdr = ctx.film_hdr(film)
    enc_s = time.perf_counter() - t
    out = {"pixels": int(film.shape[0] * film.shape[1]), "encode_seconds": round(enc_s, 4)}
    with tempfile.TemporaryDirectory() as d:
        d = pathlib.Path(d)
        t = time.perf_counter()
        ctx.write_png(d / "a.png", film, 0.0)
        ctx.write_hdr(d / "a.hdr", film)
        out["files_seconds"] = round(time.perf_counter() - t, 4)
        if with_reference:
            sys.path.insert(0, str(ROOT / "oracle"))
            import probe
            if probe.available("ref"):
                p = probe.Probe("ref")
                t = time.perf_counter()
                p.film_write(film, 0.0, d / "r.png", d / "r.hdr")
                out["reference_seconds"] = round(time.perf_counter() - t, 4)
                out["hdr_bytes_identical"] = (d / "r.hdr").read_bytes() == hdr
                import numpy as np
                out["png_pixels_identical"] = bool(np.array_equal(p.film_png8(film, 0.0), png8))
    return out


def scene_cache_form(ctx, w) -> dict:
    """Making the workload's scene resident from a scene cache file (device-format arrays,
    no finalize or BVH build) against building it: the served form's time to residency."""
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = pathlib.Path(d) / "scene.crgscene"
        sid, info = ctx.upload_scene(w.scene)
        t = time.perf_counter()
        ctx.save_scene(sid, path)
        save_s = time.perf_counter() - t
        ctx.release_scene(sid)
        t = time.perf_counter()
        lid, linfo = ctx.load_scene(path)
        load_s = time.perf_counter() - t
        ctx.release_scene(lid)
        return {"bytes": path.stat().st_size, "save_seconds": round(save_s, 3), "load_seconds": round(load_s, 3),
                "build_and_upload_seconds": round(info.build_seconds + info.upload_seconds, 3)}


HARBOR = ROOT / "tests" / "golden" / "large" / "harbor.npz"


def real_asset_form(ctx, with_cpu: bool, steps: int = 3) -> dict:
    """The reference's harbor glTF asset (358,440 triangles, 53 materials, 14 textures) as the
    reference's own loader loads it (a fixture written where the reference tree exists; this
    form is skipped elsewhere), rendered at the reference's default options (960x540, 32 spp,
    8 bounces) under the synthetic sky: time to residency (finalize + build + upload, and from a
    scene cache file) and render throughput, with the reference CPU renderer beside it."""
    if not HARBOR.exists():
        return None
    import tempfile
    from cr_amd import abi, scenes
    scene, cam = scenes.scene_fixture(HARBOR)
    env = abi.EnvData(rgba=scenes.synthetic_sky())
    opt = abi.options(960, 540, 32, 8, seed=42)
    t = time.perf_counter()
    sid, info = ctx.upload_scene(scene)
    resident_s = time.perf_counter() - t
    eid, _ = ctx.upload_env(env)
    with tempfile.TemporaryDirectory() as d:
        path = pathlib.Path(d) / "harbor.crgscene"
        ctx.save_scene(sid, path)
        t = time.perf_counter()
        lid, _ = ctx.load_scene(path)
        load_s = time.perf_counter() - t
        ctx.release_scene(lid)
    ctx.render(sid, eid, cam, opt)
    t = time.perf_counter()
    rays = 0
    for _ in range(steps):
        r = ctx.render(sid, eid, cam, opt)
        rays += r.extend_rays + r.shadow_rays
    secs = (time.perf_counter() - t) / steps
    samples = opt.width * opt.height * opt.samples
    out = {"triangles": int(info.triangles), "images": len(scene.images), "device_mib": round(info.device_bytes / 2**20, 1),
           "resident_seconds": round(resident_s, 3), "host_build_seconds": round(info.build_seconds, 3),
           "cache_load_seconds": round(load_s, 3), "render": "960x540, 32 spp, 8 bounces, synthetic sky",
           "seconds_per_render": round(secs, 4), "msamples_per_s": round(samples / secs / 1e6, 2),
           "mrays_per_s": round(rays / steps / secs / 1e6, 2)}
    ctx.release_scene(sid)
    ctx.release_env(eid)
    if with_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import probe
        if probe.available("ref"):
            p = probe.Probe("ref")
            t = time.perf_counter()
            ps = p.scene(scene)
            ref_scene_s = time.perf_counter() - t
            small = abi.options(960, 540, 1, 8, seed=42, threads=len(os.sched_getaffinity(0)))
            _, st = ps.render(p.env(env), cam, small)
            out["reference"] = {"finalize_and_build_seconds": round(ref_scene_s, 3),
                                "msamples_per_s": round(960 * 540 / st.seconds / 1e6, 4),
                                "sample": "960x540 at 1 spp (the same first sample of every pixel)",
                                "cores": small.threads}
            out["gpu_vs_cpu"] = round(out["msamples_per_s"] / out["reference"]["msamples_per_s"], 1)
    return out


def _pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q))


def quality(ctx) -> dict:
    """relMSE against the reference's converged film of the box-and-spheres workload
    (128x128, 4096 spp, other seed, binary64 accumulation): the spp at which the GPU
    render reaches relMSE <= 1e-3 and the throughput there."""
    import numpy as np
    from cr_amd import scenes
    f = ROOT / "tests" / "golden" / "converged_c1.npz"
    if not f.exists():
        return None
    g = np.load(f, allow_pickle=False)
    width, height, spp_ref, seed_ref = (int(x) for x in g["meta"])
    ref = g["film"].astype(np.float64)
    w = scenes.config1(width=width, height=height)
    sid, _ = ctx.upload_scene(w.scene)
    eid, _ = ctx.upload_env(w.env)
    ladder, hit = {}, None
    try:
        for spp in (64, 256, 1024, 2048, 4096):
            r = ctx.render(sid, eid, w.camera, w.options(samples=spp, seed=0))
            img = r.pixels.astype(np.float64)
            e = float(np.mean((img - ref) ** 2 / (ref * ref + 1e-2)))
            ladder[spp] = round(e, 7)
            if e <= 1e-3 and hit is None:
                hit = (spp, width * height * spp / r.seconds / 1e6, (r.extend_rays + r.shadow_rays) / r.seconds / 1e6)
    finally:
        ctx.release_scene(sid)
        ctx.release_env(eid)
    return {"workload": f"config1 box+spheres {width}x{height}", "reference": f"{spp_ref} spp seed {seed_ref}",
            "relmse_by_spp": ladder, "spp_at_relmse_1e-3": hit[0] if hit else None,
            "msamples_per_s_at_target": round(hit[1], 3) if hit else None,
            "mrays_per_s_at_target": round(hit[2], 2) if hit else None}


def quality_config2(ctx, sid, eid, w) -> dict:
    """relMSE of the headline workload (config 2, seed 0) over two 64x64 windows of its 1080p
    frame (inside a dielectric sphere; sun-lit terrain), rendered with their full-frame sample
    streams, against the oracle's binary64-accumulated 32768-spp render of the same windows
    with another seed: the spp at which each window reaches relMSE <= 1e-3."""
    import numpy as np
    f = ROOT / "tests" / "golden" / "converged_c2.npz"
    if not f.exists() or w.name != "config2_terrain_spheres":
        return None
    g = np.load(f, allow_pickle=False)
    spp_ref, seed_ref = (int(x) for x in g["meta"])
    out, need = {}, 0
    for name in ("sphere", "terrain"):
        win = tuple(int(v) for v in g[f"{name}/window"])
        ref = g[f"{name}/film"].astype(np.float64)
        ladder, hit, rate = {}, None, None
        for spp in (1024, 4096, 16384, 65536):
            t = time.perf_counter()
            r = ctx.render_window(sid, eid, w.camera, w.options(samples=spp, seed=0), win)
            el = time.perf_counter() - t
            img = r.pixels.astype(np.float64)
            e = float(np.mean((img - ref) ** 2 / (ref * ref + 1e-2)))
            ladder[spp] = round(e, 7)
            rate = win[2] * win[3] * spp / el / 1e6
            if e <= 1e-3 and hit is None:
                hit = spp
        out[name] = {"window": list(win), "relmse_by_spp": ladder, "spp_at_relmse_1e-3": hit,
                     "window_msamples_per_s": round(rate, 2)}
        need = max(need, hit or 1 << 30)
    out["reference"] = f"oracle window render, {spp_ref} spp, seed {seed_ref}, binary64 accumulation"
    out["spp_at_relmse_1e-3"] = need if need < (1 << 30) else None
    return out


BUILDERS = 4  # host threads preparing scenes concurrently in the batched form


def batched_form(ctx, count: int) -> dict:
    """64-request batched mix.  Plain = the drop-in call per request (scene upload with
    host BVH build, render, release), one at a time; its render-only time is the sum of
    the per-request render times.  Batched = scenes uploaded by BUILDERS host threads while
    the GPU already renders the requests whose scenes are resident, all in one shared
    wavefront; render-only = the same batch re-rendered with every scene resident."""
    import numpy as np
    from cr_amd import scenes
    t0 = time.time()
    sky = scenes.synthetic_sky()
    mix = [scenes.batch_request(1000 + i, sky=sky) for i in range(count)]
    log(f"batched: {count} requests generated in {time.time() - t0:.1f}s")
    env_id, _ = ctx.upload_env(mix[0].env)  # one shared sky
    tris = sum(int(w.scene.triangle_count) for w in mix)
    samples = sum(w.camera_samples for w in mix)
    # autotune on a mid-size request at reduced spp (results never depend on the choice)
    mid = sorted(mix, key=lambda w: w.width * w.height * int(w.scene.triangle_count))[count // 2]
    sid, _ = ctx.upload_scene(mid.scene)
    tuned, rate = ctx.autotune(sid, env_id, mid.camera, mid.options(samples=16))
    ctx.release_scene(sid)
    log(f"batched: autotuned {tuned.as_dict()} ({rate / 1e6:.1f} Msamples/s on the sample)")
    # plain drop-in, one request at a time
    t = time.perf_counter()
    plain, render_s, build_s = [], 0.0, 0.0
    for w in mix:
        sid, info = ctx.upload_scene(w.scene)
        r = ctx.render(sid, env_id, w.camera, w.options())
        ctx.release_scene(sid)
        plain.append(r.pixels)
        render_s += r.seconds
        build_s += info.build_seconds + info.upload_seconds
    plain_s = time.perf_counter() - t
    log(f"batched: plain drop-in {plain_s:.2f}s (render {render_s:.2f}s, host build+upload {build_s:.2f}s)")
    # batched end to end: builder threads upload scenes (host preparation runs outside the
    # library's lock, so several overlap each other and the GPU); each request is submitted
    # as soon as its scene is resident
    resident = [None] * count
    next_i = [0]
    lock = threading.Lock()
    t = time.perf_counter()

    def builder():
        while True:
            with lock:
                i = next_i[0]
                next_i[0] += 1
            if i >= count:
                return
            resident[i] = ctx.upload_scene(mix[i].scene)[0]

    builders = [threading.Thread(target=builder) for _ in range(BUILDERS)]
    for th in builders:
        th.start()
    tickets = [None] * count
    pending = set(range(count))
    while pending:
        ready = [i for i in sorted(pending) if resident[i] is not None]
        for i in ready:
            w = mix[i]
            tickets[i] = ctx.submit(resident[i], env_id, w.camera, w.options())
            pending.discard(i)
        if pending and not ready:
            time.sleep(0.0005)
    films = [ctx.wait(tk, film).pixels for tk, film in tickets]
    for th in builders:
        th.join()
    e2e_s = time.perf_counter() - t
    same = sum(1 for a, b in zip(plain, films) if np.array_equal(a.view(np.uint32), b.view(np.uint32)))
    log(f"batched: end to end {e2e_s:.2f}s")
    # render-only: the same batch with every scene resident
    t = time.perf_counter()
    batch = ctx.render_batch([(sid, env_id, w.camera, w.options()) for w, sid in zip(mix, resident)])
    batch_s = time.perf_counter() - t
    same2 = sum(1 for a, b in zip(plain, batch)
                if not isinstance(b, Exception) and np.array_equal(a.view(np.uint32), b.pixels.view(np.uint32)))
    rays = sum(b.extend_rays + b.shadow_rays for b in batch if not isinstance(b, Exception))
    log(f"batched: resident one shared wavefront {batch_s:.2f}s")
    out = {"requests": count, "triangles_total": tris, "msamples": round(samples / 1e6, 2),
           "msamples_per_s": round(samples / e2e_s / 1e6, 3),
           "plain_msamples_per_s": round(samples / plain_s / 1e6, 3),
           "vs_plain": round(plain_s / e2e_s, 3),
           "resident_msamples_per_s": round(samples / batch_s / 1e6, 3),
           "plain_resident_msamples_per_s": round(samples / render_s / 1e6, 3),
           "vs_plain_resident": round(render_s / batch_s, 3),
           "mrays_per_s": round(rays / batch_s / 1e6, 2),
           "seconds": round(e2e_s, 3), "plain_seconds": round(plain_s, 3),
           "resident_seconds": round(batch_s, 3), "plain_render_seconds": round(render_s, 3),
           "host_build_seconds": round(build_s, 2),
           "films_bit_identical_to_plain": f"{min(same, same2)}/{count}", "tuning": tuned.as_dict()}
    return out, list(zip(mix, resident)), env_id


def served_form(ctx, catalog, env_id, count: int, rate: float) -> dict:
    """Poisson request stream over resident scenes, continuously batched.  Plain = the
    drop-in call per request (upload, render, release) one at a time; its latencies are
    those of a FIFO single-render server with the measured per-request times under the
    same arrivals (render-only variants use the render times alone)."""
    import numpy as np
    from cr_amd import scenes
    arrivals, reqs = scenes.served_stream(count, rate, len(catalog), seed=0)
    jobs = []
    for k, side, spp, bounces in reqs:
        w, sid = catalog[k]
        jobs.append((w, sid, w.options(width=side, height=side, samples=spp, bounces=bounces)))
    samples = sum(o.width * o.height * o.samples for _, _, o in jobs)
    log(f"served: {count} requests at {rate} req/s, {samples / 1e6:.1f} Msamples")
    t = time.perf_counter()
    plain_times, render_times, plain_films = [], [], []
    for w, _, opt in jobs:
        t1 = time.perf_counter()
        sid, _ = ctx.upload_scene(w.scene)
        r = ctx.render(sid, env_id, w.camera, opt)
        ctx.release_scene(sid)
        plain_times.append(time.perf_counter() - t1)
        render_times.append(r.seconds)
        plain_films.append(r.pixels)
    plain_s = time.perf_counter() - t
    render_s = sum(render_times)
    log(f"served: plain drop-in {plain_s:.2f}s (render {render_s:.2f}s)")

    def fifo(service):
        free_at, lat = 0.0, []
        for a, s_ in zip(arrivals, service):
            start = max(a, free_at)
            free_at = start + s_
            lat.append(free_at - a)
        return lat

    plain_lat, plain_res_lat = fifo(plain_times), fifo(render_times)
    tickets = [None] * count
    t_start = time.perf_counter()

    def submitter():
        for i, (w, sid, opt) in enumerate(jobs):
            delay = t_start + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[i] = ctx.submit(sid, env_id, w.camera, opt)

    th = threading.Thread(target=submitter)
    th.start()
    lat, same, last_done = [], 0, 0.0
    for i in range(count):
        while tickets[i] is None:
            time.sleep(0.001)
        tk, film = tickets[i]
        r = ctx.wait(tk, film)
        lat.append(r.latency_seconds)
        last_done = max(last_done, time.perf_counter() - t_start)
        if np.array_equal(film.view(np.uint32), plain_films[i].view(np.uint32)):
            same += 1
        plain_films[i] = None
        if (i + 1) % 50 == 0:
            log(f"served: {i + 1}/{count} done")
    th.join()
    served_s = last_done
    plain_makespan = max(a + l for a, l in zip(arrivals, plain_lat))
    plain_res_makespan = max(a + l for a, l in zip(arrivals, plain_res_lat))
    return {"requests": count, "rate_req_per_s": rate, "msamples": round(samples / 1e6, 2),
            "msamples_per_s": round(samples / served_s / 1e6, 3),
            "plain_msamples_per_s": round(samples / plain_makespan / 1e6, 3),
            "vs_plain": round(plain_makespan / served_s, 3),
            "plain_resident_msamples_per_s": round(samples / plain_res_makespan / 1e6, 3),
            "vs_plain_resident": round(plain_res_makespan / served_s, 3),
            "p50_s": round(_pct(lat, 50), 3), "p99_s": round(_pct(lat, 99), 3),
            "plain_p50_s": round(_pct(plain_lat, 50), 3), "plain_p99_s": round(_pct(plain_lat, 99), 3),
            "plain_resident_p50_s": round(_pct(plain_res_lat, 50), 3),
            "plain_resident_p99_s": round(_pct(plain_res_lat, 99), 3),
            "makespan_s": round(served_s, 3), "plain_makespan_s": round(plain_makespan, 3),
            "films_bit_identical_to_plain": f"{same}/{count}",
            "scenes": f"{len(catalog)} resident; plain uploads (host BVH build) per request"}


def measure_plain(ctx, w, steps: int, warmup: int, cpu_budget) -> dict:
    """The plain render path on one workload: `steps` timed renders after `warmup`, work per
    query from a counting render, per-kernel device times (HIP events on the library's stream,
    kernels launched one by one) with the traversal and shading kernels against the HBM
    roofline, and the reference CPU renderer beside it."""
    sid, info = ctx.upload_scene(w.scene)
    eid, env_bytes = ctx.upload_env(w.env)
    log(f"{w.name} resident: {info.triangles} tris, {info.wide_nodes} wide nodes, "
        f"{(info.device_bytes + env_bytes) / 2**20:.1f} MiB, host build {info.build_seconds:.2f}s, "
        f"upload {info.upload_seconds:.2f}s")
    opt = w.options()
    for i in range(warmup):
        r = ctx.render(sid, eid, w.camera, opt)
        log(f"{w.name} warmup {i}: {r.seconds:.3f}s")
    rays, res = 0, None
    t_start = time.perf_counter()
    for i in range(steps):
        t = time.perf_counter()
        res = ctx.render(sid, eid, w.camera, opt)
        rays += res.extend_rays + res.shadow_rays
        log(f"{w.name} step {i}: {time.perf_counter() - t:.3f}s")
    elapsed = time.perf_counter() - t_start
    sec_per_step = elapsed / steps
    samples = w.camera_samples
    value = samples / sec_per_step / 1e6

    # Work per query from a render with the counting traversal kernels (same seed: the
    # same queries), then per-kernel device time of the timed kernels (no counters) from
    # HIP events on the library stream, kernels launched one by one.
    ctx.set_work_counters(True)
    cr_ = ctx.render(sid, eid, w.camera, opt)
    ctx.set_work_counters(False)
    ctx.set_profiling(True)
    pr = ctx.render(sid, eid, w.camera, opt)
    kt = ctx.kernel_times()
    ctx.set_profiling(False)
    assert (cr_.extend_rays, cr_.shadow_rays) == (pr.extend_rays, pr.shadow_rays)
    # Algorithmic bytes per query as the survey defines them: 48 (ray in, hit out) +
    # 128 per wide-node visit + 48 per triangle test.  The device's 4-wide nodes are
    # 64-byte quantized records, so the bytes it actually needs per visit are half that
    # figure (node_bytes below).
    def traversal_line(kernel, rays_, nodes, tris):
        launches, ms = kt[kernel]
        n_node, n_tri = nodes / max(1, rays_), tris / max(1, rays_)
        per = 48 + 128 * n_node + 48 * n_tri
        ach = per * rays_ / (ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(ach / PEAK_HBM_GBS, 4), "kernel": kernel,
                "avg_launch_ms": round(ms / max(1, launches), 4), "launches": launches,
                "bytes_per_ray": round(per, 1), "nodes_per_ray": round(n_node, 2), "tris_per_ray": round(n_tri, 2),
                "node_bytes": 64, "frac_at_node_bytes": round((per - 64 * n_node) * rays_ / (ms * 1e-3) / 1e9 /
                                                              PEAK_HBM_GBS, 4)}
    roofline = traversal_line("k_extend", pr.extend_rays, cr_.extend_nodes, cr_.extend_tris)
    roofline["traffic"] = _traffic(w.name)
    shadow_line = traversal_line("k_shadow", pr.shadow_rays, cr_.shadow_nodes, cr_.shadow_tris)
    # Shading: per shaded path the ray record (64 B) and state record (128 B) read and
    # written back, the primitive's traversal (48 B), shading (64 B) and normal (16 B)
    # records, and a 72-byte shadow record written: 584 B (misses read fewer).
    sh_launches, sh_ms = kt["k_shade"]
    sh_ach = 584 * pr.extend_rays / (sh_ms * 1e-3) / 1e9
    shade_line = {"bound": "latency of per-path scattered reads, shown against hbm", "achieved": round(sh_ach, 1),
                  "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(sh_ach / PEAK_HBM_GBS, 4), "kernel": "k_shade",
                  "avg_launch_ms": round(sh_ms / max(1, sh_launches), 4), "bytes_per_path": 584}
    n_node = cr_.extend_nodes / max(1, cr_.extend_rays)
    n_tri = cr_.extend_tris / max(1, cr_.extend_rays)
    ext_ms = kt["k_extend"][1]
    # The same kernel against the FP32 vector roofline: 12 FLOP per child slab test of a
    # 4-wide node and 40 per FP32 triangle test (6 FMA per child slab, the watertight test's operations).
    flop_per_ray = 12 * 4 * n_node + 40 * n_tri
    tflops = flop_per_ray * pr.extend_rays / (ext_ms * 1e-3) / 1e12
    valu = {"achieved": round(tflops, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / PEAK_FP32_TFLOPS, 5), "flop_per_ray": round(flop_per_ray, 1), "kernel": "k_extend"}
    kernels = {k: {"launches": v[0], "ms": round(v[1], 2)} for k, v in kt.items()}
    log(f"{w.name} kernels: {kernels}")
    cpu = None
    if cpu_budget is not None:
        cpu = cpu_baseline(ctx, sid, eid, w, cpu_budget, len(os.sched_getaffinity(0)))
    qual2 = quality_config2(ctx, sid, eid, w)
    if qual2:
        log(f"{w.name} quality: {qual2}")
    ctx.release_scene(sid)
    ctx.release_env(eid)
    return {"value": round(value, 3), "ms_per_step": round(sec_per_step * 1e3, 2), "steps": steps,
            "mrays_per_s": round(rays / elapsed / 1e6, 2),
            "rays_per_sample": round((res.extend_rays + res.shadow_rays) / samples, 3),
            "exact_fallbacks_per_mray": round(pr.exact_fallbacks / max(1, pr.extend_rays + pr.shadow_rays) * 1e6, 2),
            "exact_rays_per_mray": round(pr.exact_rays / max(1, pr.extend_rays + pr.shadow_rays) * 1e6, 2),
            "roofline": roofline, "roofline_k_shadow": shadow_line, "roofline_k_shade": shade_line,
            "valu_roofline": valu, "kernels_ms": kernels, "cpu_baseline": cpu, "quality": qual2,
            "config": {"workload": w.name, "triangles": int(info.triangles), "width": w.width, "height": w.height,
                       "spp": w.samples, "bounces": w.bounces, "seed": w.seed, "env": "synthetic 1024x512 HDR sky",
                       "bvh_width": ctx.get_tuning().bvh_width, "tuning": ctx.get_tuning().as_dict()},
            "_info": info, "_result": res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline rendering")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-config3", dest="config3", action="store_false",
                    help="skip the config-3 (8M-triangle divergence stress) block")
    ap.add_argument("--forms", default="plain,batched,served", help="comma list of plain,batched,served")
    ap.add_argument("--batch", type=int, default=64, help="requests in the batched mix")
    ap.add_argument("--served", type=int, default=200, help="requests in the served stream")
    ap.add_argument("--rate", type=float, default=20.0, help="served arrival rate (req/s)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="one render of the workload (kernels launched one by one) plus a 1 GiB device copy, "
                         "for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs")
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args.workload)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    _package()
    from cr_amd import crgpu
    forms = set(args.forms.split(","))

    t0 = time.time()
    w = workload(args.workload)
    log(f"workload {w.name}: {w.scene.triangle_count} tris generated in {time.time() - t0:.1f}s")
    ctx = crgpu.Context(0)
    default_tuning = ctx.get_tuning()  # the plain renders run with the library's default configuration
    plain = measure_plain(ctx, w, args.steps, args.warmup, None if args.no_cpu else args.cpu_budget)
    info, res = plain.pop("_info"), plain.pop("_result")
    value, sec_per_step = plain["value"], plain["ms_per_step"] / 1e3

    qual = {"config2": plain.pop("quality"), "config1": quality(ctx)}
    log(f"quality config1: {qual['config1']}")

    batched = served = None
    if "batched" in forms or "served" in forms:
        batched, catalog, env_id = batched_form(ctx, args.batch)
        log(f"batched: {batched}")
        if "served" in forms:
            served = served_form(ctx, catalog, env_id, args.served, args.rate)
            log(f"served: {served}")
        if "batched" not in forms:
            batched = None

    film_enc = film_encode_form(ctx, res.pixels, not args.no_cpu)
    log(f"film encode: {film_enc}")
    cache = scene_cache_form(ctx, w)
    log(f"scene cache: {cache}")
    asset = real_asset_form(ctx, not args.no_cpu)
    log(f"real asset: {asset}")

    cpu = plain.get("cpu_baseline")
    host_build = {"triangles": int(info.triangles), "seconds": round(info.build_seconds, 3),
                  "what": "finalize + 4-wide SAH build + quantized 4- and 8-wide device trees on all host threads"}
    if cpu:
        ref_s = cpu.pop("scene_seconds")
        host_build.update({"reference_seconds": ref_s,
                           "reference_what": f"the CPU baseline's Scene::finalize + Bvh::build ({cpu['kind']})",
                           "speedup": round(ref_s / max(info.build_seconds, 1e-9), 2)})
    config3 = None
    if args.config3 and args.workload != "c3":
        t0 = time.time()
        w3 = workload("c3")
        log(f"workload {w3.name}: {w3.scene.triangle_count} tris generated in {time.time() - t0:.1f}s")
        ctx.set_tuning(default_tuning)  # the batched form left its autotuned configuration applied
        config3 = measure_plain(ctx, w3, min(args.steps, 3), 1, None if args.no_cpu else args.cpu_budget)
        config3.pop("_result")
        config3.pop("_info")
        if config3["cpu_baseline"]:
            config3["cpu_baseline"].pop("scene_seconds")
        log(f"config3: {config3}")

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(sec_per_step * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": plain["config"],
        "mrays_per_s": plain["mrays_per_s"],
        "rays_per_sample": plain["rays_per_sample"],
        "exact_fallbacks_per_mray": plain["exact_fallbacks_per_mray"],
        "exact_rays_per_mray": plain["exact_rays_per_mray"],
        "roofline": plain["roofline"],
        "roofline_k_shadow": plain["roofline_k_shadow"],
        "roofline_k_shade": plain["roofline_k_shade"],
        "valu_roofline": plain["valu_roofline"],
        "kernels_ms": plain["kernels_ms"],
        "quality": qual,
        "batched": batched,
        "served": served,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
        "host_build": host_build,
        "film_encode": film_enc,
        "scene_cache": cache,
        "real_asset_harbor": asset,
        "config3": config3,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def traffic_probe(name: str):
    """Counter-collection workload: the copy kernel of known size calibrates the
    HBM byte counters (bytes read = bytes written = 1 GiB)."""
    _package()
    import torch
    from cr_amd import crgpu
    w = workload(name)
    ctx = crgpu.Context(0)
    sid, _ = ctx.upload_scene(w.scene)
    eid, _ = ctx.upload_env(w.env)
    ctx.set_work_counters(True)  # work per query (its own render; the counted kernels are not the timed ones)
    c = ctx.render(sid, eid, w.camera, w.options())
    ctx.set_work_counters(False)
    ctx.set_profiling(True)  # individual launches (no graph) so every dispatch is counted
    r = ctx.render(sid, eid, w.camera, w.options())
    ctx.set_profiling(False)
    kt = ctx.kernel_times()
    a = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"workload": w.name, "extend_rays": r.extend_rays, "extend_nodes": c.extend_nodes,
                      "extend_tris": c.extend_tris, "shadow_rays": r.shadow_rays, "shadow_nodes": c.shadow_nodes,
                      "shadow_tris": c.shadow_tris, "k_extend_launches": kt["k_extend"][0],
                      "copy_bytes": 1 << 30}), flush=True)
    ctx.close()


def _traffic(workload_name: str):
    """HBM bytes per k_extend launch from the committed rocprofv3 counter summary, if any
    (a line `bytes_per_launch: N` in profiles/traffic_k_extend_<workload>.md)."""
    f = ROOT / "profiles" / f"traffic_k_extend_{workload_name}.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.strip().startswith("bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


if __name__ == "__main__":
    main()
