"""Batched (C4) and served (C5) forms pinned to the reference renderer.

Requests of the bench's mix generator (terrain, spheres and soup scenes of the
64-request mix's seeds, capped at 50k triangles) were rendered by the
reference's own code (tests/golden/render_batch.npz).  Each request is
rendered on the device three ways — plain crg_render, all together in one
crg_render_batch, and through crg_submit/crg_wait with staggered submissions
while others are in flight — and each film must agree with the reference's
(per-pixel 1e-4 rule, equal precision-replay counts) and be bit-identical
across the three forms.  The reference's own determinism test compares its
renders across worker counts (reference tests/test_integrator.cpp:298-307);
the device analogue is plain vs batched vs served.

At full size, the bench's 64-request batched mix (1k..2M triangles,
128^2..1024^2 px, 4..256 spp) and its 200-request Poisson stream over those
scenes must give films bit-identical to the plain renders.
"""
from __future__ import annotations

import threading
import time

import numpy as np
import pytest

from conftest import GOLDEN, load_golden, scenes
from parity import assert_parity

pytestmark = pytest.mark.gpu


def _cases():
    import sys
    sys.path.insert(0, str(GOLDEN))
    import make_golden
    return make_golden.batch_cases()


@pytest.fixture(scope="module")
def golden():
    return load_golden("render_batch.npz")


@pytest.fixture(scope="module")
def resident(gpu, golden):
    items = []
    env = None
    for name, w, opt in _cases():
        assert w.scene.bytes_sha256() == str(golden[f"{name}/scene_sha"]), f"{name}: scene generator drifted"
        assert list(golden[f"{name}/options"]) == [opt.width, opt.height, opt.samples, opt.bounces, opt.seed]
        if env is None:
            env, _ = gpu.upload_env(w.env)  # every request shares the synthetic sky
        sid, _ = gpu.upload_scene(w.scene)
        items.append((name, w, opt, sid))
    yield items, env
    for _, _, _, sid in items:
        gpu.release_scene(sid)
    gpu.release_env(env)


@pytest.fixture(scope="module")
def plain(gpu, resident, golden):
    items, env = resident
    out = {}
    for name, w, opt, sid in items:
        r = gpu.render(sid, env, w.camera, opt)
        assert_parity(r.pixels, golden[f"{name}/film"], f"{name} plain")
        assert r.boundary_retries == int(golden[f"{name}/retries"])
        out[name] = r
    return out


def _check(name, r, plain, golden, form):
    assert not isinstance(r, Exception), f"{name} {form}: {r}"
    assert_parity(r.pixels, golden[f"{name}/film"], f"{name} {form}")
    assert r.boundary_retries == int(golden[f"{name}/retries"])
    np.testing.assert_array_equal(r.pixels.view(np.uint32), plain[name].pixels.view(np.uint32),
                                  err_msg=f"{name}: {form} film differs from the plain render")
    for k in ("camera_samples", "boundary_retries", "extend_rays", "shadow_rays"):
        assert getattr(r, k) == getattr(plain[name], k), f"{name} {form}: {k}"


def test_plain_requests_match_reference(plain):
    assert len(plain) == len(_cases())


def test_batched_requests_match_reference(gpu, resident, plain, golden):
    items, env = resident
    res = gpu.render_batch([(sid, env, w.camera, opt) for _, w, opt, sid in items])
    for (name, _, _, _), r in zip(items, res):
        _check(name, r, plain, golden, "batched")


def test_served_requests_match_reference(gpu, resident, plain, golden):
    """Staggered submissions in reverse order, each arriving while earlier ones are in flight."""
    items, env = resident
    tickets = []
    for name, w, opt, sid in reversed(items):
        tickets.append((name, gpu.submit(sid, env, w.camera, opt)))
        time.sleep(0.003)
    for name, (tk, film) in tickets:
        _check(name, gpu.wait(tk, film), plain, golden, "served")


@pytest.fixture(scope="module")
def full_mix(gpu):
    sky = scenes.synthetic_sky()
    mix = [scenes.batch_request(1000 + i, sky=sky) for i in range(64)]
    env, _ = gpu.upload_env(mix[0].env)
    ids = [gpu.upload_scene(w.scene)[0] for w in mix]
    print(f"full mix: {sum(int(w.scene.triangle_count) for w in mix)} triangles, "
          f"{sum(w.camera_samples for w in mix) / 1e6:.0f} M camera samples", flush=True)
    yield mix, ids, env
    for sid in ids:
        gpu.release_scene(sid)
    gpu.release_env(env)


def test_full_batched_mix_equals_plain(gpu, full_mix):
    """The bench's 64-request config-4 mix: one shared wavefront vs one render at a time."""
    mix, ids, env = full_mix
    t = time.perf_counter()
    plain = [gpu.render(sid, env, w.camera, w.options()) for w, sid in zip(mix, ids)]
    t_plain = time.perf_counter() - t
    t = time.perf_counter()
    batch = gpu.render_batch([(sid, env, w.camera, w.options()) for w, sid in zip(mix, ids)])
    t_batch = time.perf_counter() - t
    print(f"full mix: plain {t_plain:.2f}s, batched {t_batch:.2f}s", flush=True)
    for w, a, b in zip(mix, plain, batch):
        assert not isinstance(b, Exception), f"{w.name}: {b}"
        np.testing.assert_array_equal(a.pixels.view(np.uint32), b.pixels.view(np.uint32), err_msg=w.name)
        assert (a.boundary_retries, a.extend_rays, a.shadow_rays) == (b.boundary_retries, b.extend_rays, b.shadow_rays)


def test_full_served_stream_equals_plain(gpu, full_mix):
    """The bench's config-5 stream: 200 requests arriving as a Poisson process at 20 req/s over
    the 64 resident scenes, continuously batched; each film equals its plain render."""
    mix, ids, env = full_mix
    arrivals, reqs = scenes.served_stream(200, 20.0, len(mix), seed=0)
    jobs = [(mix[k], ids[k], mix[k].options(width=side, height=side, samples=spp, bounces=b))
            for k, side, spp, b in reqs]
    tickets = [None] * len(jobs)
    t0 = time.perf_counter()

    def submitter():
        for i, (w, sid, opt) in enumerate(jobs):
            delay = t0 + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[i] = gpu.submit(sid, env, w.camera, opt)

    th = threading.Thread(target=submitter)
    th.start()
    served = []
    for i in range(len(jobs)):
        while tickets[i] is None:
            time.sleep(0.001)
        tk, film = tickets[i]
        r = gpu.wait(tk, film)
        served.append((film.copy(), r.boundary_retries, r.extend_rays, r.shadow_rays, r.latency_seconds))
    th.join()
    lat = [s[4] for s in served]
    print(f"served stream: makespan {time.perf_counter() - t0:.2f}s p50 {np.percentile(lat, 50):.3f}s "
          f"p99 {np.percentile(lat, 99):.3f}s", flush=True)
    for i, (w, sid, opt) in enumerate(jobs):
        a = gpu.render(sid, env, w.camera, opt)
        film, retries, ext, shd, _ = served[i]
        np.testing.assert_array_equal(a.pixels.view(np.uint32), film.view(np.uint32), err_msg=f"request {i} {w.name}")
        assert (a.boundary_retries, a.extend_rays, a.shadow_rays) == (retries, ext, shd), f"request {i}"
