"""The CPU oracle (oracle/cr_oracle.cpp) against the reference's golden vectors.

The oracle is a restatement in native binary64, so every golden value is
expected bit-for-bit: random streams, camera rays, closest/any hits with
binary64 t/u/v, offsetRay, surfaceAt, BSDF samples/evaluations, environment
sampling/pdf/radiance and whole films (incl. precision-replay counts).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN, abi, load_golden
import sys

sys.path.insert(0, str(GOLDEN))
import make_golden  # noqa: E402
import probe  # noqa: E402


def test_rng_and_camera(oracle):
    g = load_golden("rng_camera.npz")
    for s, p, q, ss, dr, end in zip(g["seeds"], g["pixels"], g["samples"], g["sample_seed"], g["draws"], g["ends"]):
        assert oracle.sample_seed(int(s), int(p), int(q)) == int(ss)
        u, e = oracle.rng_draw(int(ss), 8)
        np.testing.assert_array_equal(u, dr)
        assert e == int(end)
    cams = [make_golden.scenes.config1().camera, make_golden.scenes.config3(triangles=10).camera,
            make_golden.scenes.look_at((3, 4, 5), (0, 0.5, 0), 0.8), make_golden.rs.test_camera(),
            make_golden.rs.touching_boundary()[2], abi.camera(yfov=np.pi / 2, near_clip=1, far_clip=10)]
    for ci, px, py, w, h, ok, ray in zip(g["cam_index"], g["px"], g["py"], g["w"], g["h"], g["ok"], g["rays"]):
        r, err = oracle.camera_generate(cams[ci], float(px), float(py), int(w), int(h))
        assert (r is not None) == bool(ok), err
        if r is not None:
            assert bytes(r) == ray.tobytes()


@pytest.mark.parametrize("name", [s[0] for s in make_golden.trace_sets()])
def test_trace(oracle, name):
    g = load_golden("trace.npz")
    _, scene, rays, stored = next(s for s in make_golden.trace_sets() if s[0] == name)
    assert make_golden.sha(rays) == str(g[f"{name}/rays_sha"]), "ray generator drifted"
    assert scene.bytes_sha256() == str(g[f"{name}/scene_sha"]), "scene generator drifted"
    ps = oracle.scene(scene)
    assert ps.triangles == int(g[f"{name}/triangles"])
    assert ps.nodes == int(g[f"{name}/nodes"])
    h = ps.trace(rays)
    for k in ("hit", "triangle"):
        np.testing.assert_array_equal(h[k], g[f"{name}/{k}"])
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(h[k].view(np.uint64), g[f"{name}/{k}"].view(np.uint64))
    np.testing.assert_array_equal(ps.trace(rays, True)["hit"], g[f"{name}/any"])


def test_offset_and_surface(oracle):
    g = load_golden("offset_surface.npz")
    for p, n, d, m, ray in zip(g["p"], g["n"], g["d"], g["m"], g["ray"]):
        r, err = oracle.offset_ray(p, n, d, float(m))
        assert r is not None, err
        assert bytes(r) == ray.tobytes()
    c1 = make_golden.scenes.config1().scene
    rays = make_golden.scenes.random_rays(c1, 3000, seed=1)
    ps = oracle.scene(c1)
    hits = ps.trace(rays)
    for i, surf in zip(g["surface_index"], g["surface"]):
        ray = abi.Ray.from_buffer_copy(rays[i].tobytes())
        hit = abi.Hit.from_buffer_copy(hits[i].tobytes())
        assert bytes(ps.surface_at(ray, hit)) == surf.tobytes(), f"surfaceAt mismatch at ray {i}"


def test_bsdf(oracle):
    g = load_golden("bsdf.npz")
    for s, wo, wi, st, smp, end, ev in zip(g["surface"], g["wo"], g["wi"], g["state"], g["sample"], g["end"],
                                           g["eval"]):
        surf = probe.Surface.from_buffer_copy(s.tobytes())
        b, e = oracle.bsdf_sample(surf, wo, int(st))
        assert bytes(b) == smp.tobytes()
        assert e == int(end)
        val, pdf = oracle.bsdf_eval(surf, wo, wi)
        np.testing.assert_array_equal(np.concatenate([val, [pdf]]).astype(np.float32), ev)


@pytest.mark.parametrize("name", [e[0] for e in make_golden.test_envs()])
def test_environment(oracle, name):
    g = load_golden("env.npz")
    env = dict(make_golden.test_envs())[name]
    pe = oracle.env(env)
    for st, d, p, end in zip(g[f"{name}/state"], g[f"{name}/dir"], g[f"{name}/pdf"], g[f"{name}/end"]):
        dd, pp, ee = pe.sample(int(st))
        np.testing.assert_array_equal(dd, d)
        assert pp == p and ee == int(end)
    np.testing.assert_array_equal(np.array([pe.pdf(q) for q in g[f"{name}/query"]], np.float32), g[f"{name}/qpdf"])
    np.testing.assert_array_equal(np.array([pe.radiance(q) for q in g[f"{name}/query"]]), g[f"{name}/qrad"])


@pytest.mark.parametrize("name", [c[0] for c in make_golden.render_cases()])
def test_render(oracle, name):
    g = load_golden("render.npz")
    _, scene, env, cam, opt, crop = next(c for c in make_golden.render_cases() if c[0] == name)
    assert scene.bytes_sha256() == str(g[f"{name}/scene_sha"]), "scene generator drifted"
    film, stats = oracle.scene(scene).render(oracle.env(env), cam, opt)
    assert make_golden.sha(film) == str(g[f"{name}/film_sha"]), f"{name}: film differs from the reference"
    assert stats.boundary_retries == int(g[f"{name}/retries"])


def test_textured_surfaces(oracle):
    """Textured surfaceAt (all five maps, every wrap mode, nearest and bilinear, UV set 1,
    normal mapping with rotated/mirrored slot transforms, tangents derived in finalize)
    equals the reference bit for bit (reference src/geometry.cpp:238-275, :296-459)."""
    g = load_golden("textured_surface.npz")
    for name, scene, rays in make_golden.textured_surface_sets():
        ps = oracle.scene(scene)
        hits = ps.trace(rays)
        for i, surf in zip(g[f"{name}/index"], g[f"{name}/surface"]):
            ray = abi.Ray.from_buffer_copy(rays[i].tobytes())
            hit = abi.Hit.from_buffer_copy(hits[i].tobytes())
            assert bytes(ps.surface_at(ray, hit)) == surf.tobytes(), f"{name}: surfaceAt mismatch at ray {i}"


def _film_names():
    g = load_golden("film_encode.npz")
    return sorted({k.split("/")[0] for k in g.files})


@pytest.mark.parametrize("name", _film_names())
def test_film_encode(oracle, name):
    """writePng display bytes and the complete writeHdr file (reference
    src/io/image.cpp:101-124, :201-236), bit for bit; rejected inputs with the
    reference's messages."""
    g = load_golden("film_encode.npz")
    film, ex = g[f"{name}/film"], float(g[f"{name}/exposure"])
    if name.startswith("err_"):
        for kind, fn in (("png", lambda: oracle.film_png8(film, ex)), ("hdr", lambda: oracle.film_hdr(film))):
            want = str(g[f"{name}/{kind}"])
            if want:
                with pytest.raises(RuntimeError) as e:
                    fn()
                assert str(e.value) == want
            else:
                fn()
        return
    np.testing.assert_array_equal(oracle.film_png8(film, ex), g[f"{name}/png8"])
    assert oracle.film_hdr(film) == g[f"{name}/hdr"].tobytes()


@pytest.mark.parametrize("name", [c[0] for c in make_golden.batch_cases()])
def test_render_batch_requests(oracle, name):
    """Requests of the batched/served mix generator: the oracle's films equal the
    reference's bit for bit (the device forms are compared with the same films)."""
    g = load_golden("render_batch.npz")
    _, w, opt = next(c for c in make_golden.batch_cases() if c[0] == name)
    assert w.scene.bytes_sha256() == str(g[f"{name}/scene_sha"]), "scene generator drifted"
    film, stats = oracle.scene(w.scene).render(oracle.env(w.env), w.camera, opt)
    np.testing.assert_array_equal(film.view(np.uint32), g[f"{name}/film"].view(np.uint32))
    assert stats.boundary_retries == int(g[f"{name}/retries"])


WINDOW_CASES = ("c2_small", "c3_small", "tex_factors", "fixture4", "plane_1")


@pytest.mark.parametrize("name", WINDOW_CASES)
def test_render_window_matches_reference_crops(oracle, name):
    """The oracle's window render (the pixels of a window of the frame, each with its frame
    index, so its full-frame sample stream) equals the same crop of the reference's own
    full-frame film bit for bit, at several windows including the frame's edges; the
    per-sample radiance it reports sums, in sample order in binary64, to the film."""
    g = load_golden("render.npz")
    _, scene, env, cam, opt, crop = next(c for c in make_golden.render_cases() if c[0] == name)
    ref = g[f"{name}/film"]
    ps, pe = oracle.scene(scene), oracle.env(env)
    W, H = opt.width, opt.height
    for win in ((0, 0, W, H), (W // 3, H // 4, max(1, W // 2), max(1, H // 2)), (W - 1, H - 1, 1, 1),
                (0, H // 2, W, H - H // 2)):
        x0, y0, w, h = win
        film, st, smp = ps.render_window(pe, cam, opt, win, with_samples=True)
        np.testing.assert_array_equal(film.view(np.uint32), ref[y0:y0 + h, x0:x0 + w].view(np.uint32),
                                      err_msg=f"{name} window {win}")
        assert st.camera_samples == w * h * opt.samples
        acc = np.zeros((h, w, 3), np.float64)
        for s in range(opt.samples):
            acc += smp[s]
        np.testing.assert_array_equal((acc / opt.samples).astype(np.float32).view(np.uint32), film.view(np.uint32))
    with pytest.raises(RuntimeError, match="Render window must lie inside the image"):
        ps.render_window(pe, cam, opt, (W - 1, 0, 2, 1))
