"""The device path's logic, compiled for the CPU, against the reference goldens.

tests/csrc/emu.cpp builds path.h / geom.h / shade.h / sf64.h (the exact
sources the gfx950 kernels compile) with a host compiler.  With glibc on both
sides these runs isolate the device algorithms (FP32-filtered traversal with
binary64-exact fallback, integer CDF inversion, emulated binary64 camera,
offsets and accumulation) from device-libm differences:
  * traversal: ids, flags and binary64 t/u/v bit-exact on every golden ray set;
  * 4-wide tree: every primitive in exactly one leaf, boxes nested;
  * environment inversion table: equal to the reference's upper_bound for all
    2^24 draws (exhaustive);
  * camera rays, environment samples/pdf/radiance bit-exact;
  * films: per-pixel agreement with the reference (pow(x,5) is evaluated
    correctly rounded on the device side, which can differ from libm by 1 ulp).
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from conftest import GOLDEN, abi, load_golden, scenes

sys.path.insert(0, str(GOLDEN))
import make_golden  # noqa: E402


@pytest.fixture(scope="module")
def emu():
    import emu as emu_mod
    return emu_mod.Emu()


@pytest.mark.parametrize("name", [s[0] for s in make_golden.trace_sets()])
def test_emu_trace_bit_exact(emu, name):
    g = load_golden("trace.npz")
    _, scene, rays, _ = next(s for s in make_golden.trace_sets() if s[0] == name)
    es = emu.scene(scene)
    assert es.check() == 0, "4-wide BVH structure and its quantized 4- and 8-wide trees"
    assert es.triangles == int(g[f"{name}/triangles"])
    h = es.trace(rays)
    np.testing.assert_array_equal(h["hit"], g[f"{name}/hit"])
    np.testing.assert_array_equal(h["triangle"], g[f"{name}/triangle"])
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(h[k].view(np.uint64), g[f"{name}/{k}"].view(np.uint64))
    np.testing.assert_array_equal(es.trace(rays, True)["hit"], g[f"{name}/any"])
    # Postponed-leaf traversal (plain and preloaded leaf tests): any
    # interleaving of inner and leaf rounds gives the same closest and any hits.
    # Schedules 9 and 13 walk the 8-wide tree (with and without the FP32 mask tests).
    for schedule in (1, 2, 5, 9, 13):
        hp = es.trace(rays, schedule=schedule)
        np.testing.assert_array_equal(hp, h)
        np.testing.assert_array_equal(es.trace(rays, True, schedule=schedule)["hit"], g[f"{name}/any"])


@pytest.mark.parametrize("name", [e[0] for e in make_golden.test_envs()] + ["sky1024"])
def test_env_inversion_exhaustive(emu, name):
    from conftest import scenes
    env = dict(make_golden.test_envs()).get(name) or abi.EnvData(rgba=scenes.synthetic_sky())
    assert emu.env(env).verify_inversion() == 0


@pytest.mark.parametrize("name", [e[0] for e in make_golden.test_envs()])
def test_emu_environment_bit_exact(emu, name):
    g = load_golden("env.npz")
    env = dict(make_golden.test_envs())[name]
    pe = emu.env(env)
    for st, d, p, end in zip(g[f"{name}/state"][:500], g[f"{name}/dir"], g[f"{name}/pdf"], g[f"{name}/end"]):
        dd, pp, ee = pe.sample(int(st))
        np.testing.assert_array_equal(dd, d)
        assert pp == p and ee == int(end)
    q = g[f"{name}/query"][:500]
    np.testing.assert_array_equal(np.array([pe.pdf(v) for v in q], np.float32), g[f"{name}/qpdf"][:500])
    np.testing.assert_array_equal(np.array([pe.radiance(v) for v in q]), g[f"{name}/qrad"][:500])


def _exact_slab(o, d, tmin, tmax, lo, hi):
    """Slab test in exact rational arithmetic: True when the ray meets the box
    for some t in [tmin, tmax]."""
    from fractions import Fraction as F
    tn = F(float(tmin))
    tf = None if np.isinf(tmax) else F(float(tmax))
    for a in range(3):
        oa, da = F(float(o[a])), F(float(d[a]))
        la, ha = F(float(lo[a])), F(float(hi[a]))
        if da == 0:
            if oa < la or oa > ha:
                return False
            continue
        t1, t2 = (la - oa) / da, (ha - oa) / da
        n_, f_ = min(t1, t2), max(t1, t2)
        tn = max(tn, n_)
        tf = f_ if tf is None else min(tf, f_)
    return tf is None or tn <= tf


def test_slab_test_is_conservative(emu):
    """node_hit never culls a box the ray meets in exact arithmetic (the FP32
    traversal may only over-approximate, reference src/geometry.cpp:565-586):
    grazing boxes whose faces pass through points of the ray, at origin scales
    from 1e-3 to 1e5, with tiny and zero direction components.  The box is
    quantized as a child of a wide node, half the time next to a sibling box
    up to 1e4 box sizes away (the node's origin far from the box, a coarse
    scale, the per-node distance terms large against the box's own)."""
    import ctypes as C
    rng = np.random.default_rng(11)
    fp = C.POINTER(C.c_float)
    f32 = lambda v: np.ascontiguousarray(v, np.float32)
    missed = tight_misses = 0
    for case in range(6000):
        scale = [1e-3, 1.0, 1e3, 1e5][case % 4]
        o = f32(rng.uniform(-1, 1, 3) * scale)
        d = rng.normal(size=3)
        if case % 7 == 0:
            d[case % 3] *= 1e-6
        if case % 11 == 0:
            d[(case // 11) % 3] = 0.0
        d = f32(d / np.linalg.norm(d))
        t0 = rng.uniform(0.01, 50.0) * max(scale, 1.0)
        p = o.astype(np.float64) + t0 * d.astype(np.float64)
        ext = np.abs(rng.normal(size=3)) * t0 * 10.0 ** rng.uniform(-7, -1)
        lo, hi = f32(p - ext), f32(p + ext)
        # Push one face exactly onto the ray point (grazing contact), or just past it.
        a = rng.integers(3)
        nudge = rng.integers(-2, 3)
        if rng.random() < 0.5:
            lo[a] = np.nextafter(np.float32(p[a]), np.float32(np.inf) if nudge > 0 else np.float32(-np.inf)) \
                if nudge else np.float32(p[a])
        else:
            hi[a] = np.nextafter(np.float32(p[a]), np.float32(-np.inf) if nudge > 0 else np.float32(np.inf)) \
                if nudge else np.float32(p[a])
        lo, hi = f32(np.minimum(lo, hi)), f32(np.maximum(lo, hi))
        tmax = np.float32(np.inf) if case % 3 else np.float32(t0 * (1 + rng.uniform(-1e-6, 1e-6)))
        sib_lo = sib_hi = None
        if case % 2:
            size = float(np.max(hi.astype(np.float64) - lo)) + 1e-30
            off = rng.normal(size=3) * size * 10.0 ** rng.uniform(0, 4)
            c = (lo.astype(np.float64) + hi) / 2 + off
            sib_lo, sib_hi = f32(c - size * rng.uniform(0.1, 3)), f32(c + size * rng.uniform(0.1, 3))
        got = emu.lib.emu_box_hit(o.ctypes.data_as(fp), d.ctypes.data_as(fp), 0.0, float(tmax),
                                  lo.ctypes.data_as(fp), hi.ctypes.data_as(fp),
                                  sib_lo.ctypes.data_as(fp) if sib_lo is not None else None,
                                  sib_hi.ctypes.data_as(fp) if sib_hi is not None else None)
        exact = _exact_slab(o, d, 0.0, float(tmax), lo, hi)
        if exact and not got:
            missed += 1
        if not exact:
            tight_misses += 1
    assert missed == 0
    assert tight_misses > 200  # the set really probes both sides of contact


def test_slab_test_keeps_corner_contacts(emu):
    """Boxes the ray meets at (or within rounding of) one corner: it enters
    through two faces and leaves through the third at nearly the same t, so
    every box it meets at all it meets only just.  Every plane is exactly
    decodable (coordinates on a power-of-two grid whose step is the node's
    scale), so the decoded box is the box itself and nothing but the slab
    arithmetic's widening keeps the contact.  The ray origin lies deep inside
    a node spanning 255 grid steps at coordinates up to 2^12 steps, so the
    contact distance is small against the node's planes and coordinates (a
    slab arithmetic whose rounding grew with those, not with the distance,
    would lose such contacts)."""
    import ctypes as C
    rng = np.random.default_rng(5)
    fp = C.POINTER(C.c_float)
    f32 = lambda v: np.ascontiguousarray(v, np.float32)
    hits = 0
    for case in range(6000):
        g = 2.0 ** int(rng.integers(-12, 8))             # grid step = the node's scale
        p = rng.integers(-2 ** 12, 2 ** 12, 3) * g      # node origin (low corner)
        c = p + rng.integers(64, 192, 3) * g             # the corner
        sgn = np.where(rng.random(3) < 0.5, -1.0, 1.0)
        d = f32(sgn * rng.uniform(1.0 / 64, 1.0, 3))
        t0 = g * rng.uniform(1e-3, 4.0)
        o = f32(c - t0 * d.astype(np.float64))           # rounded: the contact is only nearly exact
        k = rng.integers(1, 33, 3) * g
        lo, hi = c.copy(), c.copy()
        leave = int(rng.integers(3))                      # the axis whose slab ends at the corner
        for a in range(3):
            if (d[a] > 0) != (a == leave):                # entering faces sit at the corner, the box beyond
                hi[a] = c[a] + k[a]
            else:
                lo[a] = c[a] - k[a]
        sib_lo, sib_hi = f32(p), f32(p + 255 * g)
        lo, hi = f32(lo), f32(hi)
        if not _exact_slab(o, d, 0.0, float("inf"), lo, hi):
            continue
        got = emu.lib.emu_box_hit(o.ctypes.data_as(fp), d.ctypes.data_as(fp), 0.0, float("inf"),
                                  lo.ctypes.data_as(fp), hi.ctypes.data_as(fp),
                                  sib_lo.ctypes.data_as(fp), sib_hi.ctypes.data_as(fp))
        assert got == 1, (o, d, lo, hi, p, g)
        hits += 1
    assert hits > 1500


def test_slab_test_keeps_contacts_at_rounded_planes(emu):
    """Corner contacts where the ray leaves through the node's own far plane
    on an axis whose scale is half the ulp of the node origin p: the builder's
    255th step rounds p + 127.5 ulp up to the box's face p + 128 ulp, so the
    decoded plane is the face while the unrounded sum lies half an ulp inside
    it (a slab arithmetic that skipped the decode's rounding would lose such
    contacts).  Every box the ray meets in exact arithmetic must be kept (the
    other two axes enter at grid-exact planes)."""
    import ctypes as C
    rng = np.random.default_rng(6)
    fp = C.POINTER(C.c_float)
    f32 = lambda v: np.ascontiguousarray(v, np.float32)
    hits = 0
    for case in range(8000):
        ax = int(rng.integers(3))                         # the axis left through the rounded far plane
        e = int(rng.integers(-4, 16))
        m = 2 * int(rng.integers(2 ** 21, 2 ** 22))       # even significand: p + 127.5 ulp rounds up
        sx = 1.0 if rng.random() < 0.5 else -1.0
        px = sx * m * 2.0 ** (e - 23)
        ulp = float(np.spacing(np.float32(abs(px))))
        if float(np.float32(px + 127.5 * ulp)) != px + 128 * ulp:
            continue
        g = 2.0 ** int(rng.integers(-12, 4))
        pg = rng.integers(-2 ** 10, 2 ** 10, 3) * g       # grid-exact node origins on the other axes
        corner = pg + rng.integers(64, 192, 3) * g
        node_lo, node_hi = pg.copy(), pg + 255 * g
        node_lo[ax], node_hi[ax] = px, px + 128 * ulp
        corner[ax] = px + 128 * ulp
        d = rng.uniform(1.0 / 64, 1.0, 3) * np.where(rng.random(3) < 0.5, -1.0, 1.0)
        d[ax] = abs(d[ax])                                # leaving through the high face
        d = f32(d)
        lo, hi = corner.copy(), corner.copy()
        for a in range(3):
            if a == ax:
                lo[a] = corner[a] - int(rng.integers(1, 64)) * ulp
            elif d[a] > 0:
                hi[a] = corner[a] + int(rng.integers(1, 33)) * g
            else:
                lo[a] = corner[a] - int(rng.integers(1, 33)) * g
        t0 = min(ulp, g) * rng.uniform(0.5, 40.0)
        o = f32(corner - t0 * d.astype(np.float64))
        lo, hi = f32(lo), f32(hi)
        if not _exact_slab(o, d, 0.0, float("inf"), lo, hi):
            continue
        sib_lo, sib_hi = f32(node_lo), f32(node_hi)
        got = emu.lib.emu_box_hit(o.ctypes.data_as(fp), d.ctypes.data_as(fp), 0.0, float("inf"),
                                  lo.ctypes.data_as(fp), hi.ctypes.data_as(fp),
                                  sib_lo.ctypes.data_as(fp), sib_hi.ctypes.data_as(fp))
        assert got == 1, (o, d, lo, hi, node_lo, node_hi)
        hits += 1
    assert hits > 1000


def test_emu_camera_bit_exact(emu):
    import ctypes as C
    g = load_golden("rng_camera.npz")
    cams = [make_golden.scenes.config1().camera, make_golden.scenes.config3(triangles=10).camera,
            make_golden.scenes.look_at((3, 4, 5), (0, 0.5, 0), 0.8), make_golden.rs.test_camera(),
            make_golden.rs.touching_boundary()[2], abi.camera(yfov=np.pi / 2, near_clip=1, far_clip=10)]
    for ci, px, py, w, h, ok, ray in zip(g["cam_index"], g["px"], g["py"], g["w"], g["h"], g["ok"], g["rays"]):
        r = abi.Ray()
        rc = emu.lib.emu_camera_generate(C.byref(cams[ci]), float(px), float(py), int(w), int(h), C.byref(r))
        assert (rc == 0) == bool(ok)
        if rc == 0:
            assert bytes(r) == ray.tobytes()


def within(a, b, rel=1e-4):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    return np.all((d <= rel * np.abs(b.astype(np.float64))) | (d <= 1e-7), axis=-1)


@pytest.mark.parametrize("name", [c[0] for c in make_golden.render_cases()])
def test_emu_render(emu, name):
    g = load_golden("render.npz")
    _, scene, env, cam, opt, crop = next(c for c in make_golden.render_cases() if c[0] == name)
    film, stats = emu.scene(scene).render(emu.env(env), cam, opt)
    if crop is not None:
        y0, y1, x0, x1 = crop
        film = film[y0:y1, x0:x1]
    ref = g[f"{name}/film"]
    ok = within(film, ref)
    exact = np.mean(np.all(film == ref, axis=-1))
    print(f"{name}: within 1e-4 {ok.mean() * 100:.3f}%  bit-exact {exact * 100:.3f}%")
    if ok.size >= 1000:
        assert ok.mean() >= 0.999
    else:
        assert ok.all()
    assert stats.boundary_retries == int(g[f"{name}/retries"])



def _film_names():
    g = load_golden("film_encode.npz")
    return sorted({k.split("/")[0] for k in g.files})


@pytest.mark.parametrize("name", _film_names())
def test_emu_film_encode(emu, name):
    """The device's film encoders (film.h) with the host thresholds
    (film_host.cpp), run on the CPU: writePng display bytes and the whole
    writeHdr file bit for bit against the reference's writers, and the same
    rejections (reference src/io/image.cpp:101-124, :201-236)."""
    g = load_golden("film_encode.npz")
    film, ex = g[f"{name}/film"], float(g[f"{name}/exposure"])
    if name.startswith("err_"):
        for kind, fn in (("png", lambda: emu.film_png8(film, ex)), ("hdr", lambda: emu.film_hdr(film))):
            want = str(g[f"{name}/{kind}"])
            if want:
                with pytest.raises(RuntimeError) as e:
                    fn()
                assert str(e.value) == want
            else:
                fn()
        return
    np.testing.assert_array_equal(emu.film_png8(film, ex), g[f"{name}/png8"])
    assert emu.film_hdr(film) == g[f"{name}/hdr"].tobytes()


def test_display_thresholds_match_the_curve(emu, oracle):
    """Threshold search equals the display curve evaluated per value, on values
    at and one float step either side of every threshold and on a dense sweep,
    at several exposures (the curve's monotonicity is what makes the table
    exact)."""
    g = np.random.default_rng(3)
    for ex in (-7.0, -1.25, 0.0, 0.5, 3.0, 9.0):
        v = np.concatenate([np.geomspace(1e-9, 1e4, 20000), g.random(5000) * 3]).astype(np.float32)
        film = np.resize(v, (len(v) // 300 + 1) * 300).reshape(-1, 100, 3)
        np.testing.assert_array_equal(emu.film_png8(film, ex), oracle.film_png8(film, ex))
        # values straddling each byte boundary of the reference curve
        ref = oracle.film_png8(film, ex).ravel()
        edges = np.nonzero(np.diff(ref[np.argsort(film.ravel(), kind="stable")]))[0]
        s = np.sort(film.ravel())
        probe_v = np.concatenate([s[edges], s[edges + 1]])
        near = np.concatenate([probe_v, np.nextafter(probe_v, np.float32(np.inf)),
                               np.nextafter(probe_v, np.float32(0))]).astype(np.float32)
        near = np.resize(near, (len(near) // 300 + 1) * 300).reshape(-1, 100, 3)
        np.testing.assert_array_equal(emu.film_png8(near, ex), oracle.film_png8(near, ex))


def test_expf_is_correctly_rounded(emu):
    """The device's expf (Beer-Lambert transmittance) equals e^x correctly rounded to binary32
    on a sweep of every 1009th binary32 argument over [-104, 89] (the underflow-to-zero and
    overflow thresholds bracket it), on the subnormal-result range densely, and on special values."""
    fp = C.POINTER(C.c_float)
    neg = np.arange(0x80000000, 0xC2D00001, 1009, dtype=np.uint64).astype(np.uint32).view(np.float32)
    pos = np.arange(0, 0x42B20001, 1009, dtype=np.uint64).astype(np.uint32).view(np.float32)
    sub = np.linspace(-103.98, -87.3, 200001, dtype=np.float32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, -104.0, -103.97, 88.72, 88.73, 89.0, 1e-30, -1e-30],
                       np.float32)
    x = np.concatenate([neg, pos, sub, special])
    out = np.empty_like(x)
    emu.lib.emu_expf(x.ctypes.data_as(fp), out.ctypes.data_as(fp), len(x))
    with np.errstate(over="ignore", under="ignore"):
        want = np.exp(x.astype(np.float64)).astype(np.float32)
    bad = np.nonzero(out.view(np.uint32) != want.view(np.uint32))[0]
    assert len(bad) == 0, [(float(x[i]), float(out[i]), float(want[i])) for i in bad[:5]]
    nan = np.array([np.nan], np.float32)
    emu.lib.emu_expf(nan.ctypes.data_as(fp), out.ctypes.data_as(fp), 1)
    assert np.isnan(out[0])


def _camera_rays(emu, cam, width, height):
    rays = np.zeros(width * height, abi.RAY_DTYPE)
    r = abi.Ray()
    for y in range(height):
        for x in range(width):
            emu.lib.emu_camera_generate(C.byref(cam), x + 0.37, y + 0.61, width, height, C.byref(r))
            rays[y * width + x] = np.frombuffer(bytes(r), abi.RAY_DTYPE)[0]
    return rays


@pytest.mark.parametrize("case", ["stress", "tex_mask_random", "tex_mask_camera", "gltf_contract"])
def test_fp32_mask_decision_is_conservative(emu, case):
    """The traversal decides alpha masks in FP32 where the footprint's texels allow it
    (mask_fp32); the decision must never contradict the binary64 MASK test at the exact
    barycentrics (reference src/geometry.cpp:279-292), and must decide most hits of
    ordinary rays (random rays include many grazing ones, whose barycentrics are loose)."""
    import ref_scenes as rs
    if case == "stress":
        scene = rs.alpha_mask_stress()
        rays, need = scenes.random_rays(scene, 60000, seed=21, spread=1.2), 0.7
    elif case.startswith("tex_mask"):
        scene, _, cam = rs.textured_alpha_mask()
        if case.endswith("camera"):
            rays, need = _camera_rays(emu, cam, 200, 160), 0.95
        else:
            rays, need = scenes.random_rays(scene, 60000, seed=21, spread=1.2), 0.3
    else:
        scene = rs.gltf_export_contract()[0]
        rays, need = scenes.random_rays(scene, 20000, seed=21, spread=1.2), 0.8
    sc = emu.scene(scene)
    counts = (C.c_uint64 * 5)()
    rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
    assert emu.lib.emu_mask_check(sc.h, rays.ctypes.data, len(rays), counts) == 0
    acc, rej, unsure, wrong, missed = counts[:]
    print(f"{case}: accept {acc} reject {rej} undecided {unsure} contradictions {wrong} missed {missed}")
    assert wrong == 0 and missed == 0
    decided = acc + rej
    assert decided > 0 and decided >= need * (decided + unsure)


def test_split_references_cover_their_triangles(emu, oracle):
    """The host build splits triangles whose box is far larger than the mean into several
    references with clipped boxes.  Rays aimed at points spread over those triangles (and
    grazing their edges) must hit exactly what the oracle's unsplit BVH reports: the union of
    the reference boxes covers each triangle, and repeated references change no result."""
    rng = np.random.Generator(np.random.PCG64(404))
    b = None
    import ref_scenes as rs
    b = rs.Builder()
    b.materials = [abi.material(base_color=(0.7, 0.7, 0.7), metallic=0.0, roughness=0.9)]
    # two huge slanted triangles over a field of small ones
    b.tri((-50, -1, -50), (50, -3, -50), (0, 4, 60), 0)
    b.tri((-40, 8, -40), (45, 2, 30), (-45, 5, 50), 0)
    for _ in range(3000):
        c = rng.uniform(-40, 40, 3) * np.array([1, 0.05, 1])
        e = rng.normal(0, 0.6, (3, 3))
        b.tri(*(c + e[k] for k in range(3)), 0)
    scene = b.scene()
    p = scene.positions.reshape(-1, 3, 3).astype(np.float64)
    rays = np.zeros(8000, abi.RAY_DTYPE)
    for i in range(len(rays)):
        t = p[i % 2]
        w = rng.dirichlet((1, 1, 1)) if i % 4 else np.array([0.5, 0.5 - 1e-7, 1e-7])  # interior / near an edge
        target = w @ t
        o = target + rng.normal(0, 1, 3) * 30 + np.array([0, 40, 0])
        d = target - o
        rays[i]["origin"] = o
        rays[i]["direction"] = d / np.linalg.norm(d)
        rays[i]["tmin"] = 0.0
        rays[i]["tmax"] = np.inf
        rays[i]["ignore"] = 0xFFFFFFFF
    es = emu.scene(scene)
    assert es.check() == 0
    got = es.trace(rays)
    want = oracle.scene(scene).trace(rays)
    for k in ("hit", "triangle"):
        np.testing.assert_array_equal(got[k], want[k])
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(got[k].view(np.uint64), want[k].view(np.uint64))
    assert np.isin([0, 1], got["triangle"]).all()
    np.testing.assert_array_equal(es.trace(rays, True)["hit"], oracle.scene(scene).trace(rays, True)["hit"])


def test_pow5_float_float_equals_the_binary64_chain(emu):
    """The Schlick term's x^5 in float-float (with the binary64 chain for hard cases) equals the
    correctly rounded binary64 chain on every 37th binary32 in [2^-30, 1] (the term's domain),
    on negative arguments and on special values."""
    fp = C.POINTER(C.c_float)
    u = np.arange(np.float32(2.0 ** -30).view(np.uint32), np.float32(1.0).view(np.uint32) + 1, 37, dtype=np.uint64)
    x = u.astype(np.uint32).view(np.float32)
    x = np.concatenate([x, -x[::97], np.array([0.0, -0.0, 1.0, 2.0 ** -12, 2.0 ** -24, 2.0 ** 24, 3.0e10], np.float32)])
    ff, chain = np.empty_like(x), np.empty_like(x)
    emu.lib.emu_pow5(x.ctypes.data_as(fp), ff.ctypes.data_as(fp), chain.ctypes.data_as(fp), len(x))
    bad = np.nonzero(ff.view(np.uint32) != chain.view(np.uint32))[0]
    assert len(bad) == 0, [(float(x[i]), float(ff[i]), float(chain[i])) for i in bad[:5]]
    # spot check against exact rational arithmetic
    from fractions import Fraction
    for v in x[:: max(1, len(x) // 200)]:
        want = np.float32(float(Fraction(float(v)) ** 5))  # float() of the exact value, then to binary32
        assert abs(float(chain[x == v][0]) - float(want)) <= abs(float(want)) * 2 ** -23


def test_power_heuristic_float_float_equals_the_binary64_chain(emu):
    """The MIS weight a^2 / (a^2 + b^2) in float-float (with the binary64 chain for hard cases)
    equals the reference's binary64 expression rounded to float (reference src/math.h:124-128)
    on 3 M pairs: log-uniform pdfs over 2^-50..2^50, equal and nearly equal pairs, b = 0,
    negative and non-positive a."""
    fp = C.POINTER(C.c_float)
    g = np.random.Generator(np.random.PCG64(77))
    n = 1_000_000
    a = np.exp2(g.uniform(-50, 50, n)).astype(np.float32)
    b = np.exp2(g.uniform(-50, 50, n)).astype(np.float32)
    a2 = np.exp2(g.uniform(-20, 20, n)).astype(np.float32)
    b2 = (a2 * g.uniform(0.5, 2.0, n).astype(np.float32)).astype(np.float32)
    a3 = g.uniform(0, 4, n).astype(np.float32)
    b3 = np.nextafter(a3, np.float32(np.inf)) * (g.integers(0, 2, n) * 2 - 1).astype(np.float32)
    special = np.array([[1, 0], [0, 1], [-1, 2], [3, -3], [2.0 ** -45, 1], [1, 2.0 ** 45], [5, 5]], np.float32)
    A = np.concatenate([a, a2, a3, special[:, 0]])
    B = np.concatenate([b, b2, b3, special[:, 1]])
    ff, chain = np.empty_like(A), np.empty_like(A)
    emu.lib.emu_mis(A.ctypes.data_as(fp), B.ctypes.data_as(fp), ff.ctypes.data_as(fp), chain.ctypes.data_as(fp),
                    len(A))
    bad = np.nonzero(ff.view(np.uint32) != chain.view(np.uint32))[0]
    assert len(bad) == 0, [(float(A[i]), float(B[i]), float(ff[i]), float(chain[i])) for i in bad[:5]]
    want = (A.astype(np.float64) ** 2 / (A.astype(np.float64) ** 2 + B.astype(np.float64) ** 2)).astype(np.float32)
    ok = A > 0
    np.testing.assert_array_equal(chain[ok], want[ok])


def test_dunit_float_float_equals_the_binary64_emulation(emu):
    """Unit vectors (surface normals, ray directions) in float-float with the binary64
    emulation for hard cases equal the emulated binary64 normalization rounded to float on
    1.2 M vectors: random binary64 vectors over 2^-70..2^70, widened floats, near-axis vectors
    with tiny or zero (and negative zero) components, and the zero vector (fallback)."""
    fp = C.POINTER(C.c_float)
    g = np.random.Generator(np.random.PCG64(91))
    n = 300_000
    v1 = g.normal(size=(n, 3)) * np.exp2(g.uniform(-70, 70, (n, 1)))
    v2 = g.normal(size=(n, 3)).astype(np.float32).astype(np.float64)
    v3 = np.zeros((n, 3)); v3[:, 0] = g.uniform(0.1, 10, n); v3[:, 1] = g.normal(size=n) * 1e-9
    v3[:, 2] = np.where(g.integers(0, 2, n) == 1, -0.0, 0.0)
    v4 = g.normal(size=(n, 3)) * np.exp2(g.uniform(-3, 3, (n, 3)))
    v = np.concatenate([v1, v2, v3, v4, np.zeros((1, 3)), [[-0.0, 0.0, 1e-300]]])
    bits = np.ascontiguousarray(v.astype(np.float64)).view(np.uint64)
    ff = np.empty((len(v), 3), np.float32)
    chain = np.empty_like(ff)
    emu.lib.emu_dunit(bits.ctypes.data_as(C.POINTER(C.c_uint64)), ff.ctypes.data_as(fp), chain.ctypes.data_as(fp),
                      len(v))
    bad = np.nonzero(np.any(ff.view(np.uint32) != chain.view(np.uint32), axis=1))[0]
    assert len(bad) == 0, [(v[i].tolist(), ff[i].tolist(), chain[i].tolist()) for i in bad[:3]]


def test_offset_coordinate_float_float_equals_the_binary64_sum(emu):
    """offsetRay's shifted coordinate float(p + 2^-21 * scale * n) in float-float with the
    binary64 emulation for hard cases equals both the emulated binary64 sum and numpy's
    binary64 evaluation of the reference expression (src/geometry.cpp:193-198) on 1.8 M
    cases: random points and normals, small coordinates of far-out points, offsets placed on and a few ulps around binary32
    rounding midpoints (where rounding twice through binary64 matters), tiny and zero
    (signed) normal components, and a zero scale."""
    fp = C.POINTER(C.c_float)
    g = np.random.Generator(np.random.PCG64(123))
    m = 300_000
    f32 = np.float32
    # random coordinates; scale at least |p|
    p1 = (g.normal(size=m) * np.exp2(g.uniform(-20, 20, m))).astype(f32)
    s1 = (np.abs(p1) * np.exp2(g.uniform(0, 6, m))).astype(f32)
    n1 = g.uniform(-1, 1, m).astype(f32)
    # offsets near (k + 1/2) ulp(p): n chosen so 2^-21 * scale * n sits on a midpoint, then nudged
    p2 = (g.uniform(1, 2, m) * np.exp2(g.integers(-10, 10, m))).astype(f32)
    s2 = np.abs(p2).astype(f32)
    ulp = (np.nextafter(np.abs(p2), f32(np.inf)) - np.abs(p2)).astype(np.float64)
    k = g.integers(0, 8, m) + 0.5
    n2 = (k * ulp / (2.0 ** -21 * s2.astype(np.float64))).astype(f32)
    nudge = g.integers(-3, 4, m)
    for step in (1, -1):
        sel = np.sign(nudge) == step
        for _ in range(3):
            n2[sel & (np.abs(nudge) > _)] = np.nextafter(n2[sel & (np.abs(nudge) > _)], f32(step * np.inf))
    n2 *= np.where(g.integers(0, 2, m) == 1, f32(-1), f32(1))
    n2 = np.clip(n2, f32(-1), f32(1))
    # exact products at the midpoint, signs mixed
    p3 = -p2
    # tiny and zero components
    p4 = (g.normal(size=m) * 100).astype(f32)
    s4 = (np.abs(p4) + f32(1)).astype(f32)
    n4 = np.where(g.integers(0, 3, m) == 0, f32(0.0), (g.normal(size=m) * 1e-12).astype(f32))
    n4 = np.where(g.integers(0, 4, m) == 0, f32(-0.0), n4).astype(f32)
    # small coordinates of points far out on another axis (scale up to 2^24 |p|), with offsets
    # near the rounding midpoints of p or (random n) many times p
    p5 = (g.uniform(1, 2, m) * np.exp2(g.integers(-12, 4, m))).astype(f32)
    s5 = (np.abs(p5) * np.exp2(g.uniform(4, 24, m))).astype(f32)
    ulp5 = (np.nextafter(p5, f32(np.inf)) - p5).astype(np.float64)
    n5 = ((g.integers(0, 4, m) + 0.5) * ulp5 / (2.0 ** -21 * s5.astype(np.float64))).astype(f32)
    n5 = (n5 * np.where(g.integers(0, 2, m) == 1, f32(-1), f32(1))).astype(f32)
    n5 = np.where(g.integers(0, 2, m) == 1, n5, g.uniform(-1, 1, m).astype(f32))
    special_p = np.array([0.0, -0.0, -0.0, 0.0, 1.0, -1.0, 3e-39, 1e30], f32)
    special_s = np.array([0.0, 0.0, 0.0, 5.0, 1.0, 1.0, 3e-39, 1e30], f32)
    special_n = np.array([0.0, 0.0, -0.0, 1.0, -0.0, 0.5, 1.0, -1.0], f32)
    P = np.concatenate([p1, p2, p3, p4, p5, special_p])
    S = np.concatenate([s1, s2, s2, s4, s5, special_s])
    N = np.concatenate([n1, n2, n2, n4, n5, special_n])
    ff, chain = np.empty_like(P), np.empty_like(P)
    emu.lib.emu_offset(P.ctypes.data_as(fp), S.ctypes.data_as(fp), N.ctypes.data_as(fp), ff.ctypes.data_as(fp),
                       chain.ctypes.data_as(fp), len(P))
    bad = np.nonzero(ff.view(np.uint32) != chain.view(np.uint32))[0]
    assert len(bad) == 0, [(float(P[i]), float(S[i]), float(N[i]), float(ff[i]), float(chain[i])) for i in bad[:5]]
    want = (P.astype(np.float64) + (2.0 ** -21 * S.astype(np.float64)) * N.astype(np.float64)).astype(f32)
    bad = np.nonzero(chain.view(np.uint32) != want.view(np.uint32))[0]
    assert len(bad) == 0, [(float(P[i]), float(S[i]), float(N[i]), float(chain[i]), float(want[i])) for i in bad[:5]]
    # the midpoint cases do reach the rounding question
    assert np.count_nonzero(p2 + (2.0 ** -21 * s2) * n2 != p2) > m // 2


def test_interpolated_normal_float_float_equals_the_binary64_chain(emu):
    """surfaceAt's shading normal unit(n0 w + n1 hu + n2 hv) (reference src/geometry.cpp:407-415,
    binary64 sums) in float-float with the binary64 chain for hard cases equals the chain alone,
    and the chain equals numpy's binary64 evaluation, on 1.2 M cases: smooth vertex normals,
    flat triangles (n0 = n1 = n2), axis-aligned normals with signed zero components, nearly
    opposite vertex normals, exactly opposite ones with w ~ hu (cancellation by up to 2^30, where
    the binary64 roundings show), and barycentrics at the edges and corners."""
    fp = C.POINTER(C.c_float)
    g = np.random.Generator(np.random.PCG64(131))
    m = 200_000

    def unitv(k):
        v = g.normal(size=(k, 3))
        return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)

    base = unitv(m)
    smooth = np.stack([base, (base + 0.2 * unitv(m)), (base + 0.2 * unitv(m))], 1).astype(np.float32)
    flat = np.repeat(unitv(m)[:, None, :], 3, axis=1)
    axis = np.zeros((m, 3, 3), np.float32)
    ax = g.integers(0, 3, m)
    axis[np.arange(m), :, ax] = np.where(g.integers(0, 2, m) == 1, 1.0, -1.0)[:, None]
    axis[(axis == 0) & (g.integers(0, 2, (m, 3, 3)) == 1)] = -0.0
    opp = np.stack([base, -base + 1e-3 * unitv(m), base], 1).astype(np.float32)
    tilt = np.stack([unitv(m), unitv(m), unitv(m)], 1)
    cancel = np.stack([base, -base, unitv(m)], 1).astype(np.float32)
    normals = np.concatenate([smooth, flat, axis, opp, tilt, cancel]).reshape(-1, 9)
    k = len(normals)
    hu = g.uniform(0, 1, k)
    hv = g.uniform(0, 1, k) * (1 - hu)
    corner = g.integers(0, 6, k)
    hu = np.where(corner == 0, 0.0, np.where(corner == 1, 1.0, hu))
    hv = np.where(corner == 0, np.where(g.integers(0, 2, k) == 1, 1.0, 0.0), np.where(corner == 1, 0.0, hv))
    hv = np.where(corner == 2, 1e-12 * g.uniform(size=k), hv)
    near = slice(5 * m, 6 * m)
    hv[near] = np.exp2(g.uniform(-40, -10, m))
    hu[near] = (1 - hv[near]) / 2 + g.normal(size=m) * np.exp2(g.uniform(-45, -20, m))
    uv = np.ascontiguousarray(np.stack([hu, hv], 1)).view(np.uint64)
    n32 = np.ascontiguousarray(normals, np.float32)
    ff = np.empty((k, 3), np.float32)
    chain = np.empty_like(ff)
    emu.lib.emu_interp_unit(n32.ctypes.data_as(fp), uv.ctypes.data_as(C.POINTER(C.c_uint64)),
                            ff.ctypes.data_as(fp), chain.ctypes.data_as(fp), k)
    bad = np.nonzero(np.any(ff.view(np.uint32) != chain.view(np.uint32), axis=1))[0]
    assert len(bad) == 0, [(n32[i].tolist(), hu[i], hv[i], ff[i].tolist(), chain[i].tolist()) for i in bad[:3]]
    nd = n32.astype(np.float64).reshape(-1, 3, 3)
    w = (1.0 - hu) - hv
    N = (nd[:, 0] * w[:, None] + nd[:, 1] * hu[:, None]) + nd[:, 2] * hv[:, None]
    sq = (N[:, 0] * N[:, 0] + N[:, 1] * N[:, 1]) + N[:, 2] * N[:, 2]
    ok = sq > 0
    inv = 1.0 / np.sqrt(sq[ok])
    want = (N[ok] * inv[:, None]).astype(np.float32)
    np.testing.assert_array_equal(chain[ok].view(np.uint32), want.view(np.uint32))


def test_interpolated_point_float_float_equals_the_binary64_chain(emu):
    """surfaceAt's hit point float((a w + b hu) + c hv) (reference src/geometry.cpp:401-412,
    binary64 sums, then narrowed) in float-float, kept only where that decides the rounding,
    equals the binary64 chain on 1.2 M cases, and the chain equals numpy's binary64 evaluation:
    small and large triangles at coordinates from 1e-3 to 1e6, triangles straddling a zero
    coordinate (cancellation), slivers, axis-aligned triangles with signed zero coordinates,
    barycentrics at the edges and corners, and points put on a binary32 rounding midpoint (the
    float-float sum cannot round those; the chain must).  Most of the other cases must take
    the float-float path."""
    fp = C.POINTER(C.c_float)
    g = np.random.Generator(np.random.PCG64(137))
    m = 200_000

    def tris(scale, size):
        c = g.normal(size=(m, 1, 3)) * scale
        return (c + g.normal(size=(m, 3, 3)) * size).astype(np.float32)

    small = tris(10.0, 0.01)
    large = tris(1e6, 1e3)
    tiny = tris(1e-3, 1e-5)
    straddle = (g.normal(size=(m, 3, 3)) * 0.5).astype(np.float32)
    sliver = tris(100.0, 1.0)
    sliver[:, 2] = (sliver[:, 0].astype(np.float64) + (sliver[:, 1].astype(np.float64) - sliver[:, 0]) *
                    g.uniform(0, 1, (m, 1)) + g.normal(size=(m, 3)) * 1e-6).astype(np.float32)
    flat = tris(5.0, 1.0)
    ax = g.integers(0, 3, m)
    flat[np.arange(m), :, ax] = np.where(g.integers(0, 2, m) == 1, 0.0, -0.0)[:, None]
    # Points on a binary32 rounding midpoint: a + hu (b - a) with a, b on the same binade,
    # hu the binary64 quotient putting the point on a midpoint between them, so only the
    # binary64 chain's roundings decide which float the point narrows to.
    a = np.abs(g.normal(size=(m, 3)) * 10.0 ** g.uniform(-2, 5, (m, 1))).astype(np.float32)
    ulp = np.spacing(a).astype(np.float64)
    steps = g.integers(2, 1000, (m, 1)).astype(np.float64)
    b = (a.astype(np.float64) + steps * ulp).astype(np.float32)
    mid = a.astype(np.float64) + (np.floor(g.uniform(0, 1, (m, 1)) * (steps - 1)) + 0.5) * ulp
    axis_mid = g.integers(0, 3, m)
    hu_mid = (mid[np.arange(m), axis_mid] - a[np.arange(m), axis_mid]) / \
        (b[np.arange(m), axis_mid].astype(np.float64) - a[np.arange(m), axis_mid])
    midpoint = np.stack([a, b, (a.astype(np.float64) * 1.5).astype(np.float32)], 1)
    verts = np.concatenate([small, large, tiny, straddle, sliver, flat, midpoint]).reshape(-1, 9)
    k = len(verts)
    hu = g.uniform(0, 1, k)
    hv = g.uniform(0, 1, k) * (1 - hu)
    corner = g.integers(0, 6, k)
    hu = np.where(corner == 0, 0.0, np.where(corner == 1, 1.0, hu))
    hv = np.where(corner == 0, np.where(g.integers(0, 2, k) == 1, 1.0, 0.0), np.where(corner == 1, 0.0, hv))
    hv = np.where(corner == 2, 1e-12 * g.uniform(size=k), hv)
    hu[6 * m:] = hu_mid
    hv[6 * m:] = 0.0
    uv = np.ascontiguousarray(np.stack([hu, hv], 1)).view(np.uint64)
    v32 = np.ascontiguousarray(verts, np.float32)
    ff = np.empty((k, 3), np.float32)
    chain = np.empty_like(ff)
    decided = np.zeros(k, np.uint8)
    emu.lib.emu_interp_point(v32.ctypes.data_as(fp), uv.ctypes.data_as(C.POINTER(C.c_uint64)),
                             ff.ctypes.data_as(fp), chain.ctypes.data_as(fp),
                             decided.ctypes.data_as(C.POINTER(C.c_uint8)), k)
    bad = np.nonzero(np.any(ff.view(np.uint32) != chain.view(np.uint32), axis=1))[0]
    assert len(bad) == 0, [(v32[i].tolist(), hu[i], hv[i], ff[i].tolist(), chain[i].tolist()) for i in bad[:3]]
    assert decided[:6 * m].mean() > 0.99, decided[:6 * m].mean()
    vd = v32.astype(np.float64).reshape(-1, 3, 3)
    w = (1.0 - hu) - hv
    P = (vd[:, 0] * w[:, None] + vd[:, 1] * hu[:, None]) + vd[:, 2] * hv[:, None]
    np.testing.assert_array_equal(chain.view(np.uint32), P.astype(np.float32).view(np.uint32))


def _edge_rays(scene_data, count, seed):
    """Rays at triangle edges and vertices (barycentrics within 1e-7 of 0), from near and
    far (up to 1e4 scene sizes), some grazing the triangle's plane."""
    g = np.random.Generator(np.random.PCG64(seed))
    p = scene_data.positions.reshape(-1, 3, 3).astype(np.float64)
    idx = g.integers(0, len(p), count)
    a, b, c = p[idx, 0], p[idx, 1], p[idx, 2]
    t = g.uniform(0, 1, (count, 1))
    on_edge = a + (b - a) * t + (c - a) * (g.uniform(-1e-7, 1e-7, (count, 1)))
    corner = np.where(g.integers(0, 2, (count, 1)) == 1, a, b)
    target = np.where(g.integers(0, 3, (count, 1)) == 0, corner, on_edge)
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True) + 1e-300
    d = g.normal(size=(count, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    graze = g.integers(0, 2, count) == 1
    d[graze] -= n[graze] * np.sum(d[graze] * n[graze], 1, keepdims=True) * (1 - 1e-4)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    size = np.ptp(p.reshape(-1, 3), 0).max()
    dist = size * 10.0 ** g.uniform(-3, 4, (count, 1))
    rays = abi.make_rays((target - d * dist).astype(np.float32), d.astype(np.float32))
    return rays


def _midpoint_t_case(count, seed):
    """Triangles in planes z = f (floats in [1, 2)) and rays from z = -2^-24 along (dx, dy, 1)
    with small dx, dy: the exact hit distance f + 2^-24 lies on a binary32 rounding midpoint,
    so only the binary64 chain's roundings (its shear terms are not exact) decide float(t).
    Each ray meets its own triangle (grid cells 8 apart, triangles 2 across)."""
    g = np.random.Generator(np.random.PCG64(seed))
    side = int(np.ceil(np.sqrt(count)))
    ij = np.stack(np.meshgrid(np.arange(side), np.arange(side)), -1).reshape(-1, 2)[:count]
    f = (1.0 + g.integers(0, 2 ** 23, count) * 2.0 ** -23).astype(np.float32)
    o = np.zeros((count, 3), np.float32)
    o[:, :2] = (ij * 8.0 + g.uniform(-0.5, 0.5, (count, 2))).astype(np.float32)
    o[:, 2] = -2.0 ** -24
    d = np.ones((count, 3), np.float32)
    d[:, :2] = g.uniform(-0.1, 0.1, (count, 2)).astype(np.float32)
    cx = o[:, 0] + (f + 2.0 ** -24) * d[:, 0]
    cy = o[:, 1] + (f + 2.0 ** -24) * d[:, 1]
    tri = np.zeros((count, 3, 3), np.float32)
    tri[:, 0, :2] = np.stack([cx - 1.0, cy - 1.0], 1)
    tri[:, 1, :2] = np.stack([cx + 1.0, cy - 0.7], 1)
    tri[:, 2, :2] = np.stack([cx - 0.3, cy + 1.0], 1)
    tri[:, :, 2] = f[:, None]
    scene = abi.SceneData(tri.reshape(-1, 9), np.zeros(count, np.uint32), [abi.material(double_sided=True)])
    return scene, abi.make_rays(o, d)


def test_surface_from_float_pair_barycentrics_equals_the_binary64_chain(emu):
    """Shading's hit position, shading normal and float(t) from float-pair barycentrics
    with running error bounds (hit_uv_ff, surface_at_ff, hit_t_ff) equal those from the
    reference's binary64 chain (reference src/geometry.cpp:147-182 t and barycentrics,
    :401-415 surfaceAt) on every hit:
    random and aimed rays on terrain, sphere, soup and box scenes, rays at triangle edges and
    vertices from up to 1e4 scene sizes away, rays grazing the triangle planes, and hit
    distances on binary32 rounding midpoints.  The float-pair path must decide at least 95 %
    of the random rays' surfaces and 90 % of their float(t) (edge and grazing rays from afar
    are where its bounds give way to the chain)."""
    cases = [scenes.config2(grid=96, spheres=12, width=8, height=8).scene,
             scenes.config3(triangles=20_000, width=8, height=8).scene, scenes.config1(width=8, height=8).scene]
    hits = 0
    for k, sd in enumerate(cases):
        es = emu.scene(sd)
        for j, rays in enumerate((scenes.random_rays(sd, 20_000, 40 + k), _edge_rays(sd, 20_000, 50 + k))):
            ff, chain, decided = es.hit_surface(rays)
            hit = decided != 4
            assert not np.any(decided & 8), "a barycentric pair outside its error bound"
            bad = np.nonzero(np.any(ff.view(np.uint32) != chain.view(np.uint32), axis=1))[0]
            assert len(bad) == 0, [(rays[i], ff[i].tolist(), chain[i].tolist()) for i in bad[:2]]
            hits += int(hit.sum())
            if j == 0:
                for bit, floor in ((1, 0.95), (2, 0.9)):
                    got = ((decided & bit) != 0) & hit
                    assert got.sum() > floor * hit.sum(), (k, bit, got.sum() / hit.sum())
    assert hits > 60_000
    sd, rays = _midpoint_t_case(20_000, 60)
    ff, chain, decided = emu.scene(sd).hit_surface(rays)
    assert np.all(decided != 4)
    np.testing.assert_array_equal(ff.view(np.uint32), chain.view(np.uint32))


def test_camera_float_float_equals_the_binary64_chain(emu):
    """Camera::generate (reference src/geometry.cpp:213-234) in float-float with the binary64
    chain for hard cases equals the chain alone (direction, near and far parameters, accept)
    on 1.5 M pixels of ten cameras: axis-aligned and oblique bases, aspect from the image
    (0) and given, zero, tiny and large near clips, finite and infinite far clips, wide (up to
    3.141592 rad: corner rays nearly perpendicular to an oblique view axis, where the
    projection on it cancels) and narrow fields of view,
    random sub-pixel positions, pixel corners, image centres and points a few ulps off them."""
    fp = C.POINTER(C.c_float)
    g = np.random.Generator(np.random.PCG64(151))
    cams = [
        (abi.camera(), 640, 480),
        (abi.camera(near_clip=0.0, far_clip=float("inf")), 1920, 1080),
        (scenes.look_at((3, 4, 5), (0, 0.5, 0), 0.8), 1920, 1080),
        (scenes.look_at((-2.5, 1.0, 7.0), (0.3, -0.2, 0.1), 2.9), 333, 777),
        (abi.camera(yfov=0.01, near_clip=1e-6, far_clip=1e6), 512, 512),
        (abi.camera(forward=(0.3, -0.4, -0.86), right=(1, 0.2, 0.1), up=(0, 1, -0.5), aspect=2.35,
                    near_clip=0.5, far_clip=50), 1000, 425),
        (scenes.look_at((100, 200, -300), (0, 0, 0), 1.2), 4096, 2160),
        (abi.camera(near_clip=3e4, far_clip=3e38), 64, 64),
        (abi.camera(yfov=3.14, aspect=3.0, near_clip=0.25, far_clip=9.0), 800, 600),
        (abi.camera(forward=(0.3, -0.4, -0.86), right=(1, 0.2, 0.1), up=(0, 1, -0.5), yfov=3.141592,
                    aspect=3.0, near_clip=0.25, far_clip=9.0), 800, 600),
    ]
    for cam, w, h in cams:
        n = 150_000
        px = (g.uniform(0, 1, n) * w).astype(np.float32)
        py = (g.uniform(0, 1, n) * h).astype(np.float32)
        px[:1000] = np.float32(w / 2)
        py[1000:2000] = np.float32(h / 2)
        px[2000:3000] = g.integers(0, w + 1, 1000).astype(np.float32)
        # a few ulps off the centre lines: the subtraction of 1 cancels and the binary64
        # quotient's rounding becomes visible in the direction
        steps = g.integers(-64, 65, 20000).astype(np.float32)
        cx, cy = np.float32(w / 2), np.float32(h / 2)
        px[3000:23000] = cx + steps * (np.spacing(cx))
        py[23000:43000] = cy + steps * (np.spacing(cy))
        ff = np.empty((n, 6), np.float32)
        chain = np.empty_like(ff)
        rc = emu.lib.emu_camera_pair(C.byref(cam), px.ctypes.data_as(fp), py.ctypes.data_as(fp), n, w, h,
                                     ff.ctypes.data_as(fp), chain.ctypes.data_as(fp))
        assert rc == 0
        assert chain[:, 5].mean() > 0.99
        bad = np.nonzero(np.any(ff.view(np.uint32) != chain.view(np.uint32), axis=1))[0]
        assert len(bad) == 0, (w, h, [(px[i], py[i], ff[i].tolist(), chain[i].tolist()) for i in bad[:3]])
