"""Batched and served forms on the GPU: identical results alone and in a mix.

Every request is a job of one shared wavefront pool.  These tests check the
determinism contract of the reference (each pixel has a fixed sample order on
any worker count, reference src/main.cpp:50, src/render.cpp:384-415; the
1-vs-N worker bit-identity test, reference tests/test_integrator.cpp:300-307)
in its device form: a request's film and counters are bit-identical whether it
runs alone, in a batch of other scenes and sizes, submitted while others are
in flight, under a tight memory budget, or under any launch configuration.
Work counters (node visits, triangle tests, exact reruns) may differ: see _same.
"""
from __future__ import annotations

import dataclasses
import time

import numpy as np
import pytest

from conftest import abi, scenes

pytestmark = pytest.mark.gpu


def _mix():
    """Small requests of different scenes, sizes, spp and depths."""
    out = [dataclasses.replace(scenes.config1(), width=64, height=48, samples=4)]
    sky = scenes.synthetic_sky(256, 128)
    for seed, (w, h, spp) in zip((3, 4, 5, 6), ((40, 40, 3), (72, 24, 2), (16, 16, 9), (33, 17, 5))):
        r = scenes.batch_request(seed, sky=sky, max_triangles=20_000)
        out.append(dataclasses.replace(r, width=w, height=h, samples=spp))
    return out


@pytest.fixture(scope="module")
def resident(gpu):
    gpu.set_work_counters(True)  # _same checks that the counting kernels count
    items = []
    for w in _mix():
        sid, _ = gpu.upload_scene(w.scene)
        eid, _ = gpu.upload_env(w.env)
        items.append((w, sid, eid))
    yield items
    for _, sid, eid in items:
        gpu.release_scene(sid)
        gpu.release_env(eid)
    gpu.set_work_counters(False)


@pytest.fixture(scope="module")
def alone(gpu, resident):
    return [gpu.render(sid, eid, w.camera, w.options()) for w, sid, eid in resident]


def _same(a, b, what):
    """Film bits and the reference's own counters are identical.  The work
    counters (node visits, triangle tests, exact reruns) are not: a wave
    interleaves node and leaf visits by the state of all its lanes, so the
    order in which a query meets nodes, and with it how many it can skip,
    depends on which other queries share the wave (the result does not)."""
    assert not isinstance(b, Exception), f"{what}: {b}"
    np.testing.assert_array_equal(a.pixels.view(np.uint32), b.pixels.view(np.uint32), err_msg=f"{what}: film bits")
    for k in ("camera_samples", "boundary_retries", "extend_rays", "shadow_rays"):
        assert getattr(a, k) == getattr(b, k), f"{what}: {k}"
    for k in ("extend_nodes", "extend_tris", "shadow_nodes", "shadow_tris"):
        assert getattr(b, k) > 0, f"{what}: {k}"


def test_alone_renders_are_nontrivial(alone):
    for r in alone:
        assert np.isfinite(r.pixels).all() and r.pixels.max() > 0
        assert r.extend_rays >= r.camera_samples


def test_batch_equals_alone(gpu, resident, alone):
    res = gpu.render_batch([(sid, eid, w.camera, w.options()) for w, sid, eid in resident])
    for (w, _, _), a, b in zip(resident, alone, res):
        assert not isinstance(b, Exception), b
        _same(a, b, w.name)


def test_batch_with_repeats_equals_alone(gpu, resident, alone):
    order = [2, 0, 4, 0, 1, 3, 2]
    res = gpu.render_batch([(resident[i][1], resident[i][2], resident[i][0].camera, resident[i][0].options())
                            for i in order])
    for i, b in zip(order, res):
        _same(alone[i], b, resident[i][0].name)


def test_served_staggered_submission_equals_alone(gpu, resident, alone):
    tickets = []
    for k, (w, sid, eid) in enumerate(resident):
        tickets.append(gpu.submit(sid, eid, w.camera, w.options()))
        if k == 1:
            time.sleep(0.02)  # later requests arrive while the first ones are in flight
    for (w, _, _), a, (t, film) in zip(resident, alone, tickets):
        b = gpu.wait(t, film)
        _same(a, b, w.name)
        assert b.latency_seconds >= b.seconds >= 0


def test_results_do_not_depend_on_launch_configuration(gpu, resident, alone):
    """Pool size, workgroup size, workgroups per CU, iterations per graph and BVH width
    (the 4-wide tree or its 8-wide collapse) change speed only."""
    default = gpu.get_tuning()
    try:
        for t in (abi.Tuning(4096, 64, 2, 2, 4), abi.Tuning(1 << 16, 128, 4, 6, 4), abi.Tuning(4096, 64, 2, 2, 8),
                  abi.Tuning(1 << 16, 256, 8, 4, 8)):
            gpu.set_tuning(t)
            assert gpu.get_tuning().as_dict() == t.as_dict()
            res = gpu.render_batch([(sid, eid, w.camera, w.options()) for w, sid, eid in resident])
            for (w, _, _), a, b in zip(resident, alone, res):
                _same(a, b, f"{w.name} under {t.as_dict()}")
    finally:
        gpu.set_tuning(default)


def test_memory_budget_admits_one_request_at_a_time(gpu, resident, alone):
    info = gpu.engine_info()
    biggest = max(w.camera_samples * 16 + w.width * w.height * 12 + 4096 for w, _, _ in resident)
    try:
        gpu.set_memory_budget(info.resident_bytes + info.inflight_bytes + biggest)
        res = gpu.render_batch([(sid, eid, w.camera, w.options()) for w, sid, eid in resident])
        for (w, _, _), a, b in zip(resident, alone, res):
            _same(a, b, w.name)
            assert b.latency_seconds >= b.queue_seconds
    finally:
        gpu.set_memory_budget(info.budget_bytes)
    assert gpu.engine_info().inflight == 0


def test_request_larger_than_budget_is_refused(gpu, resident):
    from cr_amd.crgpu import CrgError
    w, sid, eid = resident[0]
    info = gpu.engine_info()
    try:
        gpu.set_memory_budget(info.resident_bytes + 1024)
        with pytest.raises(CrgError) as e:
            gpu.submit(sid, eid, w.camera, w.options())
        assert e.value.code == abi.CRG_ENOMEM
    finally:
        gpu.set_memory_budget(info.budget_bytes)


def test_path_error_fails_only_its_request(gpu, resident, alone, oracle):
    import ref_scenes as rs
    from cr_amd.crgpu import CrgError
    scene, env, _ = rs.touching_boundary()
    cam = abi.camera(position=(20, -0.5, 0), forward=(0, 1, 0), right=(1, 0, 0), up=(0, 0, 1), yfov=1e-7)
    opt = abi.options(2, 2, 4, 8, threads=1, seed=42)
    ps, pe = oracle.scene(scene), oracle.env(env)
    with pytest.raises(RuntimeError) as ref_err:
        ps.render(pe, cam, opt)
    sid, _ = gpu.upload_scene(scene)
    eid, _ = gpu.upload_env(env)
    try:
        with pytest.raises(CrgError) as alone_err:
            gpu.render(sid, eid, cam, opt)
        assert alone_err.value.code == abi.CRG_EPATH
        assert str(alone_err.value) == str(ref_err.value)
        w0, s0, e0 = resident[0]
        t_bad = gpu.submit(sid, eid, cam, opt)
        t_ok = gpu.submit(s0, e0, w0.camera, w0.options())
        with pytest.raises(CrgError) as mixed_err:
            gpu.wait(*t_bad)
        assert str(mixed_err.value) == str(ref_err.value)
        _same(alone[0], gpu.wait(*t_ok), w0.name)
    finally:
        gpu.release_scene(sid)
        gpu.release_env(eid)


def test_scene_released_while_in_flight_stays_valid(gpu, alone):
    w = _mix()[0]
    sid, _ = gpu.upload_scene(w.scene)
    eid, _ = gpu.upload_env(w.env)
    t = gpu.submit(sid, eid, w.camera, w.options())
    gpu.release_scene(sid)
    gpu.release_env(eid)
    _same(alone[0], gpu.wait(*t), w.name)


def test_autotune_returns_a_valid_configuration(gpu, resident, alone):
    w, sid, eid = resident[0]
    default = gpu.get_tuning()
    try:
        tuned, rate = gpu.autotune(sid, eid, w.camera, w.options())
        assert rate > 0
        assert tuned.workgroup in (64, 128, 256) and tuned.bvh_width in (4, 8)
        _same(alone[0], gpu.render(sid, eid, w.camera, w.options()), w.name)
    finally:
        gpu.set_tuning(default)


def test_alpha_mask_scene_in_a_batch_equals_alone(gpu, resident, alone):
    """A scene with alpha masks switches the traversal kernels to their variant with FP32
    mask tests while its job is live; requests of mask-free scenes batched with it, and the
    mask scene itself, render exactly as they do alone (the export-contract glTF asset has an
    alpha-masked material)."""
    import ref_scenes as rs
    sc, env, cam = rs.gltf_export_contract()
    sid, _ = gpu.upload_scene(sc)
    eid, _ = gpu.upload_env(env)
    try:
        opt = abi.options(64, 36, 4, 6, seed=9)
        mask_alone = gpu.render(sid, eid, cam, opt)
        reqs = [(s, e, w.camera, w.options()) for w, s, e in resident]
        reqs.insert(2, (sid, eid, cam, opt))
        res = gpu.render_batch(reqs)
        _same(mask_alone, res[2], "mask scene in a batch")
        for k, r in enumerate(res[:2] + res[3:]):
            _same(alone[k], r, f"request {k} batched with a mask scene")
    finally:
        gpu.release_scene(sid)
        gpu.release_env(eid)
