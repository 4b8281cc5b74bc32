"""Scene cache files: a scene made resident from its saved device-format arrays
renders bit-identically to the scene it was saved from (no finalize or BVH build
on load), and damaged or foreign files are refused with a clear message."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import scenes

pytestmark = pytest.mark.gpu


def _cases():
    import ref_scenes as rs
    out = [("config1", scenes.config1(width=48, height=40, samples=4))]
    for name, fn in (("tex_factors", rs.textured_factors), ("tex_normal", rs.textured_normal_map),
                     ("tex_mask", rs.textured_alpha_mask), ("gltf_contract", rs.gltf_export_contract)):
        sc, env, cam = fn()
        out.append((name, (sc, env, cam)))
    return out


@pytest.mark.parametrize("name", [c[0] for c in _cases()])
def test_loaded_scene_renders_identically(gpu, tmp_path, name):
    from cr_amd import abi
    w = dict(_cases())[name]
    if isinstance(w, tuple):
        sc, env, cam = w
        opt = abi.options(40, 32, 4, 6, threads=1, seed=3)
    else:
        sc, env, cam, opt = w.scene, w.env, w.camera, w.options()
    sid, info = gpu.upload_scene(sc)
    eid, _ = gpu.upload_env(env)
    path = tmp_path / "scene.crgscene"
    try:
        a = gpu.render(sid, eid, cam, opt)
        gpu.save_scene(sid, path)
        lid, linfo = gpu.load_scene(path)
        try:
            assert (linfo.triangles, linfo.wide_nodes, linfo.bvh_nodes) == (info.triangles, info.wide_nodes,
                                                                            info.bvh_nodes)
            b = gpu.render(lid, eid, cam, opt)
            np.testing.assert_array_equal(a.pixels.view(np.uint32), b.pixels.view(np.uint32))
            for k in ("camera_samples", "boundary_retries", "extend_rays", "shadow_rays"):
                assert getattr(a, k) == getattr(b, k), k
            assert gpu.trace(lid, _rays(sc)).tobytes() == gpu.trace(sid, _rays(sc)).tobytes()
        finally:
            gpu.release_scene(lid)
    finally:
        gpu.release_scene(sid)
        gpu.release_env(eid)


def _rays(scene):
    return scenes.random_rays(scene, 500, seed=9, spread=1.2)


def test_damaged_files_are_refused(gpu, tmp_path):
    from cr_amd.crgpu import CrgError
    w = scenes.config1(width=16, height=16, samples=1)
    sid, _ = gpu.upload_scene(w.scene)
    path = tmp_path / "s.crgscene"
    try:
        gpu.save_scene(sid, path)
    finally:
        gpu.release_scene(sid)
    data = bytearray(path.read_bytes())
    bad = tmp_path / "bad.crgscene"
    data[len(data) // 2] ^= 0xFF
    bad.write_bytes(bytes(data))
    with pytest.raises(CrgError, match="checksum mismatch"):
        gpu.load_scene(bad)
    bad.write_bytes(b"not a scene")
    with pytest.raises(CrgError, match="not a scene cache"):
        gpu.load_scene(bad)
    with pytest.raises(CrgError, match="cannot open"):
        gpu.load_scene(tmp_path / "missing.crgscene")


def test_host_written_cache_loads_and_renders_like_an_upload(gpu, tmp_path):
    """crg_scene_write_cache (no device) writes the file crg_scene_load makes
    resident; its renders equal those of the uploaded scene, and a file whose
    index was edited (checksum recomputed) is refused by the loader too."""
    import ref_scenes as rs
    import test_scene_cache_host as th
    from cr_amd import abi, crgpu
    sc, env, cam = rs.textured_factors()
    opt = abi.options(40, 32, 4, 6, threads=1, seed=5)
    path = tmp_path / "host.crgscene"
    crgpu.write_scene_cache(sc, path)
    sid, _ = gpu.upload_scene(sc)
    eid, _ = gpu.upload_env(env)
    lid, _ = gpu.load_scene(path)
    try:
        a = gpu.render(sid, eid, cam, opt)
        b = gpu.render(lid, eid, cam, opt)
        np.testing.assert_array_equal(a.pixels.view(np.uint32), b.pixels.view(np.uint32))
    finally:
        gpu.release_scene(lid)
        gpu.release_scene(sid)
        gpu.release_env(eid)
    f = th.CacheFile(path.read_bytes())
    f.put_u32("prims", 0, 44, 77)
    bad = tmp_path / "bad.crgscene"
    bad.write_bytes(f.sealed())
    with pytest.raises(crgpu.CrgError, match="primitive record refers outside"):
        gpu.load_scene(bad)
