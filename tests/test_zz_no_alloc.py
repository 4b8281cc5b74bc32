"""crg_trace_rays_noalloc: Bvh::intersect over a host ray array with its
staging in the library's fixed per-device scratch array (no device
allocation).  Called through the C ABI twice in a row — the second call with
fewer rays, so stale contents of the shared scratch would show — its hits
equal crg_trace_rays's bit for bit (flags, triangle ids, binary64 t/u/v), for
closest-hit and any-hit queries; a batch above the scratch's capacity is
refused with CRG_EINVAL before anything is launched."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import abi, scenes

pytestmark = pytest.mark.gpu


def _equal(a, b, what):
    np.testing.assert_array_equal(a["hit"], b["hit"], err_msg=f"{what}: hit flags")
    np.testing.assert_array_equal(a["triangle"], b["triangle"], err_msg=f"{what}: triangle ids")
    for k in ("t", "u", "v"):
        np.testing.assert_array_equal(a[k].view(np.uint64), b[k].view(np.uint64), err_msg=f"{what}: {k}")


def test_noalloc_trace_equals_trace_on_two_calls(gpu):
    w = scenes.config1()
    big = scenes.random_rays(w.scene, 65536, seed=11)
    small = scenes.random_rays(w.scene, 3000, seed=12)
    sid, _ = gpu.upload_scene(w.scene)
    try:
        for any_hit in (False, True):
            first = gpu.trace_noalloc(sid, big, any_hit)
            second = gpu.trace_noalloc(sid, small, any_hit)
            _equal(first, gpu.trace(sid, big, any_hit), f"65536 rays any={any_hit}")
            _equal(second, gpu.trace(sid, small, any_hit), f"3000 rays after 65536 any={any_hit}")
            assert int(first["hit"].sum()) > 1000 and int(second["hit"].sum()) > 50
        assert len(gpu.trace_noalloc(sid, small[:0])) == 0
    finally:
        gpu.release_scene(sid)


def test_noalloc_trace_refuses_batches_above_capacity(gpu):
    from cr_amd.crgpu import CrgError
    w = scenes.config1(width=16, height=16, samples=1)
    rays = scenes.random_rays(w.scene, 65537, seed=13)
    sid, _ = gpu.upload_scene(w.scene)
    try:
        with pytest.raises(CrgError, match="exceed the 65536-ray scratch") as e:
            gpu.trace_noalloc(sid, rays)
        assert e.value.code == abi.CRG_EINVAL
    finally:
        gpu.release_scene(sid)
