"""Python host mirror of the reference renderer API over the crgpu C ABI.

`Context` owns a device context; `render()` mirrors
`cr::render(scene, bvh, camera, environment, options)` (reference
src/render.h:53) and `trace()` mirrors `Bvh::intersect` over a ray array
(reference src/scene.h:255-256).  Errors come back as exceptions carrying the
library's message (the reference's wording, e.g. "Pixel (x,y), sample s: ...").

The product path is the gfx950 library only: if lib/libcrgpu.so is missing or
no gfx950 device is visible, construction raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import pathlib
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import abi

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libcrgpu.so"

_lib = None


class CrgError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def load_library(path: Optional[pathlib.Path] = None) -> C.CDLL:
    """Load libcrgpu.so and declare every entry point of include/crgpu.h."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path or LIB_PATH)
    if not p.exists():
        raise FileNotFoundError(f"{p} is not built; run cpu-renderer_amd/build.py (or __graft_entry__.build())")
    lib = C.CDLL(str(p))
    vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int
    sig = {
        "crg_context_create": (i32, [i32, C.POINTER(vp)]),
        "crg_context_destroy": (None, [vp]),
        "crg_last_error": (C.c_char_p, [vp]),
        "crg_version": (C.c_char_p, []),
        "crg_scene_upload": (i32, [vp, C.POINTER(abi.SceneDesc), C.POINTER(u64), C.POINTER(abi.SceneInfo)]),
        "crg_scene_release": (i32, [vp, u64]),
        "crg_env_upload": (i32, [vp, C.POINTER(abi.EnvDesc), C.POINTER(u64), C.POINTER(u64)]),
        "crg_env_release": (i32, [vp, u64]),
        "crg_render": (i32, [vp, u64, u64, C.POINTER(abi.Camera), C.POINTER(abi.Options), C.POINTER(C.c_float),
                             C.POINTER(abi.Stats)]),
        "crg_trace_rays": (i32, [vp, u64, vp, u64, vp, i32]),
        "crg_set_profiling": (i32, [vp, i32]),
        "crg_kernel_times": (i32, [vp, C.POINTER(abi.KernelTime), i32, C.POINTER(i32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


@dataclass
class RenderResult:
    """Mirror of cr::RenderResult (reference src/render.h:41-48)."""
    pixels: np.ndarray       # (height, width, 3) float32 linear film
    seconds: float
    workers: int
    camera_samples: int
    boundary_retries: int
    extend_rays: int
    shadow_rays: int
    iterations: int
    extend_nodes: int = 0
    extend_tris: int = 0
    shadow_nodes: int = 0
    shadow_tris: int = 0
    exact_fallbacks: int = 0
    exact_rays: int = 0


class Context:
    def __init__(self, device: int = 0):
        self.lib = load_library()
        self.handle = C.c_void_p()
        rc = self.lib.crg_context_create(device, C.byref(self.handle))
        if rc != abi.CRG_OK:
            raise CrgError(rc, f"crg_context_create failed ({rc}): no gfx950 device {device}")

    def _check(self, rc: int):
        if rc != abi.CRG_OK:
            raise CrgError(rc, self.lib.crg_last_error(self.handle).decode())

    def close(self):
        if self.handle:
            self.lib.crg_context_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_scene(self, scene: abi.SceneData):
        desc = scene.descriptor()
        sid, info = C.c_uint64(), abi.SceneInfo()
        self._check(self.lib.crg_scene_upload(self.handle, C.byref(desc), C.byref(sid), C.byref(info)))
        return sid.value, info

    def release_scene(self, scene_id: int):
        self._check(self.lib.crg_scene_release(self.handle, scene_id))

    def upload_env(self, env: abi.EnvData):
        desc = env.descriptor()
        eid, nbytes = C.c_uint64(), C.c_uint64()
        self._check(self.lib.crg_env_upload(self.handle, C.byref(desc), C.byref(eid), C.byref(nbytes)))
        return eid.value, nbytes.value

    def release_env(self, env_id: int):
        self._check(self.lib.crg_env_release(self.handle, env_id))

    def render(self, scene_id: int, env_id: int, camera: abi.Camera, options: abi.Options) -> RenderResult:
        film = np.zeros((options.height, options.width, 3), np.float32)
        stats = abi.Stats()
        self._check(self.lib.crg_render(self.handle, scene_id, env_id, C.byref(camera), C.byref(options),
                                        film.ctypes.data_as(C.POINTER(C.c_float)), C.byref(stats)))
        return RenderResult(film, stats.seconds, stats.workers, stats.camera_samples, stats.boundary_retries,
                            stats.extend_rays, stats.shadow_rays, stats.iterations, stats.extend_nodes,
                            stats.extend_tris, stats.shadow_nodes, stats.shadow_tris, stats.exact_fallbacks,
                            stats.exact_rays)

    def set_profiling(self, enable: bool):
        self._check(self.lib.crg_set_profiling(self.handle, int(bool(enable))))

    def kernel_times(self) -> dict:
        arr = (abi.KernelTime * 16)()
        n = C.c_int()
        self._check(self.lib.crg_kernel_times(self.handle, arr, 16, C.byref(n)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(min(n.value, 16))}

    def trace(self, scene_id: int, rays: np.ndarray, any_hit: bool = False) -> np.ndarray:
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        hits = np.zeros(len(rays), abi.HIT_DTYPE)
        self._check(self.lib.crg_trace_rays(self.handle, scene_id, rays.ctypes.data, len(rays), hits.ctypes.data,
                                            int(bool(any_hit))))
        return hits


def render(ctx: Context, workload) -> RenderResult:
    """Render a scenes.Workload end to end (upload, render, release)."""
    sid, _ = ctx.upload_scene(workload.scene)
    eid, _ = ctx.upload_env(workload.env)
    try:
        return ctx.render(sid, eid, workload.camera, workload.options())
    finally:
        ctx.release_scene(sid)
        ctx.release_env(eid)
