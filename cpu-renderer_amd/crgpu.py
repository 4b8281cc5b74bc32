"""Python host mirror of the reference renderer API over the crgpu C ABI.

`Context` owns a device context; `render()` mirrors
`cr::render(scene, bvh, camera, environment, options)` (reference
src/render.h:53) and `trace()` mirrors `Bvh::intersect` over a ray array
(reference src/scene.h:126-127).  Errors come back as exceptions carrying the
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


def write_scene_cache(scene: abi.SceneData, path, lib: Optional[C.CDLL] = None) -> abi.SceneInfo:
    """Scene::finalize + Bvh::build on the host, written to a scene cache file
    without a device (crg_scene_write_cache)."""
    lib = lib or load_library()
    desc = scene.descriptor()
    info = abi.SceneInfo()
    msg = C.create_string_buffer(1024)
    rc = lib.crg_scene_write_cache(C.byref(desc), str(path).encode(), C.byref(info), msg, len(msg))
    if rc != 0:
        raise CrgError(rc, msg.value.decode(errors="replace"))
    return info


def check_scene_cache(path, lib: Optional[C.CDLL] = None) -> None:
    """The checks crg_scene_load makes on a cache file, without a device
    (crg_scene_check); raises CrgError with the reason."""
    lib = lib or load_library()
    msg = C.create_string_buffer(1024)
    rc = lib.crg_scene_check(str(path).encode(), msg, len(msg))
    if rc != 0:
        raise CrgError(rc, msg.value.decode(errors="replace"))


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
        "crg_device_probe": (i32, [vp, i32, u64, u64, C.POINTER(abi.Camera), i32, i32, vp, u64, vp]),
        "crg_render_window": (i32, [vp, u64, u64, C.POINTER(abi.Camera), C.POINTER(abi.Options), i32, i32, i32, i32,
                                    C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(abi.Stats)]),
        "crg_trace_rays_noalloc": (i32, [vp, u64, vp, u64, vp, i32]),
        "crg_render_batch": (i32, [vp, C.POINTER(abi.Request), C.c_uint32, C.POINTER(i32), C.POINTER(abi.Stats)]),
        "crg_submit": (i32, [vp, C.POINTER(abi.Request), C.POINTER(u64)]),
        "crg_wait": (i32, [vp, u64, C.c_double, C.POINTER(abi.Stats)]),
        "crg_set_memory_budget": (i32, [vp, u64]),
        "crg_request_progress": (i32, [vp, u64, C.POINTER(u64), C.POINTER(u64)]),
        "crg_engine_info_get": (i32, [vp, C.POINTER(abi.EngineInfo)]),
        "crg_get_tuning": (i32, [vp, C.POINTER(abi.Tuning)]),
        "crg_set_tuning": (i32, [vp, C.POINTER(abi.Tuning)]),
        "crg_autotune": (i32, [vp, C.POINTER(abi.Request), C.POINTER(abi.Tuning), C.POINTER(C.c_double)]),
        "crg_set_profiling": (i32, [vp, i32]),
        "crg_set_work_counters": (i32, [vp, i32]),
        "crg_kernel_times": (i32, [vp, C.POINTER(abi.KernelTime), i32, C.POINTER(i32)]),
        "crg_scene_save": (i32, [vp, u64, C.c_char_p]),
        "crg_scene_load": (i32, [vp, C.c_char_p, C.POINTER(u64), C.POINTER(abi.SceneInfo)]),
        "crg_scene_write_cache": (i32, [C.POINTER(abi.SceneDesc), C.c_char_p, C.POINTER(abi.SceneInfo), C.c_char_p,
                                        C.c_size_t]),
        "crg_scene_check": (i32, [C.c_char_p, C.c_char_p, C.c_size_t]),
        "crg_film_encode": (i32, [vp, C.POINTER(C.c_float), i32, i32, C.c_float, C.POINTER(C.c_uint8),
                                  C.POINTER(C.c_uint8), u64, C.POINTER(u64)]),
        "crg_film_hdr_bound": (u64, [i32, i32]),
        "crg_write_png": (i32, [vp, C.c_char_p, C.POINTER(C.c_float), i32, i32, C.c_float]),
        "crg_write_hdr": (i32, [vp, C.c_char_p, C.POINTER(C.c_float), i32, i32]),
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
    queue_seconds: float = 0.0
    latency_seconds: float = 0.0

    @staticmethod
    def from_stats(film: np.ndarray, stats: abi.Stats) -> "RenderResult":
        return RenderResult(film, stats.seconds, stats.workers, stats.camera_samples, stats.boundary_retries,
                            stats.extend_rays, stats.shadow_rays, stats.iterations, stats.extend_nodes,
                            stats.extend_tris, stats.shadow_nodes, stats.shadow_tris, stats.exact_fallbacks,
                            stats.exact_rays, stats.queue_seconds, stats.latency_seconds)


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
        return RenderResult.from_stats(film, stats)

    def render_window(self, scene_id: int, env_id: int, camera: abi.Camera, options: abi.Options, window,
                      with_samples: bool = False):
        """crg_render_window: the pixels of window (x0, y0, w, h) of the frame with their
        full-frame sample streams; with_samples also returns every sample's radiance
        ([spp, h, w, 3]).  Returns the RenderResult (and the samples)."""
        x0, y0, w, h = window
        film = np.zeros((h, w, 3), np.float32)
        smp = np.zeros((options.samples, h, w, 3), np.float32) if with_samples else None
        stats = abi.Stats()
        self._check(self.lib.crg_render_window(self.handle, scene_id, env_id, C.byref(camera), C.byref(options),
                                               x0, y0, w, h, film.ctypes.data_as(C.POINTER(C.c_float)),
                                               smp.ctypes.data_as(C.POINTER(C.c_float)) if smp is not None else None,
                                               C.byref(stats)))
        r = RenderResult.from_stats(film, stats)
        return (r, smp) if with_samples else r

    # ---- batched and served forms -------------------------------------------------
    @staticmethod
    def request(scene_id: int, env_id: int, camera: abi.Camera, options: abi.Options):
        """(crg_request, film) with a freshly allocated film the request writes into."""
        film = np.zeros((options.height, options.width, 3), np.float32)
        req = abi.Request(scene_id, env_id, camera, options, film.ctypes.data_as(C.POINTER(C.c_float)))
        return req, film

    def render_batch(self, requests) -> list:
        """Render [(scene_id, env_id, camera, options), ...] together; returns RenderResults
        (or CrgError instances for requests that failed)."""
        made = [self.request(*r) for r in requests]
        arr = (abi.Request * len(made))(*[m[0] for m in made])
        status = (C.c_int * len(made))()
        stats = (abi.Stats * len(made))()
        rc = self.lib.crg_render_batch(self.handle, arr, len(made), status, stats)
        if rc != abi.CRG_OK and not any(status[i] != abi.CRG_OK for i in range(len(made))):
            raise CrgError(rc, self.lib.crg_last_error(self.handle).decode())
        first = self.lib.crg_last_error(self.handle).decode() if rc != abi.CRG_OK else ""
        out = []
        for i, (_, film) in enumerate(made):
            if status[i] == abi.CRG_OK:
                out.append(RenderResult.from_stats(film, stats[i]))
            else:
                msg = first if first.startswith(f"request {i}:") else f"request {i} failed (status {status[i]})"
                out.append(CrgError(status[i], msg))
        return out

    def submit(self, scene_id: int, env_id: int, camera: abi.Camera, options: abi.Options):
        """Queue a request; returns (ticket, film).  The film array is written when the request completes."""
        req, film = self.request(scene_id, env_id, camera, options)
        ticket = C.c_uint64()
        self._check(self.lib.crg_submit(self.handle, C.byref(req), C.byref(ticket)))
        return ticket.value, film

    def wait(self, ticket: int, film: np.ndarray, timeout: float = -1.0) -> Optional[RenderResult]:
        """RenderResult when done, None if still in flight after `timeout` seconds."""
        stats = abi.Stats()
        rc = self.lib.crg_wait(self.handle, ticket, float(timeout), C.byref(stats))
        if rc == abi.CRG_EBUSY:
            return None
        self._check(rc)
        return RenderResult.from_stats(film, stats)

    def progress(self, ticket: int):
        """(camera samples finished, total) of a submitted request."""
        done, total = C.c_uint64(), C.c_uint64()
        self._check(self.lib.crg_request_progress(self.handle, ticket, C.byref(done), C.byref(total)))
        return done.value, total.value

    def set_memory_budget(self, nbytes: int):
        self._check(self.lib.crg_set_memory_budget(self.handle, int(nbytes)))

    def engine_info(self) -> abi.EngineInfo:
        info = abi.EngineInfo()
        self._check(self.lib.crg_engine_info_get(self.handle, C.byref(info)))
        return info

    def get_tuning(self) -> abi.Tuning:
        t = abi.Tuning()
        self._check(self.lib.crg_get_tuning(self.handle, C.byref(t)))
        return t

    def set_tuning(self, tuning: abi.Tuning):
        self._check(self.lib.crg_set_tuning(self.handle, C.byref(tuning)))

    def autotune(self, scene_id: int, env_id: int, camera: abi.Camera, options: abi.Options):
        """Pick the fastest launch configuration on a sample request; returns (Tuning, samples/s)."""
        req, _film = self.request(scene_id, env_id, camera, options)
        t, rate = abi.Tuning(), C.c_double()
        self._check(self.lib.crg_autotune(self.handle, C.byref(req), C.byref(t), C.byref(rate)))
        return t, rate.value

    def set_profiling(self, enable: bool):
        self._check(self.lib.crg_set_profiling(self.handle, int(bool(enable))))

    def set_work_counters(self, enable: bool):
        """Per-query work counters (node visits, triangle tests) in the stats of later renders."""
        self._check(self.lib.crg_set_work_counters(self.handle, int(bool(enable))))

    def kernel_times(self) -> dict:
        arr = (abi.KernelTime * 16)()
        n = C.c_int()
        self._check(self.lib.crg_kernel_times(self.handle, arr, 16, C.byref(n)))
        return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(min(n.value, 16))}

    # --- scene cache file (device-format arrays of a resident scene)
    def save_scene(self, scene_id: int, path) -> None:
        self._check(self.lib.crg_scene_save(self.handle, scene_id, str(path).encode()))

    def load_scene(self, path):
        sid = C.c_uint64()
        info = abi.SceneInfo()
        self._check(self.lib.crg_scene_load(self.handle, str(path).encode(), C.byref(sid), C.byref(info)))
        return sid.value, info

    # --- film encode (io::writePng / io::writeHdr, reference src/io/image.cpp:201-236)
    def film_png8(self, film: np.ndarray, exposure: float = 0.0) -> np.ndarray:
        """Display bytes (H, W, 3) that writePng stores for a linear film."""
        film = np.ascontiguousarray(film, np.float32)
        h, w = film.shape[:2]
        out = np.zeros((h, w, 3), np.uint8)
        self._check(self.lib.crg_film_encode(self.handle, film.ctypes.data_as(C.POINTER(C.c_float)), w, h, exposure,
                                             out.ctypes.data_as(C.POINTER(C.c_uint8)), None, 0, None))
        return out

    def film_hdr(self, film: np.ndarray) -> bytes:
        """The complete Radiance .hdr file writeHdr writes for a linear film."""
        film = np.ascontiguousarray(film, np.float32)
        h, w = film.shape[:2]
        cap = max(int(self.lib.crg_film_hdr_bound(w, h)), 1)
        out = np.zeros(cap, np.uint8)
        size = C.c_uint64()
        self._check(self.lib.crg_film_encode(self.handle, film.ctypes.data_as(C.POINTER(C.c_float)), w, h, 0.0, None,
                                             out.ctypes.data_as(C.POINTER(C.c_uint8)), cap, C.byref(size)))
        return bytes(out[:size.value])

    def write_png(self, path, film: np.ndarray, exposure: float = 0.0) -> None:
        film = np.ascontiguousarray(film, np.float32)
        h, w = film.shape[:2]
        self._check(self.lib.crg_write_png(self.handle, str(path).encode(), film.ctypes.data_as(C.POINTER(C.c_float)),
                                           w, h, exposure))

    def write_hdr(self, path, film: np.ndarray) -> None:
        film = np.ascontiguousarray(film, np.float32)
        h, w = film.shape[:2]
        self._check(self.lib.crg_write_hdr(self.handle, str(path).encode(), film.ctypes.data_as(C.POINTER(C.c_float)),
                                           w, h))

    def trace(self, scene_id: int, rays: np.ndarray, any_hit: bool = False) -> np.ndarray:
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        hits = np.zeros(len(rays), abi.HIT_DTYPE)
        self._check(self.lib.crg_trace_rays(self.handle, scene_id, rays.ctypes.data, len(rays), hits.ctypes.data,
                                            int(bool(any_hit))))
        return hits

    def probe(self, what: int, records: np.ndarray, out_bytes: int, scene_id: int = 0, env_id: int = 0,
              camera: Optional[abi.Camera] = None, width: int = 0, height: int = 0) -> np.ndarray:
        """crg_device_probe: one device function over `records` (a contiguous array of the
        function's input records); returns the output records as raw bytes (n, out_bytes)."""
        records = np.ascontiguousarray(records)
        n = len(records)
        out = np.zeros((n, out_bytes), np.uint8)
        self._check(self.lib.crg_device_probe(self.handle, what, scene_id, env_id,
                                              C.byref(camera) if camera is not None else None, width, height,
                                              records.ctypes.data, n, out.ctypes.data))
        return out

    def trace_noalloc(self, scene_id: int, rays: np.ndarray, any_hit: bool = False) -> np.ndarray:
        """crg_trace_rays_noalloc: the same queries through the library's fixed
        per-device scratch (at most 65536 rays per call, no device allocation)."""
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        hits = np.zeros(len(rays), abi.HIT_DTYPE)
        self._check(self.lib.crg_trace_rays_noalloc(self.handle, scene_id, rays.ctypes.data, len(rays),
                                                    hits.ctypes.data, int(bool(any_hit))))
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
