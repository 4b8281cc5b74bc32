"""ctypes mirror of include/crgpu.h (the C ABI of the gfx950 path tracer).

The structs here are byte-for-byte the C descriptors; `SceneData.descriptor()`
turns numpy-backed scene arrays into a `crg_scene_desc` that stays valid while
the SceneData object is alive.  The same descriptors feed the GPU library, the
CPU oracle and the reference harness, so every comparison sees identical input.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

CRG_OK, CRG_EINVAL, CRG_ENOMEM, CRG_EPATH, CRG_EDEVICE, CRG_ENOTFOUND, CRG_EBUSY = range(7)
NO_TRIANGLE = 0xFFFFFFFF
INF = float("inf")

F3 = C.c_float * 3
F4 = C.c_float * 4
F2 = C.c_float * 2


class TextureSlot(C.Structure):
    _fields_ = [("texture", C.c_int32), ("texcoord", C.c_int32), ("offset", F2), ("scale", F2),
                ("rotation", C.c_float)]


def _no_texture() -> TextureSlot:
    return TextureSlot(-1, 0, F2(0, 0), F2(1, 1), 0.0)


class Material(C.Structure):
    """crg_material — defaults follow the reference Material (src/scene.h:58-70)."""
    _fields_ = [("base_color", F4), ("metallic", C.c_float), ("roughness", C.c_float),
                ("emission", F3), ("transmission", C.c_float), ("ior", C.c_float),
                ("thickness", C.c_float), ("attenuation_distance", C.c_float),
                ("attenuation_color", F3), ("double_sided", C.c_int32), ("alpha_mode", C.c_int32),
                ("alpha_cutoff", C.c_float), ("normal_scale", C.c_float),
                ("base_color_tex", TextureSlot), ("metallic_roughness_tex", TextureSlot),
                ("normal_tex", TextureSlot), ("emissive_tex", TextureSlot), ("transmission_tex", TextureSlot)]


def material(base_color=(1, 1, 1, 1), metallic=1.0, roughness=1.0, emission=(0, 0, 0), transmission=0.0,
             ior=1.5, thickness=0.0, attenuation_distance=INF, attenuation_color=(1, 1, 1),
             double_sided=False, alpha_mode=0, alpha_cutoff=0.5, normal_scale=1.0) -> Material:
    bc = list(base_color) + [1.0] * (4 - len(base_color))
    return Material(F4(*bc), metallic, roughness, F3(*emission), transmission, ior, thickness,
                    attenuation_distance, F3(*attenuation_color), int(bool(double_sided)), int(alpha_mode),
                    alpha_cutoff, normal_scale, _no_texture(), _no_texture(), _no_texture(), _no_texture(),
                    _no_texture())


class Light(C.Structure):
    _fields_ = [("type", C.c_int32), ("position", F3), ("direction", F3), ("color", F3),
                ("intensity", C.c_float), ("range", C.c_float), ("inner_cone", C.c_float),
                ("outer_cone", C.c_float)]


POINT, DIRECTIONAL, SPOT = 0, 1, 2


def light(type=POINT, position=(0, 0, 0), direction=(0, 0, -1), color=(1, 1, 1), intensity=1.0, range=INF,
          inner_cone=0.0, outer_cone=3.14159265358979323846 / 4) -> Light:
    return Light(type, F3(*position), F3(*direction), F3(*color), intensity, range, inner_cone, outer_cone)


class Image(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgba", C.POINTER(C.c_float))]


class Texture(C.Structure):
    _fields_ = [("image", C.c_int32), ("wrap_s", C.c_int32), ("wrap_t", C.c_int32), ("linear", C.c_int32)]


class SceneDesc(C.Structure):
    _fields_ = [("triangle_count", C.c_uint64), ("positions", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("tangents", C.POINTER(C.c_float)),
                ("uvs", C.POINTER(C.c_float)), ("material_ids", C.POINTER(C.c_uint32)),
                ("material_count", C.c_uint32), ("materials", C.POINTER(Material)),
                ("light_count", C.c_uint32), ("lights", C.POINTER(Light)),
                ("image_count", C.c_uint32), ("images", C.POINTER(Image)),
                ("texture_count", C.c_uint32), ("textures", C.POINTER(Texture))]


class Camera(C.Structure):
    _fields_ = [("position", F3), ("forward", F3), ("right", F3), ("up", F3), ("yfov", C.c_float),
                ("aspect", C.c_float), ("near_clip", C.c_float), ("far_clip", C.c_float)]


def camera(position=(0, 0, 0), forward=(0, 0, -1), right=(1, 0, 0), up=(0, 1, 0),
           yfov=3.14159265358979323846 / 4, aspect=0.0, near_clip=0.01, far_clip=INF) -> Camera:
    return Camera(F3(*position), F3(*forward), F3(*right), F3(*up), yfov, aspect, near_clip, far_clip)


class EnvDesc(C.Structure):
    _fields_ = [("constant", F3), ("width", C.c_int32), ("height", C.c_int32),
                ("rgba", C.POINTER(C.c_float)), ("rotation_degrees", C.c_float), ("intensity", C.c_float)]


class Options(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("samples", C.c_int32), ("bounces", C.c_int32),
                ("threads", C.c_int32), ("next_event_estimation", C.c_int32),
                ("russian_roulette", C.c_int32), ("seed", C.c_uint64)]


def options(width=960, height=540, samples=32, bounces=8, threads=0, seed=42, nee=True, rr=True) -> Options:
    return Options(width, height, samples, bounces, threads, int(bool(nee)), int(bool(rr)), seed)


class Stats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("camera_samples", C.c_uint64), ("boundary_retries", C.c_uint64),
                ("extend_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("workers", C.c_int32),
                ("iterations", C.c_int32), ("extend_nodes", C.c_uint64), ("extend_tris", C.c_uint64),
                ("shadow_nodes", C.c_uint64), ("shadow_tris", C.c_uint64), ("exact_fallbacks", C.c_uint64),
                ("exact_rays", C.c_uint64), ("queue_seconds", C.c_double), ("latency_seconds", C.c_double)]


class Request(C.Structure):
    """crg_request: one render of the batched / served forms."""
    _fields_ = [("scene", C.c_uint64), ("env", C.c_uint64), ("camera", Camera), ("options", Options),
                ("film_rgb", C.POINTER(C.c_float))]


class Tuning(C.Structure):
    _fields_ = [("pool_paths", C.c_uint32), ("workgroup", C.c_uint32), ("groups_per_cu", C.c_uint32),
                ("iterations_per_graph", C.c_uint32), ("bvh_width", C.c_uint32)]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class EngineInfo(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("resident_bytes", C.c_uint64), ("inflight_bytes", C.c_uint64),
                ("inflight", C.c_uint32), ("pending", C.c_uint32), ("completed", C.c_uint64),
                ("steps", C.c_uint64)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double)]


class Ray(C.Structure):
    _fields_ = [("origin", F3), ("direction", F3), ("tmin", C.c_float), ("tmax", C.c_float),
                ("origin_low", F3), ("ignore", C.c_uint32)]


class Hit(C.Structure):
    _fields_ = [("hit", C.c_uint32), ("triangle", C.c_uint32), ("t", C.c_double), ("u", C.c_double),
                ("v", C.c_double)]


class SceneInfo(C.Structure):
    _fields_ = [("triangles", C.c_uint64), ("bvh_nodes", C.c_uint64), ("wide_nodes", C.c_uint64),
                ("device_bytes", C.c_uint64), ("build_seconds", C.c_double), ("upload_seconds", C.c_double)]


RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("direction", "<f4", 3), ("tmin", "<f4"), ("tmax", "<f4"),
                      ("origin_low", "<f4", 3), ("ignore", "<u4")])
HIT_DTYPE = np.dtype([("hit", "<u4"), ("triangle", "<u4"), ("t", "<f8"), ("u", "<f8"), ("v", "<f8")])
assert RAY_DTYPE.itemsize == C.sizeof(Ray) == 48
assert HIT_DTYPE.itemsize == C.sizeof(Hit) == 32


def make_rays(origins, directions, tmin=0.0, tmax=INF, origin_low=None, ignore=NO_TRIANGLE) -> np.ndarray:
    origins = np.asarray(origins, np.float32).reshape(-1, 3)
    rays = np.zeros(len(origins), RAY_DTYPE)
    rays["origin"] = origins
    rays["direction"] = np.asarray(directions, np.float32).reshape(-1, 3)
    rays["tmin"] = tmin
    rays["tmax"] = tmax
    if origin_low is not None:
        rays["origin_low"] = np.asarray(origin_low, np.float32).reshape(-1, 3)
    rays["ignore"] = ignore
    return rays


def _fptr(a: Optional[np.ndarray]):
    if a is None:
        return C.POINTER(C.c_float)()
    return a.ctypes.data_as(C.POINTER(C.c_float))


@dataclass
class SceneData:
    """Triangle soup + materials + punctual lights, before Scene::finalize."""
    positions: np.ndarray                     # (n, 9) float32
    material_ids: np.ndarray                  # (n,) uint32
    materials: List[Material]
    normals: Optional[np.ndarray] = None      # (n, 9) float32
    tangents: Optional[np.ndarray] = None     # (n, 12) float32
    uvs: Optional[np.ndarray] = None          # (n, 12) float32
    lights: List[Light] = field(default_factory=list)
    images: List[np.ndarray] = field(default_factory=list)   # (h, w, 4) float32 each
    textures: List[Texture] = field(default_factory=list)
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, np.float32).reshape(-1, 9)
        self.material_ids = np.ascontiguousarray(self.material_ids, np.uint32).reshape(-1)
        for name in ("normals", "tangents", "uvs"):
            a = getattr(self, name)
            if a is not None:
                setattr(self, name, np.ascontiguousarray(a, np.float32).reshape(len(self.positions), -1))

    @property
    def triangle_count(self) -> int:
        return len(self.positions)

    def descriptor(self) -> SceneDesc:
        mats = (Material * max(1, len(self.materials)))(*self.materials)
        lights = (Light * max(1, len(self.lights)))(*self.lights)
        imgs = [np.ascontiguousarray(i, np.float32) for i in self.images]
        images = (Image * max(1, len(imgs)))(*[Image(i.shape[1], i.shape[0], _fptr(i)) for i in imgs])
        texs = (Texture * max(1, len(self.textures)))(*self.textures)
        self._keep = [mats, lights, images, texs, imgs]
        return SceneDesc(self.triangle_count, _fptr(self.positions), _fptr(self.normals), _fptr(self.tangents),
                         _fptr(self.uvs), self.material_ids.ctypes.data_as(C.POINTER(C.c_uint32)),
                         len(self.materials), mats, len(self.lights), lights, len(imgs), images,
                         len(self.textures), texs)

    def bytes_sha256(self) -> str:
        import hashlib
        h = hashlib.sha256()
        for a in (self.positions, self.material_ids, self.normals):
            if a is not None:
                h.update(np.ascontiguousarray(a).tobytes())
        for m in self.materials:
            h.update(bytes(m))
        return h.hexdigest()


@dataclass
class EnvData:
    constant: tuple = (0.0, 0.0, 0.0)
    rgba: Optional[np.ndarray] = None          # (h, w, 4) float32
    rotation_degrees: float = 0.0
    intensity: float = 1.0

    def descriptor(self) -> EnvDesc:
        if self.rgba is None:
            return EnvDesc(F3(*self.constant), 0, 0, C.POINTER(C.c_float)(), self.rotation_degrees, self.intensity)
        self.rgba = np.ascontiguousarray(self.rgba, np.float32)
        h, w = self.rgba.shape[:2]
        return EnvDesc(F3(*self.constant), w, h, _fptr(self.rgba), self.rotation_degrees, self.intensity)
