"""TEST INFRASTRUCTURE ONLY: ctypes binding of the probe interface (probe.h).

`Probe("orc")` loads oracle/liboracle.so, the CPU restatement of the
reference hot path; `Probe("ref")` loads oracle/_ref/libcrref.so, the
reference's own sources compiled from where they lie (present only where the
reference tree was available at build time).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg use this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent


def _abi():
    if "cr_amd" not in sys.modules:
        pkg = ROOT / "cpu-renderer_amd"
        spec = importlib.util.spec_from_file_location("cr_amd", pkg / "__init__.py",
                                                      submodule_search_locations=[str(pkg)])
        mod = importlib.util.module_from_spec(spec)
        sys.modules["cr_amd"] = mod
        spec.loader.exec_module(mod)
    from cr_amd import abi
    return abi


abi = _abi()

PATHS = {"orc": HERE / "liboracle.so", "ref": HERE / "_ref" / "libcrref.so"}


class Surface(C.Structure):
    _fields_ = [(n, C.c_float * 3) for n in ("position", "position_low", "geometric_normal", "normal", "tangent",
                                             "bitangent", "base_color", "emission", "attenuation_color")] + \
               [(n, C.c_float) for n in ("metallic", "roughness", "transmission", "ior", "thickness",
                                         "attenuation_distance")] + \
               [("material", C.c_uint32), ("front_face", C.c_int32), ("double_sided", C.c_int32)]


class BsdfSample(C.Structure):
    _fields_ = [("direction", C.c_float * 3), ("weight", C.c_float * 3), ("pdf", C.c_float), ("eta", C.c_float),
                ("delta", C.c_int32), ("transmitted", C.c_int32), ("valid", C.c_int32)]


def available(kind: str) -> bool:
    return PATHS[kind].exists()


class Probe:
    def __init__(self, kind: str = "orc"):
        if kind not in PATHS:
            raise ValueError(kind)
        if not PATHS[kind].exists():
            raise FileNotFoundError(f"{PATHS[kind]} is not built")
        self.kind = kind
        self.lib = C.CDLL(str(PATHS[kind]))
        p = kind + "_"
        vp, u64, f = C.c_void_p, C.c_uint64, C.c_float
        fp = C.POINTER(C.c_float)
        sig = {
            "last_error": (C.c_char_p, []),
            "sample_seed": (u64, [u64, u64, u64]),
            "rng_draw": (None, [C.POINTER(u64), C.c_int, fp]),
            "camera_generate": (C.c_int, [C.POINTER(abi.Camera), f, f, C.c_int, C.c_int, C.POINTER(abi.Ray)]),
            "scene_create": (vp, [C.POINTER(abi.SceneDesc)]),
            "scene_destroy": (None, [vp]),
            "scene_triangles": (u64, [vp]),
            "scene_nodes": (u64, [vp]),
            "trace": (C.c_int, [vp, vp, u64, vp, C.c_int]),
            "offset_ray": (C.c_int, [fp, fp, fp, f, C.POINTER(abi.Ray)]),
            "surface_at": (C.c_int, [vp, C.POINTER(abi.Ray), C.POINTER(abi.Hit), C.POINTER(Surface)]),
            "bsdf_sample": (None, [C.POINTER(Surface), fp, C.POINTER(u64), C.POINTER(BsdfSample)]),
            "bsdf_eval": (None, [C.POINTER(Surface), fp, fp, fp]),
            "env_create": (vp, [C.POINTER(abi.EnvDesc)]),
            "env_destroy": (None, [vp]),
            "env_sample": (None, [vp, C.POINTER(u64), fp, fp]),
            "env_pdf": (f, [vp, fp]),
            "env_radiance": (None, [vp, fp, fp]),
            "render": (C.c_int, [vp, vp, C.POINTER(abi.Camera), C.POINTER(abi.Options), fp, C.POINTER(abi.Stats)]),
            "film_png8": (C.c_int, [fp, C.c_int, C.c_int, f, C.POINTER(C.c_uint8)]),
            "film_hdr": (C.c_int, [fp, C.c_int, C.c_int, C.POINTER(C.c_uint8), u64, C.POINTER(u64)]),
        }
        if kind == "orc":
            sig["render_window"] = (C.c_int, [vp, vp, C.POINTER(abi.Camera), C.POINTER(abi.Options), C.c_int, C.c_int,
                                              C.c_int, C.c_int, fp, fp, C.POINTER(abi.Stats)])
        if kind == "ref":
            sig["film_write"] = (C.c_int, [fp, C.c_int, C.c_int, f, C.c_char_p, C.c_char_p])
            sig["gltf_load"] = (vp, [C.c_char_p])
            sig["gltf_destroy"] = (None, [vp])
            sig["gltf_desc"] = (C.POINTER(abi.SceneDesc), [vp])
            sig["gltf_camera_count"] = (C.c_int, [vp])
            sig["gltf_camera"] = (C.c_int, [vp, C.c_int, C.POINTER(abi.Camera)])
            sig["gltf_render_loaded"] = (C.c_int, [vp, vp, C.POINTER(abi.Camera), C.POINTER(abi.Options), fp,
                                                   C.POINTER(abi.Stats)])
        for name, (res, args) in sig.items():
            fn = getattr(self.lib, p + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, "_" + name, fn)

    def error(self) -> str:
        return self._last_error().decode()

    # --- stateless pieces
    def sample_seed(self, seed, pixel, sample) -> int:
        return self._sample_seed(seed, pixel, sample)

    def rng_draw(self, state: int, count: int):
        st = C.c_uint64(state)
        out = np.zeros(count, np.float32)
        self._rng_draw(C.byref(st), count, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out, st.value

    def camera_generate(self, camera, px, py, width, height):
        r = abi.Ray()
        rc = self._camera_generate(C.byref(camera), px, py, width, height, C.byref(r))
        return (r if rc == 0 else None), (self.error() if rc else "")

    def offset_ray(self, position, normal, direction, max_distance=float("inf")):
        r = abi.Ray()
        arr = lambda v: (C.c_float * 3)(*v)
        rc = self._offset_ray(arr(position), arr(normal), arr(direction), max_distance, C.byref(r))
        return (r if rc == 0 else None), (self.error() if rc else "")

    def bsdf_sample(self, surface: Surface, wo, state: int):
        st = C.c_uint64(state)
        out = BsdfSample()
        self._bsdf_sample(C.byref(surface), (C.c_float * 3)(*wo), C.byref(st), C.byref(out))
        return out, st.value

    def bsdf_eval(self, surface: Surface, wo, wi):
        out = (C.c_float * 4)()
        self._bsdf_eval(C.byref(surface), (C.c_float * 3)(*wo), (C.c_float * 3)(*wi), out)
        return np.array(out[:3], np.float32), np.float32(out[3])

    # --- film encode (writePng / writeHdr)
    def film_png8(self, film, exposure: float):
        """Display bytes (H, W, 3) the PNG writer stores, or raises with its message."""
        film = np.ascontiguousarray(film, np.float32)
        h, w = film.shape[:2]
        out = np.zeros((h, w, 3), np.uint8)
        if self._film_png8(film.ctypes.data_as(C.POINTER(C.c_float)), w, h, exposure,
                           out.ctypes.data_as(C.POINTER(C.c_uint8))):
            raise RuntimeError(self.error())
        return out

    def film_write(self, film, exposure: float, png_path, hdr_path) -> None:
        """Reference only: writePng and writeHdr to files."""
        film = np.ascontiguousarray(film, np.float32)
        h, w = film.shape[:2]
        if self._film_write(film.ctypes.data_as(C.POINTER(C.c_float)), w, h, exposure, str(png_path).encode(),
                            str(hdr_path).encode()):
            raise RuntimeError(self.error())

    def film_hdr(self, film) -> bytes:
        """The complete Radiance .hdr file the HDR writer produces."""
        film = np.ascontiguousarray(film, np.float32)
        h, w = film.shape[:2]
        cap = 256 + h * (4 + 4 * (w + w // 128 + 2))
        out = np.zeros(cap, np.uint8)
        size = C.c_uint64()
        if self._film_hdr(film.ctypes.data_as(C.POINTER(C.c_float)), w, h, out.ctypes.data_as(C.POINTER(C.c_uint8)),
                          cap, C.byref(size)):
            raise RuntimeError(self.error())
        return bytes(out[:size.value])

    # --- scene / environment objects
    def scene(self, data) -> "ProbeScene":
        return ProbeScene(self, data)

    def env(self, data) -> "ProbeEnv":
        return ProbeEnv(self, data)

    def gltf(self, path) -> "ProbeGltf":
        return ProbeGltf(self, path)


class ProbeScene:
    def __init__(self, probe: Probe, data):
        self.p = probe
        self.data = data
        desc = data.descriptor()
        self.h = probe._scene_create(C.byref(desc))
        if not self.h:
            raise RuntimeError(probe.error())

    def __del__(self):
        if getattr(self, "h", None):
            self.p._scene_destroy(self.h)
            self.h = None

    @property
    def triangles(self) -> int:
        return self.p._scene_triangles(self.h)

    @property
    def nodes(self) -> int:
        return self.p._scene_nodes(self.h)

    def trace(self, rays: np.ndarray, any_hit=False) -> np.ndarray:
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        hits = np.zeros(len(rays), abi.HIT_DTYPE)
        rc = self.p._trace(self.h, rays.ctypes.data, len(rays), hits.ctypes.data, int(bool(any_hit)))
        if rc:
            raise RuntimeError(self.p.error())
        return hits

    def surface_at(self, ray, hit):
        s = Surface()
        rc = self.p._surface_at(self.h, C.byref(ray), C.byref(hit), C.byref(s))
        if rc:
            raise RuntimeError(self.p.error())
        return s

    def render_window(self, env: "ProbeEnv", camera, options, window, with_samples=False):
        """Oracle only: the pixels of window (x0, y0, w, h) of the frame, each with its
        full-frame sample stream; also every sample's radiance ([spp, h, w, 3]) on request."""
        x0, y0, w, h = window
        film = np.zeros((h, w, 3), np.float32)
        smp = np.zeros((options.samples, h, w, 3), np.float32) if with_samples else None
        stats = abi.Stats()
        rc = self.p._render_window(self.h, env.h, C.byref(camera), C.byref(options), x0, y0, w, h,
                                   film.ctypes.data_as(C.POINTER(C.c_float)),
                                   smp.ctypes.data_as(C.POINTER(C.c_float)) if smp is not None else None,
                                   C.byref(stats))
        if rc:
            raise RuntimeError(self.p.error())
        return (film, stats, smp) if with_samples else (film, stats)

    def render(self, env: "ProbeEnv", camera, options):
        film = np.zeros((options.height, options.width, 3), np.float32)
        stats = abi.Stats()
        rc = self.p._render(self.h, env.h, C.byref(camera), C.byref(options),
                            film.ctypes.data_as(C.POINTER(C.c_float)), C.byref(stats))
        if rc:
            raise RuntimeError(self.p.error())
        return film, stats


class ProbeEnv:
    def __init__(self, probe: Probe, data):
        self.p = probe
        self.data = data
        desc = data.descriptor()
        self.h = probe._env_create(C.byref(desc))
        if not self.h:
            raise RuntimeError(probe.error())

    def __del__(self):
        if getattr(self, "h", None):
            self.p._env_destroy(self.h)
            self.h = None

    def sample(self, state: int):
        st = C.c_uint64(state)
        d = (C.c_float * 3)()
        pdf = C.c_float()
        self.p._env_sample(self.h, C.byref(st), d, C.byref(pdf))
        return np.array(d[:], np.float32), np.float32(pdf.value), st.value

    def pdf(self, direction) -> np.float32:
        return np.float32(self.p._env_pdf(self.h, (C.c_float * 3)(*direction)))

    def radiance(self, direction):
        out = (C.c_float * 3)()
        self.p._env_radiance(self.h, (C.c_float * 3)(*direction), out)
        return np.array(out[:], np.float32)


def _copy_array(ptr, count, dtype, shape):
    if not ptr or count == 0:
        return None
    return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True).reshape(shape)


class ProbeGltf:
    """A glTF/GLB file loaded by the reference's io::loadGltf (reference build only)."""

    def __init__(self, probe: Probe, path):
        if probe.kind != "ref":
            raise ValueError("glTF import is a reference-build probe")
        self.p = probe
        self.h = probe._gltf_load(str(path).encode())
        if not self.h:
            raise RuntimeError(probe.error())

    def __del__(self):
        if getattr(self, "h", None):
            self.p._gltf_destroy(self.h)
            self.h = None

    def scene_data(self):
        """The loaded scene as SceneData (arrays copied out of the descriptor)."""
        d = self.p._gltf_desc(self.h).contents
        n = int(d.triangle_count)
        images = []
        for i in range(d.image_count):
            im = d.images[i]
            images.append(_copy_array(im.rgba, 4 * im.width * im.height, np.float32, (im.height, im.width, 4)))
        return abi.SceneData(
            positions=_copy_array(d.positions, 9 * n, np.float32, (n, 9)),
            material_ids=_copy_array(d.material_ids, n, np.uint32, (n,)),
            materials=[abi.Material.from_buffer_copy(bytes(d.materials[i])) for i in range(d.material_count)],
            normals=_copy_array(d.normals, 9 * n, np.float32, (n, 9)),
            tangents=_copy_array(d.tangents, 12 * n, np.float32, (n, 12)),
            uvs=_copy_array(d.uvs, 12 * n, np.float32, (n, 12)),
            lights=[abi.Light.from_buffer_copy(bytes(d.lights[i])) for i in range(d.light_count)],
            images=images,
            textures=[abi.Texture.from_buffer_copy(bytes(d.textures[i])) for i in range(d.texture_count)])

    def cameras(self):
        out = []
        for i in range(self.p._gltf_camera_count(self.h)):
            c = abi.Camera()
            if self.p._gltf_camera(self.h, i, C.byref(c)):
                raise RuntimeError(self.p.error())
            out.append(c)
        return out

    def render_loaded(self, env: "ProbeEnv", camera, options):
        """cr::render of the loaded cr::Scene itself (finalize + build on a copy)."""
        film = np.zeros((options.height, options.width, 3), np.float32)
        stats = abi.Stats()
        rc = self.p._gltf_render_loaded(self.h, env.h, C.byref(camera), C.byref(options),
                                        film.ctypes.data_as(C.POINTER(C.c_float)), C.byref(stats))
        if rc:
            raise RuntimeError(self.p.error())
        return film, stats
