"""TEST INFRASTRUCTURE: build and bind the CPU emulation of the device path
(tests/csrc/emu.cpp + the device headers and host preparation sources)."""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import subprocess

import numpy as np

from conftest import ROOT, abi

SRC = [ROOT / "tests" / "csrc" / "emu.cpp", ROOT / "cpu-renderer_amd" / "csrc" / "host_scene.cpp",
       ROOT / "cpu-renderer_amd" / "csrc" / "film_host.cpp"]
DEPS = SRC + sorted((ROOT / "cpu-renderer_amd" / "csrc").glob("*.h")) + [ROOT / "include" / "crgpu.h"]
OUT = ROOT / "tests" / "csrc" / "_build" / "libemu.so"


def build() -> pathlib.Path:
    if OUT.exists() and OUT.stat().st_mtime >= max(p.stat().st_mtime for p in DEPS):
        return OUT
    OUT.parent.mkdir(exist_ok=True)
    tmp = OUT.with_suffix(f".so.tmp{os.getpid()}")  # concurrent test workers each build their own
    subprocess.run(["clang++", "-std=c++20", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                    "-pthread", "-o", str(tmp), *map(str, SRC), "-lz"], check=True)
    tmp.replace(OUT)
    return OUT


class Emu:
    def __init__(self):
        self.lib = C.CDLL(str(build()))
        vp, u64, fp = C.c_void_p, C.c_uint64, C.POINTER(C.c_float)
        sig = {
            "emu_last_error": (C.c_char_p, []),
            "emu_scene_create": (vp, [C.POINTER(abi.SceneDesc)]),
            "emu_scene_destroy": (None, [vp]),
            "emu_scene_triangles": (u64, [vp]),
            "emu_scene_wide_nodes": (u64, [vp]),
            "emu_scene_check": (C.c_int, [vp]),
            "emu_env_create": (vp, [C.POINTER(abi.EnvDesc)]),
            "emu_env_destroy": (None, [vp]),
            "emu_env_verify_inversion": (u64, [vp]),
            "emu_env_sample": (None, [vp, C.POINTER(u64), fp, fp]),
            "emu_env_pdf": (C.c_float, [vp, fp]),
            "emu_env_radiance": (None, [vp, fp, fp]),
            "emu_trace": (C.c_int, [vp, vp, u64, vp, C.c_int, C.POINTER(u64), u64]),
            "emu_box_hit": (C.c_int, [fp, fp, C.c_float, C.c_float, fp, fp, fp, fp]),
            "emu_expf": (None, [fp, fp, u64]),
            "emu_pow5": (None, [fp, fp, fp, u64]),
            "emu_mis": (None, [fp, fp, fp, fp, u64]),
            "emu_dunit": (None, [C.POINTER(C.c_uint64), fp, fp, u64]),
            "emu_offset": (None, [fp, fp, fp, fp, fp, u64]),
            "emu_interp_unit": (None, [fp, C.POINTER(C.c_uint64), fp, fp, u64]),
            "emu_interp_point": (None, [fp, C.POINTER(C.c_uint64), fp, fp, C.POINTER(C.c_uint8), u64]),
            "emu_hit_surface": (C.c_int, [vp, vp, u64, fp, fp, C.POINTER(C.c_uint8)]),
            "emu_camera_pair": (C.c_int, [C.POINTER(abi.Camera), fp, fp, u64, C.c_int, C.c_int, fp, fp]),
            "emu_mask_check": (C.c_int, [vp, vp, u64, C.POINTER(u64)]),
            "emu_film_png8": (C.c_int, [fp, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_uint8)]),
            "emu_film_hdr": (C.c_int, [fp, C.c_int, C.c_int, C.POINTER(C.c_uint8), u64, C.POINTER(u64)]),
            "emu_camera_generate": (C.c_int, [C.POINTER(abi.Camera), C.c_float, C.c_float, C.c_int, C.c_int,
                                              C.POINTER(abi.Ray)]),
            "emu_render": (C.c_int, [vp, vp, C.POINTER(abi.Camera), C.POINTER(abi.Options), fp,
                                     C.POINTER(abi.Stats)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(self.lib, name)
            f.restype, f.argtypes = res, args

    def error(self):
        return self.lib.emu_last_error().decode()

    def film_png8(self, film, exposure):
        film = np.ascontiguousarray(film, np.float32)
        h, w = film.shape[:2]
        out = np.zeros((h, w, 3), np.uint8)
        if self.lib.emu_film_png8(film.ctypes.data_as(C.POINTER(C.c_float)), w, h, exposure,
                                  out.ctypes.data_as(C.POINTER(C.c_uint8))):
            raise RuntimeError(self.error())
        return out

    def film_hdr(self, film) -> bytes:
        film = np.ascontiguousarray(film, np.float32)
        h, w = film.shape[:2]
        cap = 256 + h * (4 + 4 * (w + w // 128 + 2))
        out = np.zeros(cap, np.uint8)
        size = C.c_uint64()
        if self.lib.emu_film_hdr(film.ctypes.data_as(C.POINTER(C.c_float)), w, h,
                                 out.ctypes.data_as(C.POINTER(C.c_uint8)), cap, C.byref(size)):
            raise RuntimeError(self.error())
        return bytes(out[:size.value])

    def scene(self, data):
        return EmuScene(self, data)

    def env(self, data):
        return EmuEnv(self, data)


class EmuScene:
    def __init__(self, emu, data):
        self.e, self.data = emu, data
        desc = data.descriptor()
        self.h = emu.lib.emu_scene_create(C.byref(desc))
        if not self.h:
            raise RuntimeError(emu.error())

    def __del__(self):
        if getattr(self, "h", None):
            self.e.lib.emu_scene_destroy(self.h)

    def check(self) -> int:
        return self.e.lib.emu_scene_check(self.h)

    @property
    def triangles(self):
        return self.e.lib.emu_scene_triangles(self.h)

    def trace(self, rays, any_hit=False, schedule=0):
        """schedule 0: one node or leaf per step; k > 0: postponed-leaf phases in a
        pseudo-random interleaving seeded by k (FP32 pass only); bit 2 set: the traversal
        variant without FP32 alpha-mask tests."""
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        hits = np.zeros(len(rays), abi.HIT_DTYPE)
        fb = C.c_uint64()
        if self.e.lib.emu_trace(self.h, rays.ctypes.data, len(rays), hits.ctypes.data, int(any_hit), C.byref(fb),
                                schedule):
            raise RuntimeError(self.e.error())
        self.fallbacks = fb.value
        return hits

    def hit_surface(self, rays):
        """(ff, chain, decided): each ray's closest-hit position, shading normal and float(t)
        from float-pair barycentrics (decided bit 0: surface, bit 1: t) or the binary64
        chain (4 = miss)."""
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        n = len(rays)
        ff = np.zeros((n, 7), np.float32)
        chain = np.zeros((n, 7), np.float32)
        decided = np.zeros(n, np.uint8)
        fp = C.POINTER(C.c_float)
        if self.e.lib.emu_hit_surface(self.h, rays.ctypes.data, n, ff.ctypes.data_as(fp), chain.ctypes.data_as(fp),
                                      decided.ctypes.data_as(C.POINTER(C.c_uint8))):
            raise RuntimeError(self.e.error())
        return ff, chain, decided

    def render(self, env, cam, opt):
        film = np.zeros((opt.height, opt.width, 3), np.float32)
        st = abi.Stats()
        if self.e.lib.emu_render(self.h, env.h, C.byref(cam), C.byref(opt), film.ctypes.data_as(C.POINTER(C.c_float)),
                                 C.byref(st)):
            raise RuntimeError(self.e.error())
        return film, st


class EmuEnv:
    def __init__(self, emu, data):
        self.e, self.data = emu, data
        desc = data.descriptor()
        self.h = emu.lib.emu_env_create(C.byref(desc))
        if not self.h:
            raise RuntimeError(emu.error())

    def __del__(self):
        if getattr(self, "h", None):
            self.e.lib.emu_env_destroy(self.h)

    def verify_inversion(self) -> int:
        return self.e.lib.emu_env_verify_inversion(self.h)

    def sample(self, state):
        st = C.c_uint64(state)
        d = (C.c_float * 3)()
        p = C.c_float()
        self.e.lib.emu_env_sample(self.h, C.byref(st), d, C.byref(p))
        return np.array(d[:], np.float32), np.float32(p.value), st.value

    def pdf(self, d):
        return np.float32(self.e.lib.emu_env_pdf(self.h, (C.c_float * 3)(*d)))

    def radiance(self, d):
        out = (C.c_float * 3)()
        self.e.lib.emu_env_radiance(self.h, (C.c_float * 3)(*d), out)
        return np.array(out[:], np.float32)
