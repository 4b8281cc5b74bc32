"""In-memory test scenes modelled on the reference test suites' factories.

These are inputs only: quads/cubes/slabs built the way the reference tests
build them (reference tests/test_integrator.cpp:222-242, tests/test_transport_edges.cpp:25-65,
tests/test_geometry.cpp:42-62, :533-658), so parity tests exercise the same
transport regimes (closed media, touching interfaces, punctual lights,
adversarial BVH inputs).  Float arithmetic is done in binary32 with numpy
scalars in the same operation order, so the scenes are bit-reproducible.
"""
from __future__ import annotations

import numpy as np

from conftest import abi

f32 = np.float32
INF = float("inf")


def _v(x):
    return np.array(x, np.float32)


def normalize(v):
    v = _v(v)
    l = np.sqrt(f32(v[0] * v[0] + v[1] * v[1]) + f32(v[2] * v[2]), dtype=np.float32)
    inv = f32(f32(1) / l)
    return v * inv


def cross(a, b):
    a, b = _v(a), _v(b)
    return _v([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


class Builder:
    def __init__(self):
        self.pos, self.nrm, self.mid = [], [], []
        self.materials, self.lights = [], []

    def material(self, **kw) -> int:
        self.materials.append(abi.material(**kw))
        return len(self.materials) - 1

    def tri(self, a, b, c, m, normals=None):
        self.pos.append(np.concatenate([_v(a), _v(b), _v(c)]))
        self.nrm.append(np.zeros(9, np.float32) if normals is None else np.concatenate([_v(n) for n in normals]))
        self.mid.append(m)

    def quad(self, a, b, c, d, m=0):
        n = normalize(cross(_v(b) - _v(a), _v(c) - _v(a)))
        self.tri(a, b, c, m, (n, n, n))
        self.tri(a, c, d, m, (n, n, n))

    def cube(self, m, h=0.7):
        q = self.quad
        q((-h, -h, h), (h, -h, h), (h, h, h), (-h, h, h), m)
        q((h, -h, -h), (-h, -h, -h), (-h, h, -h), (h, h, -h), m)
        q((-h, -h, -h), (-h, -h, h), (-h, h, h), (-h, h, -h), m)
        q((h, -h, h), (h, -h, -h), (h, h, -h), (h, h, h), m)
        q((-h, h, h), (h, h, h), (h, h, -h), (-h, h, -h), m)
        q((-h, -h, -h), (h, -h, -h), (h, -h, h), (-h, -h, h), m)

    def scene(self) -> abi.SceneData:
        pos = np.array(self.pos, np.float32).reshape(-1, 9) if self.pos else np.zeros((0, 9), np.float32)
        nrm = np.array(self.nrm, np.float32).reshape(-1, 9) if self.nrm else None
        return abi.SceneData(pos, np.array(self.mid, np.uint32), list(self.materials), normals=nrm,
                             lights=list(self.lights))


def test_camera(position=(0, 0, 4), forward=(0, 0, -1), yfov=0.65, right=(1, 0, 0), up=(0, 1, 0)):
    return abi.camera(position=position, forward=forward, right=right, up=up, yfov=yfov)


def integrator_fixture(kind: int):
    """The diffuse / metal / water / lit-room feature scenes (kind 0,1,2,4)."""
    b = Builder()
    b.material(metallic=0, base_color=(0.4, 0.35, 0.25, 1))
    obj = dict(metallic=0, base_color=(0.55, 0.16, 0.08, 1), roughness=0.4)
    if kind == 1:
        obj = dict(metallic=1, base_color=(0.9, 0.55, 0.16, 1), roughness=0.2)
    if kind == 2:
        obj = dict(metallic=0, transmission=1, thickness=1, base_color=(1, 1, 1, 1),
                   attenuation_color=(0.25, 0.85, 0.7), attenuation_distance=2, ior=1.333, roughness=0.14)
    b.material(**obj)
    b.quad((-8, -0.8, 8), (8, -0.8, 8), (8, -0.8, -8), (-8, -0.8, -8), 0)
    b.cube(1)
    if kind == 4:
        b.quad((-2, -0.8, -2), (2, -0.8, -2), (2, 3, -2), (-2, 3, -2), 0)
        b.material(emission=(5, 5, 5), metallic=0, base_color=(0, 0, 0, 1))
        b.quad((-0.8, 2, -0.8), (0.8, 2, -0.8), (0.8, 2, 0.8), (-0.8, 2, 0.8), 2)
    else:
        b.lights.append(abi.light(type=abi.DIRECTIONAL, direction=tuple(normalize((-1, -2, -1))), intensity=2))
    env = abi.EnvData(constant=(0.01,) * 3 if kind == 4 else (0.3,) * 3)
    return b.scene(), env, test_camera()


def slab(builder=None, h=100.0):
    b = builder or Builder()
    b.material(metallic=0, roughness=0.6, transmission=1, ior=1.5, thickness=1, attenuation_distance=1,
               attenuation_color=(0.5, 0.8, 1))
    q = b.quad
    q((-h, -h, 1), (h, -h, 1), (h, h, 1), (-h, h, 1))
    q((h, -h, 0), (-h, -h, 0), (-h, h, 0), (h, h, 0))
    q((-h, -h, 0), (-h, -h, 1), (-h, h, 1), (-h, h, 0))
    q((h, -h, 1), (h, -h, 0), (h, h, 0), (h, h, 1))
    q((-h, h, 1), (h, h, 1), (h, h, 0), (-h, h, 0))
    q((-h, -h, 0), (h, -h, 0), (h, -h, 1), (-h, -h, 1))
    return b


def pixel_camera():
    return abi.camera(position=(0, 0, 2), yfov=1e-6)


def transport_internal_point_light(clear=False):
    b = slab()
    if clear:
        b.materials[0].attenuation_color[:] = [1, 1, 1]
    b.lights.append(abi.light(position=(0, 0, 0.5), intensity=0.01))
    return b.scene(), abi.EnvData(), pixel_camera()


def transport_exterior_directional(clear=False):
    b = slab()
    b.materials.append(abi.material(metallic=0, roughness=0.6, transmission=1, ior=1.5, thickness=1,
                                    attenuation_distance=1, attenuation_color=(0.5, 0.8, 1)))
    b.materials[0].roughness = 0
    b.mid[2] = b.mid[3] = 1
    if clear:
        for m in b.materials:
            m.attenuation_color[:] = [1, 1, 1]
    b.lights.append(abi.light(type=abi.DIRECTIONAL, direction=(0, 0, 1), intensity=0.01))
    return b.scene(), abi.EnvData(), pixel_camera()


def transport_enclosed_receiver():
    b = slab()
    b.materials[0].ior = 1
    b.materials[0].attenuation_color[:] = [1, 1, 1]
    b.material(metallic=0, ior=1, base_color=(0.4, 0.4, 0.4, 1))
    b.quad((-99, -99, 0.5), (99, -99, 0.5), (99, 99, 0.5), (-99, 99, 0.5), 1)
    return b.scene(), abi.EnvData(constant=(1, 1, 1)), pixel_camera()


def touching_boundary():
    """Opaque face micrometres above a closed liquid: exercises precision replay."""
    b = Builder()
    b.material(metallic=0, transmission=1, ior=1, thickness=1, attenuation_distance=1, attenuation_color=(0.5, 0.8, 1))
    b.material(metallic=0, ior=1, base_color=(0.6, 0.6, 0.6, 1))
    q = b.quad
    q((0, 0, 20), (43, 0, 20), (43, 0, -20), (0, 0, -20))
    q((0, -2, -20), (43, -2, -20), (43, -2, 20), (0, -2, 20))
    q((0, -2, 20), (43, -2, 20), (43, 0, 20), (0, 0, 20))
    q((43, -2, -20), (0, -2, -20), (0, 0, -20), (43, 0, -20))
    q((0, -2, -20), (0, -2, 20), (0, 0, 20), (0, 0, -20))
    q((43, -2, 20), (43, -2, -20), (43, 0, -20), (43, 0, 20))
    q((-1, -1, 21), (44, -1, 21), (44, -1, -21), (-1, -1, -21), 1)
    e = f32(2e-6)
    b.tri((19, f32(-1) + e, -1), (21, f32(1) + e, -1), (20, e, 1), 1)
    fwd = normalize(_v((20, e, 0)) - _v((22, 1, 0)))
    right = normalize(cross(fwd, (0, 1, 0)))
    up = cross(right, fwd)
    cam = abi.camera(position=(22, 1, 0), forward=tuple(fwd), right=tuple(right), up=tuple(up), yfov=1e-7)
    return b.scene(), abi.EnvData(constant=(1, 1, 1)), cam


def energy_plane():
    b = Builder()
    b.material(metallic=0, base_color=(0.6, 0.6, 0.6, 1), ior=1)
    b.quad((-20, -20, 0), (20, -20, 0), (20, 20, 0), (-20, 20, 0), 0)
    return b.scene(), abi.EnvData(constant=(1, 1, 1)), test_camera()


def bvh_random(seed=0x75645321, count=450):
    """Random oriented triangles with mixed materials (cf. reference testBvh)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = Builder()
    b.material(double_sided=True)
    b.material(double_sided=True)
    b.material()
    b.material(transmission=0.8)
    for i in range(count):
        c = rng.uniform(-10, 10, 3)
        n = rng.uniform(-1, 1, 3)
        n = n / np.sqrt((n * n).sum())
        helper = np.array([0, 0, 1.0]) if abs(n[2]) < 0.999 else np.array([0, 1.0, 0])
        x = np.cross(helper, n)
        x = x / np.sqrt((x * x).sum())
        y = np.cross(n, x)
        s = rng.uniform(0.1, 2)
        b.tri(c - x * s - y * s, c + x * s, c + y * s, i % 4)
    return b.scene()


def bvh_ties(shift: int):
    b = Builder()
    b.material(double_sided=True)
    x = float(shift)
    for _ in range(2):
        b.tri((x, 0, 0), (x + 1, 0, 0), (x, 1, 0), 0)
    for i in range(60):
        ox = float((i * 23) % 61 - 30) * 2
        b.tri((ox, 0, -2), (ox + 1, 0, -2), (ox, 1, -2), 0)
    return b.scene()


def bvh_exponential():
    b = Builder()
    b.material(double_sided=True)
    for e in range(-80, 80):
        x = float(np.ldexp(np.float32(1), e))
        b.tri((x, 0, 0), (x, 1, 0), (x, 0, 1), 0)
    return b.scene()


def bvh_boundary(scale: float):
    b = Builder()
    b.material(double_sided=True)
    s = f32(scale)
    for i in range(20):
        x = f32(f32(i) * f32(5)) * s
        b.tri((x, 0, 0), (x + s, 0, 0), (x, s, 0), 0)
    return b.scene()


def bvh_identical():
    b = Builder()
    b.material(double_sided=True)
    for _ in range(100):
        b.tri((0, 0, 0), (1, 0, 0), (0, 1, 0), 0)
    return b.scene()


def bvh_close_surfaces():
    b = Builder()
    b.material(double_sided=True)
    e = f32(1e-7)
    b.tri((19, e, -1), (21, e, -1), (20, e, 1), 0)
    b.tri((18, 0, -2), (20, 0, 2), (22, 0, -2), 0)
    return b.scene()


def bvh_ordered():
    b = Builder()
    b.material()
    b.tri((0, 0, 0), (1, 0, 0), (0, 1, 0), 0)
    e = f32(1e-6)
    b.tri((0, 0, e), (1, 0, e), (0, 1, e), 0)
    return b.scene()
