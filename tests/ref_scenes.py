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


# ---------------------------------------------------------------- textured scenes
def _slot(texture, texcoord=0, offset=(0.0, 0.0), scale=(1.0, 1.0), rotation=0.0):
    return abi.TextureSlot(texture, texcoord, abi.F2(*offset), abi.F2(*scale), rotation)


def _image(h, w, seed, kind):
    """Deterministic test images (h, w, 4) float32 in [0, 1]."""
    g = np.random.Generator(np.random.PCG64(seed))
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    img = np.empty((h, w, 4), np.float32)
    if kind == "checker":
        c = ((x // 2 + y // 2) % 2).astype(np.float32)
        img[..., 0] = 0.15 + 0.8 * c
        img[..., 1] = 0.9 - 0.7 * c
        img[..., 2] = g.random((h, w), dtype=np.float32) * 0.5 + 0.25
        img[..., 3] = c  # alpha 0/1 checker (bilinear gives intermediate values)
    elif kind == "noise":
        img[:] = g.random((h, w, 4), dtype=np.float32)
    else:  # tangent-space bump normals, z > 0
        nx = 0.35 * np.sin(x * 0.9) * np.cos(y * 0.7)
        ny = 0.35 * np.cos(x * 0.5 + y * 0.3)
        nz = np.sqrt(np.maximum(0.0, 1.0 - nx * nx - ny * ny))
        img[..., 0] = nx * 0.5 + 0.5
        img[..., 1] = ny * 0.5 + 0.5
        img[..., 2] = nz * 0.5 + 0.5
        img[..., 3] = 1.0
    return img


def textured_factors():
    """Base colour (sRGB, bilinear, repeat, scaled/rotated/offset), metallic-roughness
    (nearest, clamp, UV set 1), emissive (mirror) and transmission (nearest, repeat)
    maps, reference src/geometry.cpp:238-275 and :401-435."""
    images = [_image(8, 8, 1, "checker"), _image(5, 7, 2, "noise"), _image(6, 4, 3, "noise")]
    textures = [abi.Texture(0, 0, 0, 1), abi.Texture(1, 1, 1, 0), abi.Texture(2, 2, 2, 1), abi.Texture(1, 0, 0, 0)]
    mats = []
    m = abi.material(base_color=(0.9, 0.85, 0.8), metallic=1.0, roughness=1.0)
    m.base_color_tex = _slot(0, 0, (0.1, -0.2), (1.5, 2.0), 0.3)
    m.metallic_roughness_tex = _slot(1, 1, (0.0, 0.0), (1.0, 1.0), 0.0)
    mats.append(m)
    m = abi.material(base_color=(0.2, 0.2, 0.2), metallic=0.0, roughness=0.6, emission=(2.0, 1.5, 1.0))
    m.emissive_tex = _slot(2, 0, (0.25, 0.0), (-1.3, 2.2), -0.7)
    mats.append(m)
    m = abi.material(base_color=(1, 1, 1), metallic=0.0, roughness=0.1, transmission=1.0, double_sided=True)
    m.transmission_tex = _slot(3, 0, (0.0, 0.5), (3.0, 3.0), 1.1)
    mats.append(m)
    pos, uvs, mid = [], [], []

    def quad(a, b, c, d, uv_a, uv_b, uv_c, uv_d, mat, uv1_scale=1.0):
        for tri, tuv in (((a, b, c), (uv_a, uv_b, uv_c)), ((a, c, d), (uv_a, uv_c, uv_d))):
            pos.append(np.concatenate([_v(p) for p in tri]))
            uvs.append(np.concatenate([[u[0], u[1], u[0] * uv1_scale + 0.3, u[1] * uv1_scale - 0.1] for u in tuv]))
            mid.append(mat)

    quad((-2, -1, -2), (2, -1, -2), (2, -1, 2), (-2, -1, 2), (-1.2, -0.5), (2.7, -0.5), (2.7, 3.1), (-1.2, 3.1), 0, 0.8)
    quad((-2, -1, -2), (-2, 2, -2), (2, 2, -2), (2, -1, -2), (0, 0), (0, 1.7), (2.4, 1.7), (2.4, 0), 1)
    quad((-0.8, -0.9, 0.4), (0.7, -0.9, 0.1), (0.7, 0.8, 0.1), (-0.8, 0.8, 0.4), (0, 0), (1, 0), (1, 1), (0, 1), 2)
    # A spot light with a finite range exercises the cone and range falloff of
    # punctual NEE (reference src/render.cpp:182-209).
    spot = abi.light(abi.SPOT, position=(0.6, 1.7, 1.6), direction=tuple(normalize((-0.2, -1.0, -0.6))),
                     color=(1.0, 0.95, 0.9), intensity=9.0, range=6.0, inner_cone=0.25, outer_cone=0.7)
    scene = abi.SceneData(np.array(pos, np.float32), np.array(mid, np.uint32), mats,
                          uvs=np.array(uvs, np.float32), images=images, textures=textures, lights=[spot])
    env = abi.EnvData(constant=(0.3, 0.35, 0.4))
    cam = test_camera(position=(0.2, 0.6, 3.6), forward=normalize((-0.05, -0.25, -1)), yfov=0.9,
                      up=normalize((0, 1, -0.25)))
    return scene, env, cam


def textured_normal_map():
    """Normal-mapped sphere (UV set 1 with a rotated, mirrored slot transform; half the
    vertices without tangents so finalize derives them from UVs), reference
    src/geometry.cpp:296-397 (tangents) and :436-457 (mapping)."""
    from conftest import scenes
    unit = scenes.icosphere(2)
    pos = (unit * 0.9).reshape(-1, 9).astype(np.float32)
    nrm = unit.reshape(-1, 9).astype(np.float32)
    n = len(pos)
    v = unit.reshape(-1, 3)
    uu = (np.arctan2(v[:, 2], v[:, 0]) / (2 * np.pi) + 0.5).astype(np.float32)
    vv = (np.arccos(np.clip(v[:, 1], -1, 1)) / np.pi).astype(np.float32)
    uv = np.stack([uu, vv, uu * 2.0 + 0.1, vv * 1.5], 1).reshape(n, 12)
    tan = np.zeros((n, 3, 4), np.float32)
    t3 = np.cross(np.array([0.0, 1.0, 0.0]), v).reshape(n, 3, 3)
    tan[:, :, :3] = t3
    tan[:, :, 3] = np.where(np.arange(n)[:, None] % 3 == 0, -1.0, 1.0)
    tan[::2] = 0.0  # every other triangle: no tangents (derived from UVs in finalize)
    images = [_image(16, 16, 4, "normal"), _image(4, 4, 5, "noise")]
    textures = [abi.Texture(0, 0, 2, 1), abi.Texture(1, 0, 0, 1)]
    m = abi.material(base_color=(0.8, 0.6, 0.4), metallic=0.0, roughness=0.45, normal_scale=1.7)
    m.normal_tex = _slot(0, 1, (0.2, 0.1), (-1.5, 2.0), 0.4)
    m.base_color_tex = _slot(1, 0)
    floor = abi.material(base_color=(0.5, 0.5, 0.5), metallic=0.0, roughness=1.0)
    fpos = np.array([[-3, -1, -3, 3, -1, -3, 3, -1, 3], [-3, -1, -3, 3, -1, 3, -3, -1, 3]], np.float32)
    scene = abi.SceneData(np.concatenate([pos, fpos]), np.concatenate([np.zeros(n), np.ones(2)]).astype(np.uint32),
                          [m, floor], normals=np.concatenate([nrm, np.zeros((2, 9), np.float32)]),
                          tangents=np.concatenate([tan.reshape(n, 12), np.zeros((2, 12), np.float32)]),
                          uvs=np.concatenate([uv, np.zeros((2, 12), np.float32)]), images=images, textures=textures,
                          lights=[abi.light(abi.POINT, position=(1.5, 2.0, 2.0), color=(1, 0.9, 0.8), intensity=12.0)])
    env = abi.EnvData(constant=(0.15, 0.2, 0.3))
    cam = test_camera(position=(0.0, 0.3, 3.2), forward=normalize((0, -0.1, -1)), yfov=0.75,
                      up=normalize((0, 1, -0.1)))
    return scene, env, cam


def textured_alpha_mask():
    """Alpha-masked quad (checker alpha, bilinear) in front of a lit wall: the mask applies
    to camera, bounce and shadow rays alike (reference src/geometry.cpp:279-292)."""
    images = [_image(8, 8, 6, "checker")]
    textures = [abi.Texture(0, 0, 0, 1)]
    leaf = abi.material(base_color=(0.3, 0.8, 0.3, 0.9), metallic=0.0, roughness=0.8, alpha_mode=1,
                        alpha_cutoff=0.45, double_sided=True)
    leaf.base_color_tex = _slot(0, 0, (0.05, 0.0), (2.0, 2.0), 0.0)
    wall = abi.material(base_color=(0.7, 0.7, 0.7), metallic=0.0, roughness=1.0)
    b = Builder()
    b.materials = [leaf, wall]
    pos, uvs, mid = [], [], []
    for tri, tuv in ((((-1, -1, 0), (1, -1, 0), (1, 1, 0)), ((0, 0), (1, 0), (1, 1))),
                     (((-1, -1, 0), (1, 1, 0), (-1, 1, 0)), ((0, 0), (1, 1), (0, 1)))):
        pos.append(np.concatenate([_v(p) for p in tri]))
        uvs.append(np.concatenate([[u[0], u[1], 0.0, 0.0] for u in tuv]))
        mid.append(0)
    for tri in (((-3, -3, -1.5), (3, -3, -1.5), (3, 3, -1.5)), ((-3, -3, -1.5), (3, 3, -1.5), (-3, 3, -1.5))):
        pos.append(np.concatenate([_v(p) for p in tri]))
        uvs.append(np.zeros(12, np.float32))
        mid.append(1)
    scene = abi.SceneData(np.array(pos, np.float32), np.array(mid, np.uint32), [leaf, wall],
                          uvs=np.array(uvs, np.float32), images=images, textures=textures,
                          lights=[abi.light(abi.POINT, position=(0.3, 0.2, 2.5), intensity=20.0)])
    env = abi.EnvData(constant=(0.05, 0.05, 0.05))
    cam = test_camera(position=(0.4, 0.3, 3.0), forward=normalize((-0.1, -0.1, -1)), yfov=0.9,
                      up=normalize((0, 1, -0.1)))
    return scene, env, cam


def alpha_mask_stress():
    """Five alpha-masked quads over noise alpha textures (cutoff 0.5, base alpha 0.9): on a
    32x32 image, bilinear + repeat under a rotated, scaled, offset transform; nearest + clamp;
    bilinear + mirror on UV set 1; nearest + repeat with a negative scale; on a 512x512 image
    tiled 37x29 times, where the barycentric error spans a hundredth of a texel.  Bilinear
    noise puts many hits near the cutoff, so the FP32 mask decision is exercised where it
    must defer."""
    images = [_image(32, 32, 11, "noise"), _image(512, 512, 12, "noise")]
    textures = [abi.Texture(0, 0, 0, 1), abi.Texture(0, 1, 1, 0), abi.Texture(0, 2, 2, 1), abi.Texture(0, 0, 0, 0),
                abi.Texture(1, 0, 2, 1)]
    slots = [_slot(0, 0, (0.13, -0.4), (3.0, 2.5), 0.7), _slot(1, 0, (-0.2, 0.1), (1.4, 1.4), 0.0),
             _slot(2, 1, (0.05, 0.3), (2.2, -1.7), -1.1), _slot(3, 0, (0.0, 0.0), (-4.0, 3.0), 0.25),
             _slot(4, 0, (0.31, 0.77), (37.0, 29.0), 0.3)]
    mats = []
    for sl in slots:
        m = abi.material(base_color=(0.6, 0.6, 0.6, 0.9), metallic=0.0, roughness=0.7, alpha_mode=1,
                         alpha_cutoff=0.5, double_sided=True)
        m.base_color_tex = sl
        mats.append(m)
    pos, uvs, mid = [], [], []
    for q, z in enumerate((0.0, -0.7, -1.4, -2.1, -2.8)):
        a, b, c, d = (-1, -1, z), (1, -1, z), (1, 1, z), (-1, 1, z)
        for tri, tuv in (((a, b, c), ((0, 0), (1, 0), (1, 1))), ((a, c, d), ((0, 0), (1, 1), (0, 1)))):
            pos.append(np.concatenate([_v(p) for p in tri]))
            uvs.append(np.concatenate([[u[0], u[1], 0.3 + 1.7 * u[1], -0.2 + 1.3 * u[0]] for u in tuv]))
            mid.append(q)
    return abi.SceneData(np.array(pos, np.float32), np.array(mid, np.uint32), mats, uvs=np.array(uvs, np.float32),
                         images=images, textures=textures)


GLTF_FIXTURE = "gltf_export_contract.npz"
HARBOR_FIXTURE = "large/harbor.npz"


def _gltf_fixture(name):
    import pathlib
    from conftest import scenes
    scene, cam = scenes.scene_fixture(pathlib.Path(__file__).resolve().parent / "golden" / name)
    return scene, abi.EnvData(constant=(0.32, 0.36, 0.42)), cam


def gltf_export_contract():
    """The reference's glTF export-contract asset (reference scenes/fixtures/export_contract.glb:
    1602 triangles, 7 materials with base-colour, emissive, metallic-roughness and normal maps, an
    alpha-masked double-sided material, a transmissive volume, a point and a directional light),
    as the reference's io::loadGltf loads it (before Scene::finalize), with its camera.
    The arrays are a fixture written by tests/golden/make_golden.py from the reference build;
    texels are stored as the 8-bit codes k with value float32(k) / 255 (exact for this asset)."""
    return _gltf_fixture(GLTF_FIXTURE)


def gltf_harbor():
    """The reference's harbor asset (reference scenes/harbor/harbor.glb: 358,440 triangles,
    53 materials, 14 textures of up to 1024x1024, a sun and a lamp) as io::loadGltf loads it,
    with its camera.  The fixture (28 MB) is not committed: tests/golden/make_golden.py writes
    it where the reference tree exists, and the cases that use it are left out elsewhere."""
    return _gltf_fixture(HARBOR_FIXTURE)
