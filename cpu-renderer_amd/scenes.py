"""Seeded procedural scenes for the five benchmark/parity workloads.

Every generator uses only IEEE-exact numpy operations (+, -, *, /, sqrt on
float64, integer hashing) plus the arithmetic-only `_exp` below, so the arrays
are bit-identical on any x86-64 host; a scene built here and one built on the
GPU host hash to the same sha256.  Mapping decisions (materials for "diffuse",
"GGX", "dielectric", camera placement, volume validity) follow the rules the
reference renderer enforces: closed refractive volumes must not overlap and
the camera must be outside them (reference src/render.cpp:306-326), opaque
single-sided faces are back-face culled (reference src/geometry.cpp:279-292).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import abi

PI = 3.14159265358979323846


# ---------------------------------------------------------------- helpers
def _exp(x: np.ndarray) -> np.ndarray:
    """exp() from +,*,/ only (range reduction by ln2 + degree-20 Taylor)."""
    x = np.asarray(x, np.float64)
    ln2 = 0.6931471805599453
    k = np.floor(x / ln2 + 0.5)
    r = x - k * ln2
    term = np.ones_like(r)
    total = np.ones_like(r)
    for i in range(1, 21):
        term = term * r / i
        total = total + term
    return np.ldexp(total, k.astype(np.int64))  # power-of-two scaling is exact


def _normalize(v: np.ndarray) -> np.ndarray:
    n = np.sqrt((v * v).sum(axis=-1, keepdims=True))
    return v / n


def _unit_vectors(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform directions by rejection from the cube (no trigonometry)."""
    out = np.empty((0, 3))
    while len(out) < n:
        v = rng.random((2 * (n - len(out)) + 16, 3)) * 2.0 - 1.0
        r2 = (v * v).sum(axis=1)
        v = v[(r2 <= 1.0) & (r2 > 1e-4)]
        out = np.concatenate([out, v])
    return _normalize(out[:n])


def _quad(a, b, c, d) -> np.ndarray:
    """Two triangles (a,b,c), (a,c,d) sharing the quad's winding."""
    a, b, c, d = (np.asarray(p, np.float64) for p in (a, b, c, d))
    return np.stack([np.concatenate([a, b, c]), np.concatenate([a, c, d])])


def _flat_normals(tris: np.ndarray) -> np.ndarray:
    t = tris.reshape(-1, 3, 3)
    n = _normalize(np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]))
    return np.repeat(n[:, None, :], 3, axis=1).reshape(-1, 9)


_ICO_FACES = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                       [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                       [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])


def icosphere(level: int) -> np.ndarray:
    """Unit icosphere as (F, 3, 3) float64 vertex positions, outward winding.

    Midpoints are computed per face as normalize(a+b); IEEE addition commutes,
    so both faces sharing an edge get bit-identical midpoints (watertight).
    """
    t = (1.0 + np.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v = _normalize(v)
    f = v[_ICO_FACES]
    for _ in range(level):
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        ab, bc, ca = _normalize(a + b), _normalize(b + c), _normalize(c + a)
        f = np.stack([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                      np.stack([ab, bc, ca], 1)], 1).reshape(-1, 3, 3)
    n = np.cross(f[:, 1] - f[:, 0], f[:, 2] - f[:, 0])
    flip = (n * f.sum(axis=1)).sum(axis=1) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return f


def sphere_mesh(center, radius, level) -> Tuple[np.ndarray, np.ndarray]:
    unit = icosphere(level)
    pos = (np.asarray(center, np.float64) + radius * unit).reshape(-1, 9)
    return pos, unit.reshape(-1, 9)


def closed_box(lo, hi) -> np.ndarray:
    """12 triangles of an axis-aligned box with outward winding."""
    (x0, y0, z0), (x1, y1, z1) = lo, hi
    q = [
        _quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)),  # +z
        _quad((x1, y0, z0), (x0, y0, z0), (x0, y1, z0), (x1, y1, z0)),  # -z
        _quad((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)),  # -x
        _quad((x1, y0, z1), (x1, y0, z0), (x1, y1, z0), (x1, y1, z1)),  # +x
        _quad((x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (x0, y1, z0)),  # +y
        _quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)),  # -y
    ]
    return np.concatenate(q)


def look_at(position, target, yfov, up=(0.0, 1.0, 0.0)) -> abi.Camera:
    p = np.asarray(position, np.float64)
    f = _normalize(np.asarray(target, np.float64) - p)
    r = _normalize(np.cross(f, np.asarray(up, np.float64)))
    u = np.cross(r, f)
    return abi.camera(tuple(map(float, p.astype(np.float32))), tuple(map(float, f.astype(np.float32))),
                      tuple(map(float, r.astype(np.float32))), tuple(map(float, u.astype(np.float32))),
                      yfov=yfov)


def synthetic_sky(width: int = 1024, height: int = 512, sun_uv=(0.3, 0.22), sun_peak: float = 1e4,
                  sun_width: float = 0.006) -> np.ndarray:
    """Lat-long HDR sky: vertical gradient plus a sun lobe peaking at `sun_peak`.

    Row v=0 is +Y (the reference environment's north pole).  Arithmetic only.
    """
    v = (np.arange(height, dtype=np.float64) + 0.5) / height
    u = (np.arange(width, dtype=np.float64) + 0.5) / width
    vv, uu = np.meshgrid(v, u, indexing="ij")
    up = 1.0 - 2.0 * vv                                   # +1 zenith, -1 nadir
    s = np.clip(up, 0.0, 1.0)
    zenith = np.array([0.18, 0.36, 0.85])
    horizon = np.array([0.95, 0.92, 0.85])
    ground = np.array([0.16, 0.14, 0.12])
    sky = horizon[None, None] + (zenith - horizon)[None, None] * (s * (2.0 - s))[..., None]
    rgb = np.where((up >= 0)[..., None], sky * 1.5, ground[None, None] * (1.0 + up)[..., None])
    du = np.abs(uu - sun_uv[0])
    du = np.minimum(du, 1.0 - du) * 2.0
    dv = vv - sun_uv[1]
    r2 = (du * du + dv * dv) / (sun_width * sun_width)
    q = 1.0 + r2
    q2 = q * q
    lobe = sun_peak / (q2 * q2)
    rgb = rgb + lobe[..., None] * np.array([1.0, 0.93, 0.82])[None, None]
    out = np.ones((height, width, 4), np.float32)
    out[..., :3] = rgb
    return out


# ---------------------------------------------------------------- workloads
@dataclass
class Workload:
    name: str
    scene: abi.SceneData
    env: abi.EnvData
    camera: abi.Camera
    width: int
    height: int
    samples: int
    bounces: int
    seed: int

    def options(self, **kw) -> abi.Options:
        args = dict(width=self.width, height=self.height, samples=self.samples, bounces=self.bounces,
                    seed=self.seed)
        args.update(kw)
        return abi.options(**args)

    @property
    def camera_samples(self) -> int:
        return self.width * self.height * self.samples


def config1(seed: int = 0, width: int = 256, height: int = 256, samples: int = 16, bounces: int = 8
            ) -> Workload:
    """Closed box of 5 quads + quad light + GGX and dielectric icospheres (10,252 tris)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    albedo = 0.2 + 0.6 * rng.random(3)
    walls = np.concatenate([
        _quad((-1, -1, -1), (-1, -1, 1), (1, -1, 1), (1, -1, -1)),    # floor, normal +y
        _quad((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1)),        # ceiling, normal -y
        _quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)),    # back, normal +z
        _quad((-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (-1, -1, 1)),    # left, normal +x
        _quad((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1)),        # right, normal -x
    ])
    lamp = _quad((-0.3, 0.98, -0.3), (0.3, 0.98, -0.3), (0.3, 0.98, 0.3), (-0.3, 0.98, 0.3))
    ggx_p, ggx_n = sphere_mesh((-0.45, -0.58, -0.25), 0.4, 4)
    die_p, die_n = sphere_mesh((0.45, -0.58, 0.25), 0.4, 4)
    positions = np.concatenate([walls, lamp, ggx_p, die_p])
    normals = np.concatenate([_flat_normals(walls), _flat_normals(lamp), ggx_n, die_n])
    mids = np.concatenate([np.zeros(10), np.ones(2), np.full(len(ggx_p), 2), np.full(len(die_p), 3)])
    mats = [
        abi.material(base_color=tuple(albedo), metallic=0.0, roughness=1.0, ior=1.5),
        abi.material(base_color=(0, 0, 0), metallic=0.0, emission=(15, 15, 15)),
        abi.material(base_color=(0, 0, 0), metallic=0.0, roughness=0.2, ior=1.5),       # F0 = 0.04
        abi.material(base_color=(1, 1, 1), metallic=0.0, roughness=0.0, transmission=1.0, ior=1.5,
                     thickness=1.0),
    ]
    scene = abi.SceneData(positions.astype(np.float32), mids.astype(np.uint32), mats,
                          normals=normals.astype(np.float32))
    cam = abi.camera(position=(0, 0, 3.6), forward=(0, 0, -1), yfov=0.72)
    return Workload("config1_box_spheres", scene, abi.EnvData(constant=(0.1, 0.1, 0.1)), cam, width, height,
                    samples, bounces, seed)


def _hash2(ix: np.ndarray, iz: np.ndarray, salt: int) -> np.ndarray:
    h = (ix.astype(np.uint32) * np.uint32(0x8DA6B343)) ^ (iz.astype(np.uint32) * np.uint32(0xD8163841))
    h = h ^ np.uint32(salt * 0x2545F491 & 0xFFFFFFFF)
    h = h ^ (h >> np.uint32(15))
    h = h * np.uint32(0x2C1B3C6D)
    h = h ^ (h >> np.uint32(12))
    h = h * np.uint32(0x297A2D39)
    h = h ^ (h >> np.uint32(15))
    return (h >> np.uint32(8)).astype(np.float64) / 16777216.0


def fbm_height(x: np.ndarray, z: np.ndarray, seed: int, octaves: int = 6, amplitude: float = 6.0,
               frequency: float = 0.03) -> np.ndarray:
    h = np.zeros_like(x, np.float64)
    a, f = amplitude, frequency
    for o in range(octaves):
        px, pz = x * f, z * f
        ix, iz = np.floor(px), np.floor(pz)
        fx, fz = px - ix, pz - iz
        sx, sz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
        ix, iz = ix.astype(np.int64), iz.astype(np.int64)
        salt = seed * 31 + o
        n00, n10 = _hash2(ix, iz, salt), _hash2(ix + 1, iz, salt)
        n01, n11 = _hash2(ix, iz + 1, salt), _hash2(ix + 1, iz + 1, salt)
        n = (n00 + (n10 - n00) * sx) + ((n01 + (n11 - n01) * sx) - (n00 + (n10 - n00) * sx)) * sz
        h = h + a * (n * 2.0 - 1.0)
        a, f = a * 0.5, f * 2.0
    return h


def terrain(n: int, half_extent: float, seed: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """n x n vertex heightfield -> (positions (2(n-1)^2, 9), normals, height grid)."""
    g = np.linspace(-half_extent, half_extent, n)
    xx, zz = np.meshgrid(g, g, indexing="xy")
    hh = fbm_height(xx, zz, seed)
    step = g[1] - g[0]
    gx = np.zeros_like(hh)
    gz = np.zeros_like(hh)
    gx[:, 1:-1] = (hh[:, 2:] - hh[:, :-2]) / (2 * step)
    gx[:, 0], gx[:, -1] = gx[:, 1], gx[:, -2]
    gz[1:-1, :] = (hh[2:, :] - hh[:-2, :]) / (2 * step)
    gz[0, :], gz[-1, :] = gz[1, :], gz[-2, :]
    nrm = _normalize(np.stack([-gx, np.ones_like(hh), -gz], -1))
    p = np.stack([xx, hh, zz], -1)
    a, b, c, d = p[:-1, :-1], p[:-1, 1:], p[1:, 1:], p[1:, :-1]
    na, nb, nc, nd = nrm[:-1, :-1], nrm[:-1, 1:], nrm[1:, 1:], nrm[1:, :-1]
    # (a, d, c) and (a, c, b) wind counter-clockwise seen from +y (normal up).
    t1 = np.stack([a, d, c], -2).reshape(-1, 9)
    t2 = np.stack([a, c, b], -2).reshape(-1, 9)
    n1 = np.stack([na, nd, nc], -2).reshape(-1, 9)
    n2 = np.stack([na, nc, nb], -2).reshape(-1, 9)
    pos = np.stack([t1, t2], 1).reshape(-1, 9)
    nor = np.stack([n1, n2], 1).reshape(-1, 9)
    return pos, nor, hh


def _place_spheres(rng, count, region, rmin, rmax, height_fn, level):
    centers, radii = [], []
    tries = 0
    while len(centers) < count and tries < 200000:
        tries += 1
        r = rmin + (rmax - rmin) * rng.random()
        x = region[0] + (region[1] - region[0]) * rng.random()
        z = region[2] + (region[3] - region[2]) * rng.random()
        ok = all((x - cx) ** 2 + (z - cz) ** 2 > (r + cr + 0.1) ** 2 for (cx, _, cz), cr in zip(centers, radii))
        if not ok:
            continue
        y = height_fn(x, z, r) + r + 0.05
        centers.append((x, y, z))
        radii.append(r)
    return centers, radii


def config2(seed: int = 0, grid: int = 1024, spheres: int = 200, sphere_level: int = 4, width: int = 1920,
            height: int = 1080, samples: int = 64, bounces: int = 16) -> Workload:
    """fBm terrain (2(grid-1)^2 tris) + `spheres` GGX/dielectric icospheres, synthetic HDR sky."""
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    half = 64.0
    tpos, tnor, hh = terrain(grid, half, seed)
    step = 2 * half / (grid - 1)

    def footprint_max(x, z, r):
        i0 = int(np.clip(np.floor((x - r + half) / step), 0, grid - 1))
        i1 = int(np.clip(np.ceil((x + r + half) / step), 0, grid - 1))
        j0 = int(np.clip(np.floor((z - r + half) / step), 0, grid - 1))
        j1 = int(np.clip(np.ceil((z + r + half) / step), 0, grid - 1))
        return float(hh[j0:j1 + 1, i0:i1 + 1].max())

    centers, radii = _place_spheres(rng, spheres, (-30, 30, -30, 30), 0.6, 2.0, footprint_max, sphere_level)
    mats = [abi.material(base_color=(0.42, 0.36, 0.27), metallic=0.0, roughness=1.0, ior=1.5)]
    pos_list, nor_list, mid_list = [tpos], [tnor], [np.zeros(len(tpos))]
    for c, r in zip(centers, radii):
        rough = 0.05 + 0.75 * rng.random()
        if rng.random() < 0.5:
            col = 0.5 + 0.45 * rng.random(3)
            mats.append(abi.material(base_color=tuple(col), metallic=1.0, roughness=rough))
        else:
            att = 0.6 + 0.4 * rng.random(3)
            mats.append(abi.material(base_color=(1, 1, 1), metallic=0.0, roughness=rough, transmission=1.0,
                                     ior=1.5, thickness=1.0, attenuation_distance=3.0,
                                     attenuation_color=tuple(att)))
        p, n = sphere_mesh(c, r, sphere_level)
        pos_list.append(p)
        nor_list.append(n)
        mid_list.append(np.full(len(p), len(mats) - 1))
    scene = abi.SceneData(np.concatenate(pos_list).astype(np.float32),
                          np.concatenate(mid_list).astype(np.uint32), mats,
                          normals=np.concatenate(nor_list).astype(np.float32))
    cam_y = float(fbm_height(np.array([0.0]), np.array([44.0]), seed)[0]) + 14.0
    cam = look_at((0.0, cam_y, 44.0), (0.0, 0.0, 0.0), 0.8)
    env = abi.EnvData(rgba=synthetic_sky())
    return Workload("config2_terrain_spheres", scene, env, cam, width, height, samples, bounces, seed)


def soup(rng: np.random.Generator, n: int, lo: float = 1e-4, hi: float = 1e-2) -> np.ndarray:
    centers = rng.random((n, 3))
    scale = _exp(math.log(lo) + (math.log(hi) - math.log(lo)) * rng.random(n))
    pos = np.empty((n, 9))
    for k in range(3):
        pos[:, 3 * k:3 * k + 3] = centers + scale[:, None] * _unit_vectors(rng, n) * 0.5773502691896258
    return pos


def config3(seed: int = 0, triangles: int = 8_000_000, width: int = 1280, height: int = 720, samples: int = 32,
            bounces: int = 24) -> Workload:
    """Random triangle soup (50% dielectric sheets / 30% GGX / 20% diffuse) + 4 closed dielectric slabs."""
    rng = np.random.Generator(np.random.PCG64(seed + 2000))
    pos = soup(rng, triangles)
    pick = rng.random(triangles)
    mids = np.where(pick < 0.5, 0, np.where(pick < 0.8, 1, 2))
    mats = [
        abi.material(base_color=(1, 1, 1), metallic=0.0, roughness=0.1, transmission=1.0, ior=1.5,
                     thickness=0.0, double_sided=True),
        abi.material(base_color=(0.9, 0.62, 0.3), metallic=1.0, roughness=0.3, double_sided=True),
        abi.material(base_color=(0.6, 0.6, 0.6), metallic=0.0, roughness=1.0, ior=1.5, double_sided=True),
        abi.material(base_color=(1, 1, 1), metallic=0.0, roughness=0.05, transmission=1.0, ior=1.5,
                     thickness=1.0, attenuation_distance=0.5, attenuation_color=(0.8, 0.9, 1.0)),
    ]
    slabs = [closed_box((x0, y0, 1.2), (x0 + 0.45, y0 + 0.45, 1.23))
             for x0 in (0.02, 0.53) for y0 in (0.02, 0.53)]
    slab_pos = np.concatenate(slabs)
    positions = np.concatenate([pos, slab_pos]).astype(np.float32)
    mids = np.concatenate([mids, np.full(len(slab_pos), 3)]).astype(np.uint32)
    scene = abi.SceneData(positions, mids, mats)
    cam = abi.camera(position=(0.5, 0.5, 3.2), forward=(0, 0, -1), yfov=0.55)
    env = abi.EnvData(rgba=synthetic_sky())
    return Workload("config3_soup_slabs", scene, env, cam, width, height, samples, bounces, seed)


def _loguniform_int(rng, lo, hi) -> int:
    return int(np.floor(_exp(math.log(lo) + (math.log(hi) - math.log(lo)) * rng.random()) + 0.5))


def batch_request(seed: int, sky: Optional[np.ndarray] = None, max_triangles: int = 2_000_000) -> Workload:
    """One request of the batched/served mix: terrain, spheres or soup of 1k..2M tris,
    128^2..1024^2 px, 4..256 spp, 8..16 bounces (all seeded)."""
    rng = np.random.Generator(np.random.PCG64(seed + 5000))
    kind = int(rng.integers(0, 3))
    tris = min(_loguniform_int(rng, 1000, 2_000_000), max_triangles)
    side = int(rng.integers(128, 1025))
    spp = max(4, min(256, _loguniform_int(rng, 4, 256)))
    bounces = int(rng.integers(8, 17))
    if kind == 0:
        grid = max(4, int(np.sqrt(tris / 2)) + 1)
        pos, nor, hh = terrain(grid, 32.0, seed)
        mats = [abi.material(base_color=(0.4, 0.38, 0.3), metallic=0.0, roughness=1.0)]
        scene = abi.SceneData(pos.astype(np.float32), np.zeros(len(pos), np.uint32), mats,
                              normals=nor.astype(np.float32))
        cam = look_at((0.0, float(hh.max()) + 10.0, 40.0), (0.0, 0.0, 0.0), 0.9)
    elif kind == 1:
        count = max(1, tris // 1280)
        cols = int(np.ceil(np.sqrt(count)))
        mats, pos_list, nor_list, mid_list = [], [], [], []
        floor = _quad((-40, 0, -40), (-40, 0, 40), (40, 0, 40), (40, 0, -40))
        mats.append(abi.material(base_color=(0.5, 0.5, 0.5), metallic=0.0, roughness=1.0))
        pos_list.append(floor)
        nor_list.append(_flat_normals(floor))
        mid_list.append(np.zeros(2))
        for i in range(count):
            cx = (i % cols - cols / 2) * 2.2
            cz = (i // cols - cols / 2) * 2.2
            rough = 0.05 + 0.75 * rng.random()
            if rng.random() < 0.5:
                mats.append(abi.material(base_color=tuple(0.5 + 0.45 * rng.random(3)), metallic=1.0,
                                         roughness=rough))
            else:
                mats.append(abi.material(base_color=(1, 1, 1), metallic=0.0, roughness=rough, transmission=1.0,
                                         ior=1.5, thickness=1.0))
            p, n = sphere_mesh((cx, 1.05, cz), 1.0, 3)
            pos_list.append(p)
            nor_list.append(n)
            mid_list.append(np.full(len(p), len(mats) - 1))
        scene = abi.SceneData(np.concatenate(pos_list).astype(np.float32),
                              np.concatenate(mid_list).astype(np.uint32), mats,
                              normals=np.concatenate(nor_list).astype(np.float32))
        ext = cols * 1.1 + 2
        cam = look_at((0.0, ext * 0.8 + 2.0, ext * 1.6 + 4.0), (0.0, 0.0, 0.0), 0.8)
    else:
        p = soup(rng, tris)
        pick = rng.random(tris)
        mids = np.where(pick < 0.5, 0, np.where(pick < 0.8, 1, 2)).astype(np.uint32)
        mats = [
            abi.material(base_color=(1, 1, 1), metallic=0.0, roughness=0.1, transmission=1.0, ior=1.5,
                         double_sided=True),
            abi.material(base_color=(0.9, 0.62, 0.3), metallic=1.0, roughness=0.3, double_sided=True),
            abi.material(base_color=(0.6, 0.6, 0.6), metallic=0.0, roughness=1.0, double_sided=True),
        ]
        scene = abi.SceneData(p.astype(np.float32), mids, mats)
        cam = abi.camera(position=(0.5, 0.5, 3.2), forward=(0, 0, -1), yfov=0.55)
    env = abi.EnvData(rgba=sky if sky is not None else synthetic_sky())
    return Workload(f"batch_{seed}_{['terrain', 'spheres', 'soup'][kind]}", scene, env, cam, side, side, spp,
                    bounces, seed)


def served_stream(count: int = 200, rate: float = 20.0, catalog: int = 64, seed: int = 0):
    """Poisson arrivals (rate req/s) and per-request (catalog scene index, side, spp, bounces)
    of the served mix; sizes follow batch_request's ranges (128^2..1024^2 px, 4..256 spp,
    8..16 bounces)."""
    rng = np.random.Generator(np.random.PCG64(seed + 9000))
    gaps = rng.exponential(1.0 / rate, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = []
    for _ in range(count):
        k = int(rng.integers(0, catalog))
        side = int(rng.integers(128, 1025))
        spp = max(4, min(256, _loguniform_int(rng, 4, 256)))
        bounces = int(rng.integers(8, 17))
        reqs.append((k, side, spp, bounces))
    return arrivals, reqs


def random_rays(scene: abi.SceneData, count: int, seed: int, spread: float = 1.5) -> np.ndarray:
    """Ray set in the style of the reference BVH test (tests/test_geometry.cpp:533-568):
    random origins around the scene bounds, every other ray aimed at a point of a
    triangle, some with non-unit directions, tmin 0.5 and finite tmax mixes."""
    rng = np.random.Generator(np.random.PCG64(seed + 9000))
    p = scene.positions.reshape(-1, 3, 3).astype(np.float64)
    lo, hi = p.reshape(-1, 3).min(0), p.reshape(-1, 3).max(0)
    mid, ext = (lo + hi) / 2, (hi - lo) / 2 * spread + 1e-3
    origin = mid + ext * (rng.random((count, 3)) * 2 - 1)
    direction = _unit_vectors(rng, count)
    idx = rng.integers(0, len(p), count)
    target = p[idx, 0] * 0.5 + p[idx, 1] * 0.2 + p[idx, 2] * 0.3
    aimed = np.arange(count) % 2 == 0
    direction[aimed] = _normalize(target[aimed] - origin[aimed])
    i = np.arange(count)
    direction[i % 5 == 0] *= 3.0
    rays = abi.make_rays(origin.astype(np.float32), direction.astype(np.float32))
    rays["tmin"] = np.where(i % 7 == 0, 0.5, 0.0)
    tmax = np.where(i % 3 == 0, 1.0 + 29.0 * rng.random(count), np.inf)
    rays["tmax"] = tmax.astype(np.float32)
    return rays


def scene_fixture(path) -> Tuple[abi.SceneData, abi.Camera]:
    """A scene stored as a fixture file: the arrays of a crg_scene_desc (materials, lights and
    textures as their C records' bytes, texels as 8-bit codes k with value float32(k) / 255) and
    one camera, as written from a glTF asset loaded by the reference's own loader."""
    z = np.load(path)
    rec = lambda cls, key: [cls.from_buffer_copy(z[key][i].tobytes()) for i in range(len(z[key]))]
    images = [z[f"image{i}"].astype(np.float32) / np.float32(255) for i in range(int(z["image_count"]))]
    scene = abi.SceneData(z["positions"], z["material_ids"], rec(abi.Material, "materials"), normals=z["normals"],
                          tangents=z["tangents"], uvs=z["uvs"], lights=rec(abi.Light, "lights"), images=images,
                          textures=rec(abi.Texture, "textures"))
    return scene, abi.Camera.from_buffer_copy(z["camera"].tobytes())
