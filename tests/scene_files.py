"""Scenes as flat binary files for the C++ test programs (adapter_check,
sanitize_check): counts, positions, normals, material ids, crg_material and
crg_light records, a crg_camera, a constant environment radiance and crg_options."""
from __future__ import annotations

import pathlib

import numpy as np


def write_scene_file(path: pathlib.Path, scene, env, cam, opt) -> None:
    n = scene.triangle_count
    normals = scene.normals if scene.normals is not None else np.zeros((n, 9), np.float32)
    with open(path, "wb") as f:
        f.write(np.array([n, len(scene.materials), len(scene.lights)], np.uint64).tobytes())
        f.write(np.ascontiguousarray(scene.positions, np.float32).tobytes())
        f.write(np.ascontiguousarray(normals, np.float32).tobytes())
        f.write(np.ascontiguousarray(scene.material_ids, np.uint32).tobytes())
        for m in scene.materials:
            f.write(bytes(m))
        for l in scene.lights:
            f.write(bytes(l))
        f.write(bytes(cam))
        f.write(np.array(env.constant, np.float32).tobytes())
        f.write(bytes(opt))
