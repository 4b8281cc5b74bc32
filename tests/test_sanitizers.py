"""Host code under sanitizers (CPU only).

tests/csrc/sanitize_check.cpp drives the library's host preparation
(cpu-renderer_amd/csrc/host_scene.cpp: finalize, the parallel SAH build, the
quantized trees, light and environment tables) and the device logic compiled
for the CPU (tests/csrc/emu.cpp: traversal over both tree widths, exact
resolution, shading, a render) over test scenes.  It is built twice with
clang: AddressSanitizer + UndefinedBehaviorSanitizer (any report aborts the
run), and ThreadSanitizer over a scene large enough for the threaded build.
"""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

import ref_scenes as rs
from conftest import ROOT, abi, scenes
from scene_files import write_scene_file

SRC = [ROOT / "tests" / "csrc" / "sanitize_check.cpp", ROOT / "tests" / "csrc" / "emu.cpp",
       ROOT / "cpu-renderer_amd" / "csrc" / "host_scene.cpp", ROOT / "cpu-renderer_amd" / "csrc" / "film_host.cpp"]
FLAGS = ["-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-fno-fast-math", "-pthread"]

pytestmark = pytest.mark.skipif(shutil.which("clang++") is None, reason="clang++ not available")


def _build(tmp_path, sanitize: str):
    exe = tmp_path / f"sanitize_{sanitize.replace(',', '_')}"
    subprocess.run(["clang++", *FLAGS, f"-fsanitize={sanitize}", "-fno-sanitize-recover=all", "-o", str(exe),
                    *map(str, SRC), "-lz"], check=True, capture_output=True, text=True)
    return exe


def _scenes(tmp_path, big: bool):
    out = []
    cases = [("fixture4", rs.integrator_fixture(4)), ("fixture2", rs.integrator_fixture(2)),
             ("internal_point", rs.transport_internal_point_light(False)), ("bvh_random", None)]
    for name, case in cases:
        if case is None:
            scene = rs.bvh_random()
            env, cam = abi.EnvData(constant=(0.5, 0.5, 0.5)), rs.test_camera()
        else:
            scene, env, cam = case
        if env.rgba is not None:
            env = abi.EnvData(constant=(0.3, 0.3, 0.3))
        p = tmp_path / f"{name}.bin"
        write_scene_file(p, scene, env, cam, abi.options(12, 10, 2, 6, seed=3))
        out.append(p)
    if big:  # enough triangles for the parallel build paths (chunked partitions, subtree tasks)
        w = scenes.config3(triangles=300_000)
        p = tmp_path / "soup.bin"
        write_scene_file(p, w.scene, abi.EnvData(constant=(0.4, 0.4, 0.4)), w.camera, abi.options(8, 8, 1, 8, seed=1))
        out.append(p)
    return out


def _run(exe, files, env):
    p = subprocess.run([str(exe), *map(str, files)], capture_output=True, text=True, timeout=600,
                       env={**os.environ, **env})
    print(p.stdout[-3000:], p.stderr[-3000:])
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in p.stderr and "runtime error:" not in p.stderr
    assert "WARNING: ThreadSanitizer" not in p.stderr


def test_host_code_under_address_and_undefined_behavior_sanitizers(tmp_path):
    exe = _build(tmp_path, "address,undefined")
    _run(exe, _scenes(tmp_path, big=True),
         {"ASAN_OPTIONS": "detect_leaks=1:halt_on_error=1", "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})


def test_threaded_host_build_under_thread_sanitizer(tmp_path):
    exe = _build(tmp_path, "thread")
    _run(exe, _scenes(tmp_path, big=True), {"TSAN_OPTIONS": "halt_on_error=1"})
