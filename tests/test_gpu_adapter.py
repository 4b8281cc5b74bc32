"""The reference-side drop-in on the device.

oracle/_ref/adapter_check is tests/csrc/adapter_check.cpp compiled with the
reference's own sources (its cr::Scene, cr::Bvh, cr::Camera, cr::Environment,
cr::RenderOptions, cr::Progress) and include/crgpu_render.hpp, and linked
against libcrgpu.so; it is built where the reference tree exists
(`make -C oracle ref`, run by __graft_entry__.build()).  It rebuilds a
cr::Scene from a scene file and calls cr::gpu::render with exactly the
parameter list of cr::render (reference src/render.h:53).

* The integrator fixtures render to the reference's films (the render
  goldens' per-pixel rule, equal replay counts).
* Progress follows the reference's contract (reference src/render.cpp:419-427,
  :454): calls at least 2 s apart with finished <= total samples, then one
  final (total, total, seconds) call.
* An image Environment and options outside the reference's ranges are refused
  with the drop-in's and the reference's messages.
"""
from __future__ import annotations

import pathlib
import subprocess

import numpy as np
import pytest

import ref_scenes as rs
from conftest import ROOT, abi, load_golden
from parity import assert_parity
from scene_files import write_scene_file

pytestmark = pytest.mark.gpu

EXE = ROOT / "oracle" / "_ref" / "adapter_check"


def _run(tmp_path, scene, env, cam, opt):
    if not EXE.exists():
        pytest.fail(f"{EXE} is not built (make -C oracle ref where the reference tree exists)")
    src, film = tmp_path / "scene.bin", tmp_path / "film.bin"
    write_scene_file(src, scene, env, cam, opt)
    p = subprocess.run([str(EXE), str(src), str(film)], capture_output=True, text=True, timeout=240)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    pixels = np.fromfile(film, np.float32).reshape(opt.height, opt.width, 3)
    return pixels, p.stdout.splitlines()


def _progress(lines):
    return [(int(a), int(b), float(c)) for a, b, c in (l.split()[1:] for l in lines if l.startswith("progress "))]


@pytest.mark.parametrize("kind", [0, 2, 4])
def test_drop_in_renders_the_reference_film(tmp_path, kind):
    name = f"fixture{kind}"
    g = load_golden("render.npz")
    scene, env, cam = rs.integrator_fixture(kind)
    opt = abi.options(32, 32, 32, 8, seed=42)
    pixels, lines = _run(tmp_path, scene, env, cam, opt)
    assert_parity(pixels, g[f"{name}/film"], f"{name} through cr::gpu::render")
    head = next(l for l in lines if l.startswith("rendered"))
    assert f"{32 * 32 * 32} samples, {int(g[f'{name}/retries'])} replays, workers 0" in head
    calls = _progress(lines)
    assert calls[-1][:2] == (32 * 32 * 32, 32 * 32 * 32)
    assert any(l.startswith("image environment: cr::gpu::render: an image Environment keeps its texels private")
               for l in lines)
    assert "bad options: Samples must be in [1, 1048576]" in lines


def test_drop_in_reports_progress_every_two_seconds(tmp_path):
    """A render long enough for intermediate reports (1024x1024, 2048 spp of the
    lit-room fixture: 2.1 G camera samples)."""
    scene, env, cam = rs.integrator_fixture(0)
    opt = abi.options(1024, 1024, 2048, 8, seed=1)
    _, lines = _run(tmp_path, scene, env, cam, opt)
    total = 1024 * 1024 * 2048
    calls = _progress(lines)
    head = next(l for l in lines if l.startswith("rendered"))
    seconds = float(head.split(",")[-1].split()[0])
    assert calls[-1][:2] == (total, total)
    inter = calls[:-1]
    if seconds >= 2.5:
        assert inter, f"no intermediate progress in a {seconds:.2f} s render"
    for a, b in zip(inter, inter[1:]):
        assert b[2] - a[2] >= 1.99 and b[0] >= a[0]
    for done, tot, sec in inter:
        assert tot == total and 0 <= done <= total and sec >= 1.99
