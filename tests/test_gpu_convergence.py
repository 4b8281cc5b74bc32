"""Converged-image quality: relMSE of GPU renders against the reference's converged film.

The reference film (tests/golden/converged_c1.npz, written by
tests/golden/make_converged.py with the reference CPU renderer) is the
10,252-triangle box-and-spheres workload at 128x128, 4096 spp, seed 7, each
pixel accumulated in binary64.  GPU renders of the same workload use seed 0,
so they share no sample stream with it: relMSE here measures convergence, not
parity.  relMSE = mean over pixels and channels of (I - R)^2 / (R^2 + 1e-2).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden, scenes

pytestmark = pytest.mark.gpu

RELMSE_TARGET = 1e-3


def relmse(img: np.ndarray, ref: np.ndarray) -> float:
    img = img.astype(np.float64)
    ref = ref.astype(np.float64)
    return float(np.mean((img - ref) ** 2 / (ref * ref + 1e-2)))


def test_relmse_definition_on_known_values():
    ref = np.full((2, 2, 3), 0.1)
    img = ref.copy()
    assert relmse(img, ref) == 0.0
    img[0, 0, 0] = 0.2  # one of 12 entries off by 0.1: 0.01 / (0.01 + 0.01) / 12
    assert abs(relmse(img, ref) - 0.5 / 12) < 1e-15


def test_gpu_render_converges_to_reference(gpu):
    g = load_golden("converged_c1.npz")
    width, height, spp_ref, seed_ref = (int(x) for x in g["meta"])
    ref = g["film"]
    w = scenes.config1(width=width, height=height, samples=16)
    sid, _ = gpu.upload_scene(w.scene)
    eid, _ = gpu.upload_env(w.env)
    errs = {}
    try:
        for spp in (64, 256, 1024, 4096):
            r = gpu.render(sid, eid, w.camera, w.options(samples=spp, seed=0))
            errs[spp] = relmse(r.pixels, ref)
    finally:
        gpu.release_scene(sid)
        gpu.release_env(eid)
    print("relMSE by spp:", {k: round(v, 6) for k, v in errs.items()})
    # Monte Carlo error falls with spp (variance ~ 1/spp plus the reference's own floor).
    assert errs[64] > errs[256] > errs[1024] > errs[4096]
    assert errs[4096] <= RELMSE_TARGET
