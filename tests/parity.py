"""Film comparison shared by the device parity tests.

A pixel agrees when every channel is within REL_TOL relative of the reference
(or 1e-7 absolute near zero); a render passes when at least PIXEL_FRACTION of
its pixels agree (all of them for films under 1000 pixels).  Pixels that do
not agree are counted as diverged paths: a last-bit difference of a libm call
that steers a sample onto another triangle changes that pixel by far more than
1e-4, while the rest stay within the tolerance.
"""
from __future__ import annotations

import numpy as np

REL_TOL = 1e-4
PIXEL_FRACTION = 0.999


def within(a, b, rel=REL_TOL):
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    ok = (d <= rel * np.abs(b.astype(np.float64))) | (d <= 1e-7)
    return np.all(ok, axis=-1)


def compare(film, ref) -> dict:
    ok = within(film, ref)
    exact = np.all(film.view(np.uint32) == ref.view(np.uint32), axis=-1)
    return {"pixels": int(ok.size), "within": float(ok.mean()), "bit_exact": float(exact.mean()),
            "diverged": int((~ok).sum())}


def assert_parity(film, ref, what: str) -> dict:
    c = compare(film, ref)
    print(f"{what}: within 1e-4 {c['within'] * 100:.3f}%  bit-exact {c['bit_exact'] * 100:.3f}%  "
          f"diverged pixels {c['diverged']}/{c['pixels']}")
    if c["pixels"] >= 1000:
        assert c["within"] >= PIXEL_FRACTION, f"{what}: only {c['within'] * 100:.3f}% of pixels within {REL_TOL}"
    else:
        assert c["diverged"] == 0, f"{what}: {c['diverged']} pixels differ"
    return c
