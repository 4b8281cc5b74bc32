"""Film encode on the GPU against the reference's writers.

crg_film_encode produces the display bytes io::writePng stores and the whole
.hdr file io::writeHdr writes (reference src/io/image.cpp:101-124, :201-236):
both must equal the reference's output bit for bit — on the golden films
written by the reference itself (tests/golden/film_encode.npz), and on a full
1920x1080 film against the oracle's restatement (which the golden films pin).
"""
from __future__ import annotations

import struct
import zlib

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _names():
    g = load_golden("film_encode.npz")
    return sorted({k.split("/")[0] for k in g.files})


@pytest.mark.parametrize("name", _names())
def test_film_encode_matches_reference(gpu, name):
    from cr_amd.crgpu import CrgError
    g = load_golden("film_encode.npz")
    film, ex = g[f"{name}/film"], float(g[f"{name}/exposure"])
    if name.startswith("err_"):
        for kind, fn in (("png", lambda: gpu.film_png8(film, ex)), ("hdr", lambda: gpu.film_hdr(film))):
            want = str(g[f"{name}/{kind}"])
            if want:
                with pytest.raises(CrgError) as e:
                    fn()
                assert str(e.value) == want
            else:
                fn()
        return
    np.testing.assert_array_equal(gpu.film_png8(film, ex), g[f"{name}/png8"])
    assert gpu.film_hdr(film) == g[f"{name}/hdr"].tobytes()


def _big_film():
    g = np.random.default_rng(8)
    film = (g.exponential(0.6, (1080, 1920, 3)) ** 2).astype(np.float32)
    film[:, :300] = 0.0             # long runs
    film[500:520] = 1.75            # equal rows
    film[::97, ::89] = 3.0e4        # bright specks
    return film


def test_full_hd_film_matches_oracle(gpu, oracle):
    film = _big_film()
    for ex in (0.0, -1.5, 2.25):
        np.testing.assert_array_equal(gpu.film_png8(film, ex), oracle.film_png8(film, ex))
    assert gpu.film_hdr(film) == oracle.film_hdr(film)


def _png_pixels(data: bytes) -> np.ndarray:
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        assert zlib.crc32(data[pos + 4:pos + 8 + n]) == struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        if kind == b"IHDR":
            w, h, depth, color = struct.unpack(">IIBB", body[:10])
            assert (depth, color) == (8, 2)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
    assert (rows[:, 0] == 0).all()  # filter type 0
    return rows[:, 1:].reshape(h, w, 3)


def test_written_files(gpu, tmp_path):
    g = load_golden("film_encode.npz")
    film, ex = g["render_c1_64/film"], float(g["render_c1_64/exposure"])
    gpu.write_png(tmp_path / "a.png", film, ex)
    np.testing.assert_array_equal(_png_pixels((tmp_path / "a.png").read_bytes()), g["render_c1_64/png8"])
    gpu.write_hdr(tmp_path / "a.hdr", film)
    assert (tmp_path / "a.hdr").read_bytes() == g["render_c1_64/hdr"].tobytes()
