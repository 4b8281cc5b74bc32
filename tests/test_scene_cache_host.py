"""Scene cache files checked on the CPU (no GPU needed).

crg_scene_write_cache prepares a scene on the host (finalize, SAH build,
flattening) and writes the cache file crg_scene_load reads; crg_scene_check
makes the loader's checks without a device.  A file whose checksum matches but
whose contents no longer describe a scene (an index edited, the checksum then
recomputed, as a hand-edited file or a writer with another meaning of a field
would produce) must be refused with a message naming what is wrong, before
anything reaches the device: every index a kernel or the host follows is
checked (node children and leaf ranges, triangle ids, materials, flags, the
leaf-order maps, texture slots, textures, images and texel ranges, emissive
triangles, light types).
"""
from __future__ import annotations

import struct

import numpy as np
import pytest

import ref_scenes
from conftest import abi

MAGIC_LEN, RECORD_SIZES = 8, 9
CHILD4 = 40  # byte offset of a 4-wide node's child refs (origin, exponents, 12 + 12 quantized planes)


def _fnv_step(h, w):
    h = ((h ^ w) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return ((h << 23) | (h >> 41)) & 0xFFFFFFFFFFFFFFFF


def _block_hash(b: bytes) -> int:
    basis = 0xCBF29CE484222325
    h = [basis, basis + 1, basis + 2, basis + 3]
    full = len(b) // 32 * 32
    words = np.frombuffer(b[:full], dtype="<u8").reshape(-1, 4) if full else np.zeros((0, 4), np.uint64)
    for row in words.tolist():
        for k in range(4):
            h[k] = _fnv_step(h[k], row[k])
    for byte in b[full:]:
        h[0] = _fnv_step(h[0], byte)
    r = basis
    for k in range(4):
        r = _fnv_step(r, h[k])
    return r


def cache_checksum(body: bytes) -> int:
    """The cache file checksum: FNV-style lanes over 1 MiB blocks, then over the
    byte count and the block hashes (the writer's documented construction)."""
    r = _fnv_step(0xCBF29CE484222325, len(body))
    for i in range(0, len(body), 1 << 20):
        r = _fnv_step(r, _block_hash(body[i:i + (1 << 20)]))
    return r


class CacheFile:
    """Section offsets of a cache file, parsed from its own header."""

    NAMES = ("nodes", "nodes8", "prims", "prim_ng", "aux", "materials", "uvs", "tangents", "texels", "images",
             "textures")

    def __init__(self, data: bytes):
        self.data = bytearray(data)
        sizes = struct.unpack_from("<9I", data, MAGIC_LEN)
        rec = {"nodes": sizes[0], "nodes8": sizes[1], "prims": sizes[2], "aux": sizes[3], "materials": sizes[4],
               "images": sizes[6], "textures": sizes[7], "prim_ng": sizes[8], "uvs": sizes[8], "tangents": sizes[8],
               "texels": sizes[8]}
        self.rec = rec
        self.punctual_size = sizes[5]
        pos = MAGIC_LEN + 4 * RECORD_SIZES
        self.triangles, self.binary_nodes = struct.unpack_from("<QQ", data, pos)
        pos += 16
        self.sec = {}
        for name in self.NAMES:
            (count,) = struct.unpack_from("<Q", data, pos)
            self.sec[name] = (pos + 8, count)
            pos += 8 + count * rec[name]
        for name, size in (("order", 4), ("tri_to_prim", 4)):
            (count,) = struct.unpack_from("<Q", data, pos)
            self.sec[name] = (pos + 8, count)
            pos += 8 + count * size
        pos += 8  # radius2
        for name, size in (("punctual_weight", 8), ("emissive", 4), ("emissive_weight", 8),
                           ("punctual", self.punctual_size)):
            (count,) = struct.unpack_from("<Q", data, pos)
            self.sec[name] = (pos + 8, count)
            pos += 8 + count * size
        assert pos == len(data) - 8

    def offset(self, name, index, field):
        start, count = self.sec[name]
        assert index < count
        size = self.rec.get(name, {"order": 4, "tri_to_prim": 4, "emissive": 4, "punctual": self.punctual_size}.get(name))
        return start + index * size + field

    def put_u32(self, name, index, field, value):
        struct.pack_into("<I", self.data, self.offset(name, index, field), value)

    def put_u64(self, name, index, field, value):
        struct.pack_into("<Q", self.data, self.offset(name, index, field), value)

    def get_u32(self, name, index, field):
        return struct.unpack_from("<I", self.data, self.offset(name, index, field))[0]

    def sealed(self) -> bytes:
        body = bytes(self.data[:-8])
        return body + struct.pack("<Q", cache_checksum(body))


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    from cr_amd import crgpu
    scene, _, _ = ref_scenes.textured_factors()
    path = tmp_path_factory.mktemp("cache") / "textured.crgscene"
    info = crgpu.write_scene_cache(scene, path)
    assert info.triangles == 6 and info.wide_nodes >= 1
    return path


def test_written_cache_passes_the_loader_checks(written):
    from cr_amd import crgpu
    crgpu.check_scene_cache(written)
    f = CacheFile(written.read_bytes())
    assert f.sec["images"][1] == 3 and f.sec["textures"][1] == 4 and f.sec["emissive"][1] == 2
    assert f.sealed() == written.read_bytes()  # the checksum restatement matches the writer


def _tampered(written, tmp_path, edit):
    f = CacheFile(written.read_bytes())
    edit(f)
    out = tmp_path / "tampered.crgscene"
    out.write_bytes(f.sealed())
    return out


def _inner_or_leaf(f, want_leaf):
    for i in range(f.sec["nodes"][1]):
        for c in range(4):
            r = f.get_u32("nodes", i, CHILD4 + 4 * c)
            if r != 0x7FFFFFFF and bool(r & 0x80000000) == want_leaf:
                return i, c, r
    return None


def _stretch_leaf(f):
    """A leaf ref whose 16-primitive range starts at the last primitive."""
    node, child, _ = _inner_or_leaf(f, True)
    f.put_u32("nodes", node, CHILD4 + 4 * child, 0x80000000 | (15 << 27) | (f.sec["prims"][1] - 1))


EDITS = {
    "leaf range past the primitives": (lambda f: _stretch_leaf(f), "BVH node refers outside"),
    "node child back to the root": (
        lambda f: f.put_u32("nodes", 0, CHILD4, 0), "BVH node refers outside"),
    "8-wide node child past the nodes": (
        lambda f: f.put_u32("nodes8", 0, 64, f.sec["nodes8"][1] + 3), "BVH node refers outside"),
    "scale exponent out of range": (
        lambda f: f.put_u32("nodes", 0, 12, f.get_u32("nodes", 0, 12) & 0xFFFFFF00), "BVH node refers outside"),
    "triangle id past the triangles": (
        lambda f: f.put_u32("prims", 1, 12, f.triangles), "primitive record refers outside"),
    "material past the materials": (
        lambda f: f.put_u32("prims", 2, 44, 99), "primitive record refers outside"),
    "unknown primitive flag": (
        lambda f: f.put_u32("prims", 0, 28, f.get_u32("prims", 0, 28) | 0x40), "primitive record refers outside"),
    "leaf-order map": (
        lambda f: f.put_u32("order", 0, 0, (f.get_u32("order", 0, 0) + 1) % f.triangles), "refers outside"),
    "triangle to primitive map": (
        lambda f: f.put_u32("tri_to_prim", 0, 0, f.sec["prims"][1] + 5), "triangle to primitive map"),
    "image past the texels": (
        lambda f: f.put_u64("images", 2, 0, f.sec["texels"][1] - 3), "image 2 lies outside the texel array"),
    "texture past the images": (
        lambda f: f.put_u32("textures", 1, 0, 7), "texture 1 refers outside the image array"),
    "emissive triangle past the triangles": (
        lambda f: f.put_u32("emissive", 0, 0, 1000), "emissive triangle outside"),
}


@pytest.mark.parametrize("case", sorted(EDITS))
def test_tampered_cache_is_refused_with_a_clear_message(written, tmp_path, case):
    from cr_amd import crgpu
    edit, message = EDITS[case]
    bad = _tampered(written, tmp_path, edit)
    assert bad.read_bytes() != written.read_bytes()
    with pytest.raises(crgpu.CrgError) as e:
        crgpu.check_scene_cache(bad)
    assert e.value.code == abi.CRG_EINVAL
    assert message in str(e.value), str(e.value)
    assert str(bad) in str(e.value)


def test_material_slot_past_the_textures_is_refused(written, tmp_path):
    """Material 0's base-colour slot (DevSlot.texture, at byte 96 of the device
    material: 23 four-byte factors, then the 8-byte aligned slots) edited to a
    texture index that does not exist."""
    from cr_amd import crgpu
    f = CacheFile(written.read_bytes())
    start, _ = f.sec["materials"]
    assert struct.unpack_from("<i", f.data, start + 96)[0] == 0  # material 0 binds texture 0
    struct.pack_into("<i", f.data, start + 96, 9)
    bad = tmp_path / "slot.crgscene"
    bad.write_bytes(f.sealed())
    with pytest.raises(crgpu.CrgError) as e:
        crgpu.check_scene_cache(bad)
    assert "material 0 refers outside the texture array" in str(e.value)


def test_damaged_and_foreign_files_are_refused(written, tmp_path):
    from cr_amd import crgpu
    data = bytearray(written.read_bytes())
    data[200] ^= 1
    (tmp_path / "flip").write_bytes(bytes(data))
    with pytest.raises(crgpu.CrgError, match="checksum mismatch"):
        crgpu.check_scene_cache(tmp_path / "flip")
    old = bytearray(written.read_bytes())
    old[6:8] = b"1\0"  # the previous format version's magic
    (tmp_path / "old").write_bytes(bytes(old))
    with pytest.raises(crgpu.CrgError, match="another format version"):
        crgpu.check_scene_cache(tmp_path / "old")
    (tmp_path / "foreign").write_bytes(b"not a cache file at all" * 10)
    with pytest.raises(crgpu.CrgError, match="not a scene cache"):
        crgpu.check_scene_cache(tmp_path / "foreign")
    with pytest.raises(crgpu.CrgError, match="cannot open"):
        crgpu.check_scene_cache(tmp_path / "missing")
