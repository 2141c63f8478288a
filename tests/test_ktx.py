"""KTX textures (yapt_amd/ktx.py): files laid out by the KTX 2.0 and KTX 1.1 specifications, written
here byte by byte, decode to the texels put in (channel expansion, BGR order, row padding, the sRGB
flag of the format); unsupported formats are refused by name; KTX images load through glTF
(KHR_texture_basisu sources) and through the py_yapt Resources."""
import base64
import json
import struct

import numpy as np
import pytest

from yapt_amd import gltf, ktx
from yapt_amd import py_yapt as Y


def ktx2_bytes(texels, vk_format, levels=1, scheme=0):
    """KTX 2.0: identifier, 9 header words, the index (dfd, kvd, sgd), the level index, then the
    levels, smallest first; only level 0 carries the given texels"""
    h, w, c = texels.shape
    header = ktx.KTX2_IDENTIFIER + struct.pack("<9I", vk_format, 1, w, h, 0, 0, 1, levels, scheme)
    level_index_size = 24 * levels
    data_start = 80 + level_index_size
    blobs, offsets = [], []
    cur = data_start
    sizes = [texels.nbytes] + [max(1, (w >> k) * (h >> k) * c) for k in range(1, levels)]
    for k in reversed(range(levels)):         # mip tail first
        offsets.append((k, cur))
        blobs.append(texels.tobytes() if k == 0 else b"\x7f" * sizes[k])
        cur += sizes[k]
    offs = dict(offsets)
    index = struct.pack("<IIII QQ", 0, 0, 0, 0, 0, 0)
    levels_index = b"".join(struct.pack("<3Q", offs[k], sizes[k], sizes[k]) for k in range(levels))
    return header + index + levels_index + b"".join(blobs)


def ktx1_bytes(texels, gl_format, internal):
    h, w, c = texels.shape
    row = w * c
    pitch = (row + 3) & ~3
    data = b"".join(texels[y].tobytes() + b"\0" * (pitch - row) for y in range(h))
    header = ktx.KTX1_IDENTIFIER + struct.pack("<13I", 0x04030201, ktx.GL_UNSIGNED_BYTE, 1, gl_format, internal,
                                               gl_format, w, h, 0, 0, 1, 1, 0)
    return header + struct.pack("<I", len(data)) + data


def _img(h, w, c, seed=0):
    return np.random.default_rng(seed).integers(0, 256, (h, w, c)).astype(np.uint8)


def test_ktx2_header_is_80_bytes():
    assert len(ktx2_bytes(_img(1, 1, 4), 37)) == 80 + 24 + 4


@pytest.mark.parametrize("vk_format,srgb", [(37, False), (43, True)])
def test_ktx2_rgba8_round_trip(vk_format, srgb):
    t = _img(5, 7, 4)
    out, s = ktx.decode(ktx2_bytes(t, vk_format, levels=3))
    assert s == srgb and np.array_equal(out, t)


def test_ktx2_channel_expansion_and_bgr_order():
    r = _img(3, 4, 1)
    out, _ = ktx.decode(ktx2_bytes(r, 9))
    assert np.array_equal(out[..., 0], r[..., 0]) and np.array_equal(out[..., 2], r[..., 0]) and (out[..., 3] == 255).all()
    rg = _img(3, 4, 2, 1)
    out, _ = ktx.decode(ktx2_bytes(rg, 16))
    assert np.array_equal(out[..., :2], rg) and (out[..., 2] == 0).all()
    bgra = _img(3, 4, 4, 2)
    out, s = ktx.decode(ktx2_bytes(bgra, 50))
    assert s and np.array_equal(out, bgra[..., [2, 1, 0, 3]])
    bgr = _img(2, 3, 3, 3)
    out, _ = ktx.decode(ktx2_bytes(bgr, 30))
    assert np.array_equal(out[..., :3], bgr[..., ::-1]) and (out[..., 3] == 255).all()


def test_ktx2_refuses_what_it_cannot_decode():
    with pytest.raises(ValueError, match="Zstandard"):
        ktx.decode(ktx2_bytes(_img(2, 2, 4), 37, scheme=2))
    with pytest.raises(ValueError, match="vkFormat 145"):                    # BC7_UNORM_BLOCK
        ktx.decode(ktx2_bytes(_img(2, 2, 4), 145))
    with pytest.raises(ValueError, match="shorter"):
        ktx.decode(ktx2_bytes(_img(4, 4, 4), 37)[:-8])
    with pytest.raises(ValueError, match="not a KTX"):
        ktx.decode(b"\x89PNG\r\n\x1a\n" + b"\0" * 100)


def test_ktx1_rows_are_padded_to_four_bytes():
    t = _img(3, 5, 3, 4)                                                     # 15-byte rows, padded to 16
    out, s = ktx.decode(ktx1_bytes(t, ktx.GL_RGB, ktx.GL_SRGB8))
    assert s and np.array_equal(out[..., :3], t) and (out[..., 3] == 255).all()
    t4 = _img(2, 3, 4, 5)
    out, s = ktx.decode(ktx1_bytes(t4, ktx.GL_RGBA, 0x8058))                 # GL_RGBA8: linear
    assert not s and np.array_equal(out, t4)


def test_ktx_loads_through_gltf_and_resources(tmp_path):
    t = _img(4, 4, 4, 6)
    data = ktx2_bytes(t, 43)
    P = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1]], np.float32)
    UV = np.array([[0, 0], [1, 0], [1, 1]], np.float32)
    blob = P.tobytes() + UV.tobytes() + data
    blob += b"\0" * ((4 - len(blob) % 4) % 4)
    doc = {"asset": {"version": "2.0"}, "scenes": [{"nodes": [0]}], "nodes": [{"mesh": 0}],
           "meshes": [{"primitives": [{"attributes": {"POSITION": 0, "TEXCOORD_0": 1}, "material": 0}]}],
           "materials": [{"name": "k", "pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}],
           "textures": [{"extensions": {"KHR_texture_basisu": {"source": 0}}}],
           "images": [{"bufferView": 2, "mimeType": "image/ktx2"}],
           "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 36}, {"buffer": 0, "byteOffset": 36, "byteLength": 24},
                           {"buffer": 0, "byteOffset": 60, "byteLength": len(data)}],
           "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"},
                         {"bufferView": 1, "componentType": 5126, "count": 3, "type": "VEC2"}],
           "buffers": [{"byteLength": len(blob), "uri": "data:application/octet-stream;base64," + base64.b64encode(blob).decode()}]}
    p = tmp_path / "k.gltf"
    p.write_text(json.dumps(doc))
    scene, info = gltf.load(str(p))
    assert not info["images_failed"] and scene.materials[0].albedo_tex == 0
    assert np.array_equal(scene.texel_arrays[0], t) and scene.texture_array[0].srgb == 1
    f = tmp_path / "t.ktx2"
    f.write_bytes(ktx2_bytes(t, 37))
    tex = Y.Application(8, 8).get_resources().load_texture_2D(str(f), "R8G8B8A8_SRGB")
    assert np.array_equal(tex.texels, t) and tex.srgb is False                 # the file's format wins
